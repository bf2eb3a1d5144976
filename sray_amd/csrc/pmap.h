/*
 * pmap.h — photon maps on gfx950: a balanced kd-tree built on the device and
 * the gathers of PhotonMap::irradiance_est / radiance_est
 * (src/pmap.cc:107-178), which in the reference call the external kdtree
 * library's kd_nearest_range3f (positions and radius passed as float).
 *
 * Tree layout: photons are permuted so that the node of a segment [lo,hi)
 * is the photon at m = (lo+hi)/2, with its left subtree in [lo,m) and its
 * right subtree in [m+1,hi); axis[m] is its split axis (largest extent of
 * the segment).  Built level by level: every open segment is sorted along
 * its axis with one device radix sort of (segment, coordinate) keys.
 *
 * Inclusion test: dist^2 <= r^2 in double on the float positions, r float
 * (kd_nearest_range3f's float API); sums in double.  Fixed-radius range
 * (the reference) or the k nearest within the radius, ties by photon index.
 */
#ifndef SRAY_PMAP_H_
#define SRAY_PMAP_H_

#include "device.h"

namespace srd {

struct PmapTree {
    const float4 *pos;   /* xyz, w = original photon index (bits) */
    const float4 *dir;   /* xyz, w = colour r */
    const float4 *nrm;   /* xyz, w = colour g */
    const float *colb;   /* colour b */
    const uint8_t *axis; /* split axis per node */
    uint32_t n;
    uint32_t depth;      /* levels (stack bound) */
};

constexpr int PMAP_STACK = 32;
constexpr double MAX_COS_ANGLE = 0.30901699437494745; /* cos(pi/2.5), src/pmap.cc:147 */

/* accumulates the reference's estimate over all photons within r of q;
 * mode 0 irradiance (photons whose dir faces n), 1 radiance (-dir.n > 0 and
 * norm.n > cos(pi/2.5), weighted by -dir.n) */
struct GatherAcc {
    double r, g, b;
    uint32_t count;
};

__device__ __forceinline__ void gather_photon(const PmapTree &T, uint32_t m, int mode, double nx, double ny,
                                              double nz, GatherAcc &a)
{
    float4 d = T.dir[m], nn = T.nrm[m];
    float cb = T.colb[m];
    a.count++;
    if (mode == SRAY_GATHER_IRRADIANCE) {
        double dn = (double)d.x * nx + (double)d.y * ny + (double)d.z * nz;
        if (dn < 0.0) {
            a.r += (double)d.w;
            a.g += (double)nn.w;
            a.b += (double)cb;
        }
    } else {
        double ndotl = -(double)d.x * nx + -(double)d.y * ny + -(double)d.z * nz;
        double pn = (double)nn.x * nx + (double)nn.y * ny + (double)nn.z * nz;
        if (ndotl > 0.0 && pn > MAX_COS_ANGLE) {
            a.r += (double)d.w * ndotl;
            a.g += (double)nn.w * ndotl;
            a.b += (double)cb * ndotl;
        }
    }
}

/* range query; stk: this lane's LDS stack (stride apart) */
__device__ __forceinline__ void gather_range(const PmapTree &T, float qx, float qy, float qz, float radius, int mode,
                                             double nx, double ny, double nz, GatherAcc &a, uint2 *stk, int stride)
{
    a.r = a.g = a.b = 0.0;
    a.count = 0;
    if (T.n == 0) return;
    double r2 = (double)radius * (double)radius;
    double q[3] = {(double)qx, (double)qy, (double)qz};
    int sp = 0;
    stk[0] = make_uint2(0u, T.n);
    sp = 1;
    while (sp > 0) {
        uint2 s = stk[(--sp) * stride];
        uint32_t lo = s.x, hi = s.y;
        while (lo < hi) {
            uint32_t m = (lo + hi) >> 1;
            float4 p = T.pos[m];
            double dx = (double)p.x - q[0], dy = (double)p.y - q[1], dz = (double)p.z - q[2];
            double d2 = dx * dx + dy * dy + dz * dz;
            if (d2 <= r2) gather_photon(T, m, mode, nx, ny, nz, a);
            uint32_t ax = T.axis[m];
            double pa = ax == 0 ? (double)p.x : ax == 1 ? (double)p.y : (double)p.z;
            double diff = q[ax] - pa;
            /* near side first; the far side only if the split plane is within r */
            uint32_t nlo, nhi, flo, fhi;
            if (diff <= 0.0) {
                nlo = lo, nhi = m, flo = m + 1, fhi = hi;
            } else {
                nlo = m + 1, nhi = hi, flo = lo, fhi = m;
            }
            if (diff * diff <= r2 && flo < fhi && sp < PMAP_STACK) stk[(sp++) * stride] = make_uint2(flo, fhi);
            lo = nlo;
            hi = nhi;
        }
    }
}

/* Uniform grid over a photon map for fixed-radius gathers in shading: the
 * photons sorted by cell (x fastest), cell_start[c] = first photon of cell
 * c (ncells + 1 entries).  A query with radius r scans, for every (y, z)
 * row of cells its box meets, the one contiguous photon range of the x
 * run, with the same inclusion test and estimate as the kd-tree gather,
 * so the found set is the same. */
struct PmapGrid {
    PmapTree v;               /* photon arrays in cell order (axis unused) */
    const uint32_t *cell_start;
    float ox, oy, oz;         /* grid origin */
    float inv_cs;             /* 1 / cell size */
    int32_t dx, dy, dz;
};

__device__ __forceinline__ int grid_cell(float x, float o, float inv, int d)
{
    float f = floorf((x - o) * inv);
    if (!(f > 0.0f)) return 0; /* below the grid (or NaN) */
    return f > (float)(d - 1) ? d - 1 : (int)f;
}

__device__ __forceinline__ void gather_grid(const PmapGrid &G, float qx, float qy, float qz, float radius, int mode,
                                            double nx, double ny, double nz, GatherAcc &a)
{
    a.r = a.g = a.b = 0.0;
    a.count = 0;
    if (G.v.n == 0) return;
    double r2 = (double)radius * (double)radius;
    /* the cell range of the query box, widened so float rounding of the
     * box ends cannot drop a cell holding an included photon */
    float mx = 1e-5f * (fabsf(qx) + radius), my = 1e-5f * (fabsf(qy) + radius), mz = 1e-5f * (fabsf(qz) + radius);
    int x0 = grid_cell(qx - radius - mx, G.ox, G.inv_cs, G.dx), x1 = grid_cell(qx + radius + mx, G.ox, G.inv_cs, G.dx);
    int y0 = grid_cell(qy - radius - my, G.oy, G.inv_cs, G.dy), y1 = grid_cell(qy + radius + my, G.oy, G.inv_cs, G.dy);
    int z0 = grid_cell(qz - radius - mz, G.oz, G.inv_cs, G.dz), z1 = grid_cell(qz + radius + mz, G.oz, G.inv_cs, G.dz);
    for (int z = z0; z <= z1; z++)
        for (int y = y0; y <= y1; y++) {
            uint32_t row = ((uint32_t)z * (uint32_t)G.dy + (uint32_t)y) * (uint32_t)G.dx;
            uint32_t b = G.cell_start[row + (uint32_t)x0], e = G.cell_start[row + (uint32_t)x1 + 1u];
            for (uint32_t m = b; m < e; m++) {
                float4 p = G.v.pos[m];
                double ddx = (double)p.x - (double)qx, ddy = (double)p.y - (double)qy, ddz = (double)p.z - (double)qz;
                double d2 = ddx * ddx + ddy * ddy + ddz * ddz;
                if (d2 <= r2) gather_photon(G.v, m, mode, nx, ny, nz, a);
            }
        }
}

/* the same gather by a group of G consecutive lanes (G a power of two, at
 * most 64, all G lanes active): the lanes stride over each contiguous
 * photon range and their partial sums are combined with cross-lane
 * shuffles, so a query's photons are read G at a time.  Every lane of the
 * group returns the total. */
template <int G>
__device__ __forceinline__ void gather_grid_coop(const PmapGrid &Gd, float qx, float qy, float qz, float radius,
                                                 int mode, double nx, double ny, double nz, int lane, GatherAcc &a)
{
    a.r = a.g = a.b = 0.0;
    a.count = 0;
    if (Gd.v.n == 0) return;
    double r2 = (double)radius * (double)radius;
    float mx = 1e-5f * (fabsf(qx) + radius), my = 1e-5f * (fabsf(qy) + radius), mz = 1e-5f * (fabsf(qz) + radius);
    int x0 = grid_cell(qx - radius - mx, Gd.ox, Gd.inv_cs, Gd.dx), x1 = grid_cell(qx + radius + mx, Gd.ox, Gd.inv_cs, Gd.dx);
    int y0 = grid_cell(qy - radius - my, Gd.oy, Gd.inv_cs, Gd.dy), y1 = grid_cell(qy + radius + my, Gd.oy, Gd.inv_cs, Gd.dy);
    int z0 = grid_cell(qz - radius - mz, Gd.oz, Gd.inv_cs, Gd.dz), z1 = grid_cell(qz + radius + mz, Gd.oz, Gd.inv_cs, Gd.dz);
    for (int z = z0; z <= z1; z++)
        for (int y = y0; y <= y1; y++) {
            uint32_t row = ((uint32_t)z * (uint32_t)Gd.dy + (uint32_t)y) * (uint32_t)Gd.dx;
            uint32_t b = Gd.cell_start[row + (uint32_t)x0], e = Gd.cell_start[row + (uint32_t)x1 + 1u];
            for (uint32_t m = b + (uint32_t)lane; m < e; m += (uint32_t)G) {
                float4 p = Gd.v.pos[m];
                double ddx = (double)p.x - (double)qx, ddy = (double)p.y - (double)qy, ddz = (double)p.z - (double)qz;
                double d2 = ddx * ddx + ddy * ddy + ddz * ddz;
                if (d2 <= r2) gather_photon(Gd.v, m, mode, nx, ny, nz, a);
            }
        }
    for (int off = G / 2; off > 0; off >>= 1) {
        a.r += __shfl_xor(a.r, off, G);
        a.g += __shfl_xor(a.g, off, G);
        a.b += __shfl_xor(a.b, off, G);
        a.count += (uint32_t)__shfl_xor((int)a.count, off, G);
    }
}

/* ------------------------------------------------------------------ */
/* k nearest within the radius (PhotonMap::*_est's max_photons, unused */
/* by the reference, src/pmap.h:59-60): the photons of the range gather */
/* ordered by (dist^2, photon index), the first k kept                  */
/* ------------------------------------------------------------------ */
/* the photons of the query's cell rows, each lane of a group of G a
 * strided share, passed to f(m, d2) when within the radius */
template <int G, class F>
__device__ __forceinline__ void grid_scan(const PmapGrid &Gd, float qx, float qy, float qz, float radius, int lane, F f)
{
    const double r2 = (double)radius * (double)radius;
    float mx = 1e-5f * (fabsf(qx) + radius), my = 1e-5f * (fabsf(qy) + radius), mz = 1e-5f * (fabsf(qz) + radius);
    int x0 = grid_cell(qx - radius - mx, Gd.ox, Gd.inv_cs, Gd.dx), x1 = grid_cell(qx + radius + mx, Gd.ox, Gd.inv_cs, Gd.dx);
    int y0 = grid_cell(qy - radius - my, Gd.oy, Gd.inv_cs, Gd.dy), y1 = grid_cell(qy + radius + my, Gd.oy, Gd.inv_cs, Gd.dy);
    int z0 = grid_cell(qz - radius - mz, Gd.oz, Gd.inv_cs, Gd.dz), z1 = grid_cell(qz + radius + mz, Gd.oz, Gd.inv_cs, Gd.dz);
    for (int z = z0; z <= z1; z++)
        for (int y = y0; y <= y1; y++) {
            uint32_t row = ((uint32_t)z * (uint32_t)Gd.dy + (uint32_t)y) * (uint32_t)Gd.dx;
            uint32_t b = Gd.cell_start[row + (uint32_t)x0], e = Gd.cell_start[row + (uint32_t)x1 + 1u];
            for (uint32_t m = b + (uint32_t)lane; m < e; m += (uint32_t)G) {
                float4 p = Gd.v.pos[m];
                double ddx = (double)p.x - (double)qx, ddy = (double)p.y - (double)qy, ddz = (double)p.z - (double)qz;
                double d2 = ddx * ddx + ddy * ddy + ddz * ddz;
                if (d2 <= r2) f(m, d2, __float_as_uint(p.w));
            }
        }
}

template <int G> __device__ __forceinline__ uint32_t group_sum(uint32_t x)
{
    for (int off = G / 2; off > 0; off >>= 1) x += (uint32_t)__shfl_xor((int)x, off, G);
    return x;
}

/* LDS written by some lanes of this wave and read by others: every lane's
 * operation is issued and completed in program order, so only the
 * compiler's ordering needs holding */
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

/* the digit of the need-th smallest (1-based) of a group's 256-bin
 * histogram: digit and the count of smaller digits */
template <int G>
__device__ __forceinline__ void hist_select(const uint32_t *hist, int lane, uint32_t need, uint32_t &digit,
                                            uint32_t &below)
{
    constexpr int B = 256 / G;
    uint32_t loc = 0;
    for (int b = 0; b < B; b++) loc += hist[lane * B + b];
    uint32_t incl = loc;
    for (int d = 1; d < G; d <<= 1) {
        uint32_t y = (uint32_t)__shfl_up((int)incl, d, G);
        if (lane >= d) incl += y;
    }
    const uint32_t excl = incl - loc;
    uint32_t dg = 0, bl = 0;
    if (excl < need && need <= incl) {
        uint32_t run = excl;
        for (int b = 0; b < B; b++) {
            const uint32_t h = hist[lane * B + b];
            if (run + h >= need) {
                dg = (uint32_t)(lane * B + b);
                bl = run;
                break;
            }
            run += h;
        }
    }
    digit = group_sum<G>(dg); /* one lane holds the crossing */
    below = group_sum<G>(bl);
}

/* clears a group's histogram */
template <int G> __device__ __forceinline__ void hist_clear(uint32_t *hist, int lane)
{
    constexpr int B = 256 / G;
    for (int b = 0; b < B; b++) hist[lane * B + b] = 0u;
}

/* a group's LDS for the k-nearest selection: the histogram and the list
 * of the photons in the k-th photon's first-byte bin */
template <int G> struct KcapLds {
    static constexpr uint32_t CAP = 4 * G;
    uint32_t hist[256];
    uint32_t cnt, pad_[3];
    double d2[CAP];
    uint32_t idx[CAP], m[CAP];
};

/* The gather of the renderer's k-nearest mode by a group of G lanes:
 * the range gather (the same sums, in the same order, as
 * gather_grid_coop) and, when more than k photons are within the radius,
 * the estimate over the k nearest by (dist^2, photon index).  Selection:
 * a 32-bit key, dist^2 / r^2 in fixed point (monotone in dist^2), with a
 * 256-bin histogram of its first byte taken during the range gather.
 * The k-th photon's bin is then, most often, small: one more scan sums
 * the photons of the bins below it and lists the bin's photons in LDS,
 * which are ranked by (dist^2, index) there.  A larger bin is radix
 * selected a byte at a time, stopping as soon as the k-th photon's bin
 * holds exactly the photons still needed; photons that share the k-th key
 * go to an exact select of dist^2's double bits and then of the photon
 * index.  Every lane returns the total.  SUMS0 = false: the first scan
 * only counts and takes the histogram, and a radius holding fewer than k
 * photons returns false with no estimate (a query known to hold more than
 * k photons in a larger radius looks first in a smaller one, where its k
 * nearest most often are: every photon outside a radius is farther than
 * every photon inside it). */
template <int G, bool SUMS0 = true>
__device__ __forceinline__ bool gather_grid_kcap(const PmapGrid &Gd, float qx, float qy, float qz, float radius,
                                                 int mode, double nx, double ny, double nz, int lane, uint32_t k,
                                                 KcapLds<G> &L, GatherAcc &a)
{
    uint32_t *hist = L.hist;
    a.r = a.g = a.b = 0.0;
    a.count = 0;
    if (Gd.v.n == 0) return true;
    const double r2 = (double)radius * (double)radius;
    const double kscale = 4294967296.0 / r2;
    auto key32 = [&](double d2) -> uint32_t {
        const double f = d2 * kscale;
        return f >= 4294967295.0 ? 0xffffffffu : (uint32_t)f;
    };
    auto reduce = [&]() {
        for (int off = G / 2; off > 0; off >>= 1) {
            a.r += __shfl_xor(a.r, off, G);
            a.g += __shfl_xor(a.g, off, G);
            a.b += __shfl_xor(a.b, off, G);
            a.count += (uint32_t)__shfl_xor((int)a.count, off, G);
        }
    };
    hist_clear<G>(hist, lane);
    wave_lds_sync();
    grid_scan<G>(Gd, qx, qy, qz, radius, lane, [&](uint32_t m, double d2, uint32_t) {
        if (SUMS0) gather_photon(Gd.v, m, mode, nx, ny, nz, a);
        else a.count++;
        atomicAdd(&hist[key32(d2) >> 24], 1u);
    });
    wave_lds_sync();
    reduce();
    if (!SUMS0 && a.count < k) return false;
    if (a.count <= k) {
        if (!SUMS0) {
            a.r = a.g = a.b = 0.0;
            a.count = 0;
            grid_scan<G>(Gd, qx, qy, qz, radius, lane,
                         [&](uint32_t m, double, uint32_t) { gather_photon(Gd.v, m, mode, nx, ny, nz, a); });
            reduce();
        }
        return true;
    }
    /* the first byte of the k-th key */
    uint32_t need = k, digit, below;
    hist_select<G>(hist, lane, need, digit, below);
    need -= below;
    const uint32_t in_bin0 = hist[digit];
    wave_lds_sync();
    if (in_bin0 <= KcapLds<G>::CAP) {
        if (lane == 0) L.cnt = 0u;
        wave_lds_sync();
        const bool all = in_bin0 == need;
        a.r = a.g = a.b = 0.0;
        a.count = 0;
        grid_scan<G>(Gd, qx, qy, qz, radius, lane, [&](uint32_t m, double d2, uint32_t idx) {
            const uint32_t kb = key32(d2) >> 24;
            if (kb < digit || (kb == digit && all)) {
                gather_photon(Gd.v, m, mode, nx, ny, nz, a);
            } else if (kb == digit) {
                const uint32_t sl = atomicAdd(&L.cnt, 1u);
                L.d2[sl] = d2;
                L.idx[sl] = idx;
                L.m[sl] = m;
            }
        });
        wave_lds_sync();
        if (!all)
            for (uint32_t e = (uint32_t)lane; e < in_bin0; e += (uint32_t)G) {
                const double de = L.d2[e];
                const uint32_t ie = L.idx[e];
                uint32_t rank = 0;
                for (uint32_t f = 0; f < in_bin0; f++) {
                    const double df = L.d2[f];
                    rank += (df < de || (df == de && L.idx[f] < ie)) ? 1u : 0u;
                }
                if (rank < need) gather_photon(Gd.v, L.m[e], mode, nx, ny, nz, a);
            }
        wave_lds_sync();
        reduce();
        return true;
    }
    /* a large bin: the rest of the k-th key, a byte at a time */
    uint32_t prefix = digit << 24, pmask = 0xff000000u;
    bool whole = in_bin0 == need; /* every photon of the bin is taken */
    for (int shift = 24; !whole && shift > 0;) {
        hist_clear<G>(hist, lane);
        wave_lds_sync();
        grid_scan<G>(Gd, qx, qy, qz, radius, lane, [&](uint32_t, double d2, uint32_t) {
            const uint32_t kk = key32(d2);
            if ((kk & pmask) == prefix) atomicAdd(&hist[(kk >> (shift - 8)) & 255u], 1u);
        });
        wave_lds_sync();
        shift -= 8;
        hist_select<G>(hist, lane, need, digit, below);
        need -= below;
        prefix |= digit << shift;
        pmask |= 0xffu << shift;
        const uint32_t in_bin = hist[digit];
        wave_lds_sync();
        whole = in_bin == need;
    }
    /* the photons at the k-th key: need of them, by dist^2's bits, then
     * by photon index */
    uint64_t p64 = ~0ull;
    uint32_t ilim = 0xffffffffu;
    if (!whole) {
        uint64_t pre = 0, msk = 0;
        for (int shift = 56; shift >= 0; shift -= 8) {
            hist_clear<G>(hist, lane);
            wave_lds_sync();
            grid_scan<G>(Gd, qx, qy, qz, radius, lane, [&](uint32_t, double d2, uint32_t) {
                const uint64_t kb = (uint64_t)__double_as_longlong(d2);
                if (key32(d2) == prefix && (kb & msk) == pre) atomicAdd(&hist[(uint32_t)(kb >> shift) & 255u], 1u);
            });
            wave_lds_sync();
            hist_select<G>(hist, lane, need, digit, below);
            wave_lds_sync();
            need -= below;
            pre |= (uint64_t)digit << shift;
            msk |= 0xffull << shift;
        }
        p64 = pre;
        const uint32_t at_key = hist[digit];
        wave_lds_sync();
        if (need < at_key) {
            uint32_t ipre = 0, imask = 0;
            for (int shift = 24; shift >= 0; shift -= 8) {
                hist_clear<G>(hist, lane);
                wave_lds_sync();
                grid_scan<G>(Gd, qx, qy, qz, radius, lane, [&](uint32_t, double d2, uint32_t idx) {
                    if ((uint64_t)__double_as_longlong(d2) == p64 && (idx & imask) == ipre)
                        atomicAdd(&hist[(idx >> shift) & 255u], 1u);
                });
                wave_lds_sync();
                hist_select<G>(hist, lane, need, digit, below);
                wave_lds_sync();
                need -= below;
                ipre |= digit << shift;
                imask |= 0xffu << shift;
            }
            ilim = ipre;
        }
    }
    a.r = a.g = a.b = 0.0;
    a.count = 0;
    grid_scan<G>(Gd, qx, qy, qz, radius, lane, [&](uint32_t m, double d2, uint32_t idx) {
        const uint32_t kk = key32(d2);
        bool take;
        if (whole) take = (kk & pmask) <= prefix;
        else {
            const uint64_t kb = (uint64_t)__double_as_longlong(d2);
            take = kk < prefix || (kk == prefix && (kb < p64 || (kb == p64 && idx <= ilim)));
        }
        if (take) gather_photon(Gd.v, m, mode, nx, ny, nz, a);
    });
    reduce();
    return true;
}

} // namespace srd

#endif
