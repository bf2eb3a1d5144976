/*
 * sray_gpu.hip — gfx950 kernels and the C ABI runtime of the S-ray hot path.
 *
 * One context = one HIP stream (replaces the reference's pthread tile pool,
 * src/tpool.cc:48-225).  Scenes are built on the host (octrees of both
 * levels, src/octree.inl:100-191) and kept resident in HBM; the render is a
 * wavefront of bounce levels (render.h).  Built with -ffp-contract=off so
 * every float operation is the IEEE one the CPU checker performs.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cerrno>
#include <cstdarg>
#include <unistd.h>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include <new>

#include "runtime.h"
#include "pmap.h"
#include "surface.h"
#include "photon.h"

using namespace srd;
using namespace srt;

/* ================================================================== */
/* kernels                                                              */
/* ================================================================== */

template <bool ANY, bool COUNT>
__global__ __launch_bounds__(TRACE_BLOCK) void k_trace_rays(SceneDev S, const sray_ray *__restrict__ rays, uint32_t n,
                                                          sray_hit *__restrict__ hits,
                                                          unsigned long long *__restrict__ counts)
{
    __shared__ uint32_t stk[STACK * TRACE_BLOCK];
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    TraceCounts cnt = {0, 0, 0, 0, 0, 0, 0};
    if (i < n) {
        sray_ray r = rays[i];
        HitRec h = traverse<ANY, COUNT>(S, v3(r.origin[0], r.origin[1], r.origin[2]), v3(r.dir[0], r.dir[1], r.dir[2]),
                                        r.time, &stk[threadIdx.x], TRACE_BLOCK, cnt);
        sray_hit o;
        o.t = h.t;
        o.obj = h.obj;
        o.prim = (h.obj >= 0 && h.tri != 0xffffffffu) ? (int32_t)(h.tri - S.objs[h.obj].tri_base) : -1;
        o.rank = h.obj >= 0 ? h.rank : 0u;
        if (h.obj < 0) {
            o.t = FLT_MAX;
            o.rank = 0;
        }
        hits[i] = o;
    }
    flush_counts<COUNT>(cnt, counts);
}

/* the same through the instance of the render path's level kernels: every
 * lane of a wave walks until all are done, and waves with few lanes in
 * mesh leaves test those leaves together (SPARSE: four triangles in flight
 * per lane in its own leaf) */
template <bool ANY, bool SPARSE>
__global__ __launch_bounds__(TRACE_BLOCK) void k_trace_rays_wave(SceneDev S, const sray_ray *__restrict__ rays,
                                                               uint32_t n, sray_hit *__restrict__ hits)
{
    __shared__ uint32_t stk[STACK * TRACE_BLOCK];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = i < n;
    TraceCounts cnt = {0, 0, 0, 0, 0, 0, 0};
    RayRegs src;
    if (live) {
        const sray_ray r = rays[i];
        src = RayRegs{v3(r.origin[0], r.origin[1], r.origin[2]), v3(r.dir[0], r.dir[1], r.dir[2]), r.time};
    } else {
        src = RayRegs{v3(0.0f, 0.0f, 0.0f), v3(1.0f, 1.0f, 1.0f), 0};
    }
    const HitRec h = traverse_src<ANY, false, true, SPARSE>(trav_scene_uniform(S), src, live, &stk[threadIdx.x],
                                                            TRACE_BLOCK, cnt);
    if (!live) return;
    sray_hit o;
    o.t = h.t;
    o.obj = h.obj;
    o.prim = (h.obj >= 0 && h.tri != 0xffffffffu) ? (int32_t)(h.tri - S.objs[h.obj].tri_base) : -1;
    o.rank = h.obj >= 0 ? h.rank : 0u;
    if (h.obj < 0) {
        o.t = FLT_MAX;
        o.rank = 0;
    }
    hits[i] = o;
}

/* ------------------------------------------------------------------ */
/* photon-map kd-tree build (level by level) and gathers                */
/* ------------------------------------------------------------------ */
__device__ __forceinline__ uint32_t fkey(float f)
{
    uint32_t b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

__global__ void k_pm_init(const sray_photon *__restrict__ ph, uint32_t n, float4 *__restrict__ pos0,
                          uint32_t *__restrict__ ord)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    pos0[i] = make_float4(ph[i].pos[0], ph[i].pos[1], ph[i].pos[2], 0.0f);
    ord[i] = i;
}

/* segment of every array position at this level (-1: a finished node) */
__global__ void k_pm_seg(const uint32_t *__restrict__ slo, const uint32_t *__restrict__ shi, uint32_t nseg,
                         uint32_t n, int32_t *__restrict__ seg_of)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t a = 0, b = nseg;
    while (b - a > 1) {
        uint32_t m = (a + b) >> 1;
        if (slo[m] <= i) a = m;
        else b = m;
    }
    seg_of[i] = (nseg && slo[a] <= i && i < shi[a]) ? (int32_t)a : -1;
}

__global__ void k_pm_bbox(const uint32_t *__restrict__ ord, const float4 *__restrict__ pos0,
                          const int32_t *__restrict__ seg_of, uint32_t n, uint32_t *__restrict__ bb)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    int s = i < n ? seg_of[i] : -1;
    float4 p = s >= 0 ? pos0[ord[i]] : make_float4(0, 0, 0, 0);
    uint32_t k[3] = {fkey(p.x), fkey(p.y), fkey(p.z)};
    /* whole-wave fast path: one segment for all active lanes */
    int s0 = __shfl(s, 0);
    bool same = __all(s == s0) && s0 >= 0;
    if (same) {
        for (int a = 0; a < 3; a++) {
            uint32_t mn = k[a], mx = k[a];
            for (int off = 32; off > 0; off >>= 1) {
                mn = min(mn, (uint32_t)__shfl_xor((int)mn, off));
                mx = max(mx, (uint32_t)__shfl_xor((int)mx, off));
            }
            if ((threadIdx.x & 63) == 0) {
                atomicMin(&bb[s0 * 6 + a], mn);
                atomicMax(&bb[s0 * 6 + 3 + a], mx);
            }
        }
    } else if (s >= 0) {
        for (int a = 0; a < 3; a++) {
            atomicMin(&bb[s * 6 + a], k[a]);
            atomicMax(&bb[s * 6 + 3 + a], k[a]);
        }
    }
}

__device__ __forceinline__ float unkey(uint32_t k)
{
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}

__global__ void k_pm_bbinit(uint32_t *bb, uint32_t nseg)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nseg * 6) return;
    bb[i] = (i % 6) < 3 ? 0xffffffffu : 0u;
}

/* split axis = largest extent of the segment (ties: lowest axis) */
__global__ void k_pm_axis(const uint32_t *__restrict__ bb, uint32_t nseg, uint8_t *__restrict__ saxis)
{
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nseg) return;
    float e[3];
    for (int a = 0; a < 3; a++) e[a] = unkey(bb[s * 6 + 3 + a]) - unkey(bb[s * 6 + a]);
    uint8_t ax = 0;
    if (e[1] > e[ax]) ax = 1;
    if (e[2] > e[ax]) ax = 2;
    saxis[s] = ax;
}

__global__ void k_pm_keys(const uint32_t *__restrict__ ord, const float4 *__restrict__ pos0,
                          const int32_t *__restrict__ seg_of, const uint32_t *__restrict__ slo,
                          const uint8_t *__restrict__ saxis, uint32_t n, unsigned long long *__restrict__ key)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int s = seg_of[i];
    if (s < 0) {
        key[i] = (unsigned long long)i << 32;
        return;
    }
    float4 p = pos0[ord[i]];
    uint32_t a = saxis[s];
    float c = a == 0 ? p.x : a == 1 ? p.y : p.z;
    key[i] = ((unsigned long long)slo[s] << 32) | fkey(c);
}

__global__ void k_pm_nodes(const uint32_t *__restrict__ slo, const uint32_t *__restrict__ shi,
                           const uint8_t *__restrict__ saxis, uint32_t nseg, uint8_t *__restrict__ axis)
{
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nseg) return;
    axis[(slo[s] + shi[s]) >> 1] = saxis[s];
}

__global__ void k_pm_final(const sray_photon *__restrict__ ph, const uint32_t *__restrict__ ord, uint32_t n,
                           float4 *__restrict__ pos, float4 *__restrict__ dir, float4 *__restrict__ nrm,
                           float *__restrict__ colb)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t o = ord[i];
    const sray_photon &p = ph[o];
    pos[i] = make_float4(p.pos[0], p.pos[1], p.pos[2], __uint_as_float(o));
    dir[i] = make_float4(p.dir[0], p.dir[1], p.dir[2], p.col[0]);
    nrm[i] = make_float4(p.norm[0], p.norm[1], p.norm[2], p.col[1]);
    colb[i] = p.col[2];
}

/* shading-gather grid build */
__global__ void k_grid_bbox(const float4 *__restrict__ pos, uint32_t n, uint32_t *__restrict__ bb)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 p = pos[i];
    uint32_t k[3] = {fkey(p.x), fkey(p.y), fkey(p.z)};
    for (int a = 0; a < 3; a++) {
        atomicMin(&bb[a], k[a]);
        atomicMax(&bb[3 + a], k[a]);
    }
}

__global__ void k_grid_key(const float4 *__restrict__ pos, uint32_t n, PmapGrid G, uint32_t *__restrict__ key,
                           uint32_t *__restrict__ val)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 p = pos[i];
    uint32_t x = (uint32_t)grid_cell(p.x, G.ox, G.inv_cs, G.dx), y = (uint32_t)grid_cell(p.y, G.oy, G.inv_cs, G.dy),
             z = (uint32_t)grid_cell(p.z, G.oz, G.inv_cs, G.dz);
    key[i] = (z * (uint32_t)G.dy + y) * (uint32_t)G.dx + x;
    val[i] = i;
}

/* cell_start[c] = first sorted photon with key >= c, for c in [0, ncell]:
 * one lower-bound search per cell (runs of empty cells cost nothing extra) */
__global__ void k_grid_start(const uint32_t *__restrict__ key, uint32_t n, uint32_t ncell, uint32_t *__restrict__ start)
{
    uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c > ncell) return;
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        uint32_t mid = lo + ((hi - lo) >> 1);
        if (key[mid] < c)
            lo = mid + 1;
        else
            hi = mid;
    }
    start[c] = lo;
}

__global__ void k_grid_final(PmapTree T, const uint32_t *__restrict__ order, uint32_t n, float4 *__restrict__ pos,
                             float4 *__restrict__ dir, float4 *__restrict__ nrm, float *__restrict__ colb)
{
    uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    uint32_t o = order[j];
    pos[j] = T.pos[o];
    dir[j] = T.dir[o];
    nrm[j] = T.nrm[o];
    colb[j] = T.colb[o];
}

__global__ void k_iota(uint32_t *p, uint32_t n)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = i;
}

__device__ __forceinline__ void finish_estimate(const GatherAcc &a, int mode, double scale, float *out)
{
    if (mode == SRAY_GATHER_IRRADIANCE && a.count < 8) {
        out[0] = out[1] = out[2] = 0.0f;
        return;
    }
    out[0] = (float)(a.r * scale);
    out[1] = (float)(a.g * scale);
    out[2] = (float)(a.b * scale);
}

/* range gather; with k > 0, queries holding more than k photons are flagged
 * for the k-nearest pass */
__global__ __launch_bounds__(128) void k_pm_gather(PmapTree T, const sray_gather_query *__restrict__ q, uint32_t n,
                                                   float radius, int mode, int k, float *__restrict__ rgb,
                                                   uint32_t *__restrict__ count, uint8_t *__restrict__ need)
{
    __shared__ uint2 stk[PMAP_STACK * 128];
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const sray_gather_query &Q = q[i];
    GatherAcc a;
    gather_range(T, Q.pos[0], Q.pos[1], Q.pos[2], radius, mode, (double)Q.norm[0], (double)Q.norm[1],
                 (double)Q.norm[2], a, &stk[threadIdx.x], 128);
    count[i] = a.count;
    bool sel = k > 0 && a.count > (uint32_t)k;
    need[i] = sel ? 1 : 0;
    if (!sel) finish_estimate(a, mode, 1.0 / (M_PI * ((double)radius * (double)radius)), rgb + (size_t)i * 3);
}

/* fixed-radius gather over the uniform grid (same found set and estimate),
 * 16 lanes per query */
__global__ __launch_bounds__(256) void k_pm_gather_grid(PmapGrid G, const sray_gather_query *__restrict__ q,
                                                        uint32_t n, float radius, int mode, int k,
                                                        float *__restrict__ rgb)
{
    __shared__ KcapLds<16> lds[256 / 16];
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t i = t / 16u;
    int lane = (int)(t % 16u);
    if (i >= n) return;
    const sray_gather_query &Q = q[i];
    GatherAcc a;
    if (k > 0)
        gather_grid_kcap<16>(G, Q.pos[0], Q.pos[1], Q.pos[2], radius, mode, (double)Q.norm[0], (double)Q.norm[1],
                             (double)Q.norm[2], lane, (uint32_t)k, lds[threadIdx.x / 16], a);
    else
        gather_grid_coop<16>(G, Q.pos[0], Q.pos[1], Q.pos[2], radius, mode, (double)Q.norm[0], (double)Q.norm[1],
                             (double)Q.norm[2], lane, a);
    if (lane == 0) finish_estimate(a, mode, 1.0 / (M_PI * ((double)radius * (double)radius)), rgb + (size_t)i * 3);
}

struct Cand {
    double d2;
    uint32_t idx;  /* original photon index (tie-break) */
    uint32_t pos;  /* tree position */
};

__global__ __launch_bounds__(128) void k_pm_collect(PmapTree T, const sray_gather_query *__restrict__ q,
                                                    const uint32_t *__restrict__ fq, uint32_t nf,
                                                    const uint32_t *__restrict__ off, float radius,
                                                    Cand *__restrict__ cand)
{
    __shared__ uint2 stk[PMAP_STACK * 128];
    uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nf) return;
    const sray_gather_query &Q = q[fq[j]];
    uint2 *st = &stk[threadIdx.x];
    const int stride = 128;
    double r2 = (double)radius * (double)radius;
    double qq[3] = {(double)Q.pos[0], (double)Q.pos[1], (double)Q.pos[2]};
    uint32_t w = off[j];
    int sp = 1;
    st[0] = make_uint2(0u, T.n);
    while (sp > 0) {
        uint2 s = st[(--sp) * stride];
        uint32_t lo = s.x, hi = s.y;
        while (lo < hi) {
            uint32_t m = (lo + hi) >> 1;
            float4 p = T.pos[m];
            double dx = (double)p.x - qq[0], dy = (double)p.y - qq[1], dz = (double)p.z - qq[2];
            double d2 = dx * dx + dy * dy + dz * dz;
            if (d2 <= r2) {
                Cand c;
                c.d2 = d2;
                c.idx = __float_as_uint(p.w);
                c.pos = m;
                cand[w++] = c;
            }
            uint32_t ax = T.axis[m];
            double pa = ax == 0 ? (double)p.x : ax == 1 ? (double)p.y : (double)p.z;
            double diff = qq[ax] - pa;
            uint32_t nlo, nhi, flo, fhi;
            if (diff <= 0.0) {
                nlo = lo, nhi = m, flo = m + 1, fhi = hi;
            } else {
                nlo = m + 1, nhi = hi, flo = lo, fhi = m;
            }
            if (diff * diff <= r2 && flo < fhi && sp < PMAP_STACK) st[(sp++) * stride] = make_uint2(flo, fhi);
            lo = nlo;
            hi = nhi;
        }
    }
}

__device__ __forceinline__ bool cand_less(const Cand &a, const Cand &b)
{
    return a.d2 < b.d2 || (a.d2 == b.d2 && a.idx < b.idx);
}

/* k smallest (d2, photon index) of each flagged query's candidates
 * (in-place quickselect), then the estimate over them */
__global__ void k_pm_select(PmapTree T, const sray_gather_query *__restrict__ q, const uint32_t *__restrict__ fq,
                            uint32_t nf, const uint32_t *__restrict__ off, const uint32_t *__restrict__ cnt,
                            Cand *__restrict__ cand, int k, float radius, int mode, float *__restrict__ rgb)
{
    uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nf) return;
    uint32_t qi = fq[j];
    Cand *c = cand + off[j];
    int lo = 0, hi = (int)cnt[qi] - 1, kk = k - 1;
    while (lo < hi) {
        Cand pv = c[(lo + hi) >> 1];
        int a = lo, b = hi;
        while (a <= b) {
            while (cand_less(c[a], pv)) a++;
            while (cand_less(pv, c[b])) b--;
            if (a <= b) {
                Cand t = c[a];
                c[a] = c[b];
                c[b] = t;
                a++;
                b--;
            }
        }
        if (kk <= b) hi = b;
        else if (kk >= a) lo = a;
        else break;
    }
    const sray_gather_query &Q = q[qi];
    GatherAcc acc;
    acc.r = acc.g = acc.b = 0.0;
    acc.count = 0;
    for (int e = 0; e < k; e++)
        gather_photon(T, c[e].pos, mode, (double)Q.norm[0], (double)Q.norm[1], (double)Q.norm[2], acc);
    finish_estimate(acc, mode, 1.0 / (M_PI * ((double)radius * (double)radius)), rgb + (size_t)qi * 3);
}

/* ================================================================== */
/* runtime                                                              */
/* ================================================================== */
namespace srt {

namespace {
/* The item box codes of one triangle (v: its three float4 records) in a
 * leaf frame (per axis base and step as the device computes them).  The
 * planes base + q * step are evaluated exactly (in double, from the same
 * float base and step), and the codes taken on the safe side: the lower
 * plane at or below the triangle's minimum, the upper at or above its
 * maximum.  A triangle outside the frame's planes, or whose float test
 * could accept points measurably off its box (a non-finite vertex, or twice
 * its area below 2^-10 of its longest edge squared: a sliver), is marked
 * IBOX_OPEN. */
uint2 ibox_codes(const float4 *v, const float base[3], const float step[3])
{
    const uint2 open = make_uint2(0xff000000u, 0xffffu | IBOX_OPEN); /* x0 y0 z0 = 0, x1 y1 z1 = 255 */
    double p[3][3];
    for (int i = 0; i < 3; i++) {
        p[i][0] = v[i].x;
        p[i][1] = v[i].y;
        p[i][2] = v[i].z;
        for (int a = 0; a < 3; a++)
            if (!std::isfinite(p[i][a])) return open;
    }
    double e[3][3], len2 = 0;
    for (int a = 0; a < 3; a++) {
        e[0][a] = p[1][a] - p[0][a];
        e[1][a] = p[2][a] - p[0][a];
        e[2][a] = p[2][a] - p[1][a];
    }
    for (int k = 0; k < 3; k++) len2 = std::max(len2, e[k][0] * e[k][0] + e[k][1] * e[k][1] + e[k][2] * e[k][2]);
    const double nx = e[0][1] * e[1][2] - e[0][2] * e[1][1], ny = e[0][2] * e[1][0] - e[0][0] * e[1][2],
                 nz = e[0][0] * e[1][1] - e[0][1] * e[1][0];
    if (!(std::sqrt(nx * nx + ny * ny + nz * nz) >= len2 * 0x1p-10)) return open;
    uint32_t code[6];
    for (int a = 0; a < 3; a++) {
        const double lo = std::min(p[0][a], std::min(p[1][a], p[2][a]));
        const double hi = std::max(p[0][a], std::max(p[1][a], p[2][a]));
        const double b = base[a], s = step[a];
        auto plane = [&](int q) { return b + (double)q * s; };
        if (!(plane(0) <= lo) || !(plane(255) >= hi)) return open;
        int ql = s > 0 ? (int)std::min(255.0, std::max(0.0, std::floor((lo - b) / s))) : 0;
        while (ql > 0 && plane(ql) > lo) ql--;
        while (ql < 255 && plane(ql + 1) <= lo) ql++;
        int qh = s > 0 ? (int)std::min(255.0, std::max(0.0, std::ceil((hi - b) / s))) : 0;
        while (qh < 255 && plane(qh) < hi) qh++;
        while (qh > 0 && plane(qh - 1) >= hi) qh--;
        code[a] = (uint32_t)ql;
        code[3 + a] = (uint32_t)qh;
    }
    return make_uint2(code[0] | code[1] << 8 | code[2] << 16 | code[3] << 24, code[4] | code[5] << 8);
}

/* item box codes of every mesh leaf's item positions (scene positions and
 * unused slots stay zero) */
void leaf_item_boxes(const sray::FlatTree &ft, size_t scene_items, const std::vector<float4> &tri,
                     std::vector<uint2> &out)
{
    out.assign(ft.items.size(), make_uint2(0, 0));
    const size_t nn = ft.lo.size() / 4;
    for (size_t nd = 0; nd < nn; nd++) {
        uint32_t first, meta;
        memcpy(&first, &ft.lo[nd * 4 + 3], 4);
        memcpy(&meta, &ft.hi[nd * 4 + 3], 4);
        if (!(meta & LEAF) || first < scene_items) continue;
        float base[3], step[3];
        for (int a = 0; a < 3; a++) {
            const float lo = ft.lo[nd * 4 + a], hi = ft.hi[nd * 4 + a];
            const float ext = hi - lo;
            base[a] = lo - ext * IBOX_GROW;
            step[a] = ext * IBOX_STEP;
        }
        const uint32_t end = first + (meta & ~LEAF);
        for (uint32_t k = first; k < end; k++) out[k] = ibox_codes(&tri[(size_t)ft.items[k] * 3], base, step);
    }
}
} // namespace

float host_unkey(uint32_t k)
{
    uint32_t b = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
    float f;
    memcpy(&f, &b, 4);
    return f;
}

int fail(sray_gpu_ctx *c, int code, const char *fmt, ...)
{
    char b[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(b, sizeof b, fmt, ap);
    va_end(ap);
    if (c) c->err = b;
    return code;
}

int ensure(sray_gpu_ctx *c, DevBuf &b, size_t bytes)
{
    if (bytes == 0) bytes = 16;
    if (b.bytes >= bytes) return SRAY_OK;
    if (b.p) {
        HIPCHK(c, hipStreamSynchronize(c->stream));
        HIPCHK(c, hipFree(b.p));
        b.p = nullptr;
        b.bytes = 0;
    }
    size_t want = bytes + bytes / 4;
    hipError_t e = hipMalloc(&b.p, want);
    if (e != hipSuccess) {
        b.p = nullptr;
        return fail(c, SRAY_ERR_NOMEM, "hipMalloc(%zu) failed: %s", want, hipGetErrorString(e));
    }
    b.bytes = want;
    return SRAY_OK;
}

int ensure_zero(sray_gpu_ctx *c, DevBuf &b, size_t bytes)
{
    if (b.bytes >= bytes) return SRAY_OK;
    int rc = ensure(c, b, bytes);
    if (rc) return rc;
    HIPCHK(c, hipMemsetAsync(b.p, 0, b.bytes, c->stream));
    return SRAY_OK;
}

void release(DevBuf &b)
{
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.bytes = 0;
}

int upload(sray_gpu_ctx *c, GpuScene &g, const void *src, size_t bytes, const void **dst)
{
    DevBuf b;
    if (bytes == 0) {
        *dst = nullptr;
        return SRAY_OK;
    }
    hipError_t e = hipMalloc(&b.p, bytes);
    if (e != hipSuccess) return fail(c, SRAY_ERR_NOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    b.bytes = bytes;
    g.bufs.push_back(b);
    g.bytes += bytes;
    HIPCHK(c, hipMemcpyAsync(b.p, src, bytes, hipMemcpyHostToDevice, c->stream));
    *dst = b.p;
    return SRAY_OK;
}

hipEvent_t prof_event(sray_gpu_ctx *c)
{
    if (!c->evpool.empty()) {
        hipEvent_t e = c->evpool.back();
        c->evpool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

/* after a stream synchronisation: fold pending event pairs into the stats */
void prof_resolve(sray_gpu_ctx *c)
{
    for (auto &p : c->pending) {
        float ms = 0;
        if (hipEventSynchronize(p.b) == hipSuccess && hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
            auto &k = c->kstats[p.name];
            k.launches++;
            k.ms += ms;
        }
        c->evpool.push_back(p.a);
        c->evpool.push_back(p.b);
    }
    c->pending.clear();
}

void emit_status(sray_gpu_ctx *c, char st, int a1, int a2, int a3, int a4)
{
    const int fd = c->status_fd.load();
    if (fd < 0) return;
    char rec[SRAY_STATUS_RECORD];
    if (sray_status_record(st, a1, a2, a3, a4, rec) != SRAY_OK) return;
    std::lock_guard<std::mutex> lk(c->status_mu);
    size_t off = 0;
    while (off < sizeof rec) {
        const ssize_t w = write(fd, rec + off, sizeof rec - off);
        if (w < 0 && errno == EINTR) continue;
        if (w <= 0) break; /* a closed or full stream drops the record */
        off += (size_t)w;
    }
}

GpuScene *find_scene(sray_gpu_ctx *c, uint32_t id)
{
    auto it = c->scenes.find(id);
    return it == c->scenes.end() ? nullptr : it->second.get();
}

int validate_desc(sray_gpu_ctx *c, const sray_scene_desc *s)
{
    if (!s) return fail(c, SRAY_ERR_ARG, "scene description is NULL");
    if (s->num_objects && !s->objects) return fail(c, SRAY_ERR_ARG, "objects is NULL");
    if (s->num_meshes && !s->meshes) return fail(c, SRAY_ERR_ARG, "meshes is NULL");
    if (s->num_triangles && !s->tri_pos) return fail(c, SRAY_ERR_ARG, "tri_pos is NULL");
    if (s->num_lights && !s->lights) return fail(c, SRAY_ERR_ARG, "lights is NULL");
    if (s->num_materials && !s->materials) return fail(c, SRAY_ERR_ARG, "materials is NULL");
    if (s->scene_oct_depth < 0 || s->mesh_oct_depth < 0 || s->scene_oct_items < 0 || s->mesh_oct_items < 0)
        return fail(c, SRAY_ERR_ARG, "negative octree limits");
    if (s->num_lights > 255) return fail(c, SRAY_ERR_LIMIT, "more than 255 lights");
    for (uint32_t i = 0; i < s->num_objects; i++) {
        const sray_object_desc &o = s->objects[i];
        if (o.type < 0 || o.type > 2) return fail(c, SRAY_ERR_ARG, "object %u: bad type %d", i, o.type);
        if (o.material >= (int)s->num_materials) return fail(c, SRAY_ERR_ARG, "object %u: bad material", i);
        if (o.type == SRAY_OBJ_MESH && (o.mesh < 0 || (uint32_t)o.mesh >= s->num_meshes))
            return fail(c, SRAY_ERR_ARG, "object %u: bad mesh index", i);
    }
    for (uint32_t m = 0; m < s->num_meshes; m++) {
        uint64_t e = (uint64_t)s->meshes[m].tri_start + s->meshes[m].tri_count;
        if (e > s->num_triangles) return fail(c, SRAY_ERR_ARG, "mesh %u: triangle range out of bounds", m);
    }
    if (s->num_textures && !s->textures) return fail(c, SRAY_ERR_ARG, "textures is NULL");
    for (uint32_t i = 0; i < s->num_textures; i++) {
        const sray_texture_desc &t = s->textures[i];
        if (t.width <= 0 || t.height <= 0 || (int64_t)t.width * t.height > (1ll << 28) || !t.texels)
            return fail(c, SRAY_ERR_ARG, "texture %u: bad size %dx%d or no texels", i, t.width, t.height);
        if (t.filter < 0 || t.filter > 1 || t.wrap < 0 || t.wrap > 1)
            return fail(c, SRAY_ERR_ARG, "texture %u: bad filter or wrap", i);
    }
    for (uint32_t i = 0; i < s->num_materials; i++)
        for (int k = 0; k < SRAY_NUM_MAPS; k++)
            if (s->materials[i].map[k] < 0 || (uint32_t)s->materials[i].map[k] > s->num_textures)
                return fail(c, SRAY_ERR_ARG, "material %u: map %d names texture %d of %u", i, k,
                            s->materials[i].map[k], s->num_textures);
    bool rows = s->camera.time_xforms != nullptr;
    for (uint32_t i = 0; i < s->num_objects; i++) rows = rows || s->objects[i].time_xforms;
    for (uint32_t i = 0; i < s->num_lights; i++) rows = rows || s->lights[i].time_pos;
    if (rows && (s->frame_t1 < s->frame_t0 || (int64_t)s->frame_t1 - s->frame_t0 >= (1 << 20)))
        return fail(c, SRAY_ERR_ARG, "keyframed rows need a frame interval of 1 .. 2^20 msec (frame_t0 %d, frame_t1 %d)",
                    s->frame_t0, s->frame_t1);
    return SRAY_OK;
}

int do_scene_upload(sray_gpu_ctx *c, const sray_scene_desc *s, uint32_t *scene_id)
{
    int rc = validate_desc(c, s);
    if (rc) return rc;
    auto g = std::make_unique<GpuScene>();
    /* octrees: scene level first (root 0), then one per mesh */
    sray::FlatTree ft;
    std::vector<float> boxes;
    sray::object_boxes(s, boxes);
    uint32_t sroot = sray::build_flat_octree(boxes, s->scene_oct_items, s->scene_oct_depth, 0, ft);
    if (ft.overflow) return fail(c, SRAY_ERR_LIMIT, "scene octree passes 2^23 nodes or 2^28 item references");
    const size_t scene_items = ft.items.size(); /* mesh item positions follow */
    g->scene_height = ft.height;
    double dx = (double)ft.root_max[0] - ft.root_min[0], dy = (double)ft.root_max[1] - ft.root_min[1],
           dz = (double)ft.root_max[2] - ft.root_min[2];
    g->diag = s->num_objects ? std::sqrt(dx * dx + dy * dy + dz * dz) : 0.0;
    std::vector<uint32_t> mesh_root(s->num_meshes);
    uint32_t mh = 0;
    for (uint32_t m = 0; m < s->num_meshes; m++) {
        sray::triangle_boxes(s, m, boxes);
        sray::FlatTree tmp;
        /* append into ft: build into tmp then splice with offsets */
        uint32_t node_off = (uint32_t)(ft.lo.size() / 4), item_off = (uint32_t)ft.items.size();
        sray::build_flat_octree(boxes, s->mesh_oct_items, s->mesh_oct_depth, s->meshes[m].tri_start, tmp);
        if (tmp.overflow)
            return fail(c, SRAY_ERR_LIMIT, "octree of mesh %u passes 2^23 nodes or 2^28 item references", m);
        mh = std::max(mh, tmp.height);
        for (size_t k = 0; k < tmp.lo.size() / 4; k++) {
            uint32_t first, meta;
            memcpy(&first, &tmp.lo[k * 4 + 3], 4);
            memcpy(&meta, &tmp.hi[k * 4 + 3], 4);
            first += (meta & LEAF) ? item_off : node_off;
            memcpy(&tmp.lo[k * 4 + 3], &first, 4);
        }
        ft.lo.insert(ft.lo.end(), tmp.lo.begin(), tmp.lo.end());
        ft.hi.insert(ft.hi.end(), tmp.hi.begin(), tmp.hi.end());
        ft.items.insert(ft.items.end(), tmp.items.begin(), tmp.items.end());
        mesh_root[m] = node_off;
    }
    g->mesh_height = mh;
    /* deepest stack: pseudo root + scene levels + resume (2) + marker +
     * mesh pseudo root + mesh levels */
    if ((int)(5 + g->scene_height + g->mesh_height) > STACK)
        return fail(c, SRAY_ERR_LIMIT, "octree heights %u + %u exceed the traversal stack (%d entries)",
                    g->scene_height, g->mesh_height, STACK);
    if (ft.items.size() >= (1u << 30)) return fail(c, SRAY_ERR_LIMIT, "more than 2^30 octree item references");
    if (ft.lo.size() / 4 >= (1u << 23)) return fail(c, SRAY_ERR_LIMIT, "more than 2^23 octree nodes");
    /* objects */
    uint32_t nmat = s->num_materials;
    std::vector<DevObj> objs(s->num_objects);
    for (uint32_t i = 0; i < s->num_objects; i++) {
        const sray_object_desc &o = s->objects[i];
        DevObj &d = objs[i];
        memset(&d, 0, sizeof d);
        memcpy(d.inv, o.inv_xform, sizeof d.inv);
        memcpy(d.xf, o.xform, sizeof d.xf);
        d.rad = o.radius;
        d.start = o.start;
        d.end = o.end;
        d.type = o.type;
        d.mat = o.material >= 0 ? o.material : (int32_t)nmat; /* default material appended */
        d.moving = ((o.motion[0] != 0 || o.motion[1] != 0 || o.motion[2] != 0) && o.motion_t1 > o.motion_t0) ? 1 : 0;
        memcpy(d.motion, o.motion, sizeof d.motion);
        d.mt0 = o.motion_t0;
        d.mt1 = o.motion_t1;
        d.mesh = o.mesh;
        if (o.type == SRAY_OBJ_MESH) {
            d.root = mesh_root[o.mesh];
            d.tri_base = s->meshes[o.mesh].tri_start;
            d.has_attr = (s->tri_normal ? 1 : 0) | (s->tri_texcoord ? 2 : 0) | (s->tri_tangent ? 4 : 0);
            if (d.has_attr) g->attr_meshes = true;
        }
    }
    std::vector<DevMat> mats(nmat + 1);
    for (uint32_t i = 0; i <= nmat; i++) {
        sray_material_desc md;
        if (i < nmat) md = s->materials[i];
        else {
            /* Scene::default_mat = Material(): shade_undef, src/material.cc:79-89 */
            memset(&md, 0, sizeof md);
            md.diffuse[0] = md.diffuse[1] = md.diffuse[2] = md.diffuse[3] = 1.0f;
            md.specular[3] = 1.0f;
            md.shininess = 1.0f;
            md.ior = 1.5f;
            md.flags = SRAY_MAT_HAS_SPECULAR | SRAY_MAT_HAS_REFLECT | SRAY_MAT_HAS_REFRACT;
        }
        if (md.flags & SRAY_MAT_PHONG) {
            g->has_phong = true;
            if ((md.flags & SRAY_MAT_HAS_REFLECT) && md.reflect > 0.0f) g->pos_reflect = true;
            if ((md.flags & SRAY_MAT_HAS_REFRACT) && md.refract > 0.0f) g->pos_refract = true;
        }
        DevMat &d = mats[i];
        memset(&d, 0, sizeof d);
        for (int k = 0; k < SRAY_NUM_MAPS; k++) {
            d.map[k] = md.map[k];
            if (md.map[k] > 0) d.flags |= MAT_MAPPED;
        }
        d.nscale = md.normal;
        /* mapped materials need the shading instances with surface attributes */
        if (d.flags & MAT_MAPPED) g->attr_meshes = true;
        d.kd = make_float4(md.diffuse[0], md.diffuse[1], md.diffuse[2], md.diffuse[3]);
        d.ks = make_float4(md.specular[0], md.specular[1], md.specular[2], md.specular[3]);
        d.shininess = md.shininess;
        d.reflect = md.reflect;
        d.refract = md.refract;
        d.ior = md.ior;
        d.flags |= md.flags;
    }
    std::vector<DevLight> lights(s->num_lights);
    for (uint32_t i = 0; i < s->num_lights; i++) {
        const sray_light_desc &l = s->lights[i];
        DevLight &d = lights[i];
        memset(&d, 0, sizeof d);
        d.type = l.type;
        memcpy(d.pos, l.pos, sizeof d.pos);
        memcpy(d.color, l.color, sizeof d.color);
        d.radius = l.radius;
        memcpy(d.dim, l.dim, sizeof d.dim);
        memcpy(d.xf, l.xform, sizeof d.xf);
    }
    /* triangles as 3 float4 */
    /* three float4 per triangle: the vertices, and in their w components the
     * unnormalised face normal (v1-v0) x (v2-v0), computed here with the
     * float operations the intersection test would perform */
    std::vector<float4> tri((size_t)s->num_triangles * 3);
    for (size_t t = 0; t < s->num_triangles; t++) {
        const float *v = s->tri_pos + t * 9;
        float e1[3] = {v[3] - v[0], v[4] - v[1], v[5] - v[2]}, e2[3] = {v[6] - v[0], v[7] - v[1], v[8] - v[2]};
        float n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
        for (int k = 0; k < 3; k++) tri[t * 3 + k] = make_float4(v[k * 3], v[k * 3 + 1], v[k * 3 + 2], n[k]);
    }
    /* mesh leaves' triangles copied in item order (scene-level item
     * positions hold zeros and are never read as triangles) */
    std::vector<float4> ltri(ft.items.size() * 3, make_float4(0, 0, 0, 0));
    for (size_t k = scene_items; k < ft.items.size(); k++)
        for (int v = 0; v < 3; v++) ltri[k * 3 + v] = tri[(size_t)ft.items[k] * 3 + v];
    std::vector<uint2> lbox;
    leaf_item_boxes(ft, scene_items, tri, lbox);
    /* scene item records, filled in below once the objects' keyframe
     * tables are known */
    std::vector<SceneItem> sitem(scene_items);
    /* node records interleaved, lo then hi: one 32-byte record per node */
    std::vector<float> nodes(ft.lo.size() * 2);
    for (size_t k = 0; k < ft.lo.size() / 4; k++) {
        memcpy(&nodes[k * 8], &ft.lo[k * 4], 16);
        memcpy(&nodes[k * 8 + 4], &ft.hi[k * 4], 16);
    }
    /* the host vectors above are the sources of asynchronous copies: on
     * every way out (errors included) the stream drains before they go */
    struct Drain {
        sray_gpu_ctx *c;
        ~Drain() { (void)hipStreamSynchronize(c->stream); }
    } drain{c};
    SceneDev &D = g->dev;
    memset(&D, 0, sizeof D);
    const void *p;
    /* keyframed rows (objects, lights, camera) in one buffer, each table
     * 16-byte aligned: per msec of [frame_t0, frame_t1] */
    {
        const int nrow = s->frame_t1 - s->frame_t0 + 1;
        std::vector<float> tab;
        std::vector<size_t> oofs(s->num_objects, SIZE_MAX), lofs(s->num_lights, SIZE_MAX);
        size_t cofs = SIZE_MAX;
        auto add = [&](const float *src, size_t n) {
            size_t at = (tab.size() + 3) & ~(size_t)3;
            tab.resize(at + n);
            memcpy(&tab[at], src, n * sizeof(float));
            return at;
        };
        for (uint32_t i = 0; i < s->num_objects; i++)
            if (s->objects[i].time_xforms) oofs[i] = add(s->objects[i].time_xforms, (size_t)nrow * 24);
        for (uint32_t i = 0; i < s->num_lights; i++)
            if (s->lights[i].time_pos && s->lights[i].type != SRAY_LIGHT_BOX)
                lofs[i] = add(s->lights[i].time_pos, (size_t)nrow * 3);
        if (s->camera.time_xforms) cofs = add(s->camera.time_xforms, (size_t)nrow * 12);
        if (!tab.empty()) {
            if ((rc = upload(c, *g, tab.data(), tab.size() * sizeof(float), &p))) return rc;
            const float *base = (const float *)p;
            for (uint32_t i = 0; i < s->num_objects; i++)
                if (oofs[i] != SIZE_MAX) {
                    objs[i].moving = 2;
                    objs[i].tab = base + oofs[i];
                    objs[i].tab_t0 = s->frame_t0;
                    objs[i].tab_n = nrow;
                }
            for (uint32_t i = 0; i < s->num_lights; i++)
                if (lofs[i] != SIZE_MAX) {
                    lights[i].tab = base + lofs[i];
                    lights[i].tab_t0 = s->frame_t0;
                    lights[i].tab_n = nrow;
                }
            if (cofs != SIZE_MAX) {
                D.cam_tab = base + cofs;
                D.cam_t0 = s->frame_t0;
                D.cam_n = nrow;
            }
            g->timed = true;
        }
        g->frame_t0 = s->frame_t0;
        g->frame_t1 = s->frame_t1;
    }
    if ((rc = upload(c, *g, nodes.data(), nodes.size() * 4, &p))) return rc;
    D.node = (const float4 *)p;
    for (size_t k = 0; k < scene_items; k++) {
        const DevObj &o = objs[ft.items[k]];
        SceneItem &d = sitem[k];
        memset(&d, 0, sizeof d);
        memcpy(d.inv, o.inv, sizeof d.inv);
        d.rad = o.rad;
        d.start = o.start;
        d.end = o.end;
        d.type = o.type;
        d.obj = (int32_t)ft.items[k];
        d.root = o.root;
        d.moving = o.moving;
    }
    if ((rc = upload(c, *g, sitem.data(), sitem.size() * sizeof(SceneItem), &p))) return rc;
    D.sitem = (const SceneItem *)p;
    if ((rc = upload(c, *g, ft.items.data(), ft.items.size() * 4, &p))) return rc;
    D.items = (const uint32_t *)p;
    if ((rc = upload(c, *g, objs.data(), objs.size() * sizeof(DevObj), &p))) return rc;
    D.objs = (const DevObj *)p;
    if ((rc = upload(c, *g, tri.data(), tri.size() * sizeof(float4), &p))) return rc;
    D.tri = (const float4 *)p;
    if ((rc = upload(c, *g, ltri.data(), ltri.size() * sizeof(float4), &p))) return rc;
    D.ltri = (const float4 *)p;
    if ((rc = upload(c, *g, lbox.data(), lbox.size() * sizeof(uint2), &p))) return rc;
    D.lbox = (const uint2 *)p;
    if (s->tri_normal) {
        if ((rc = upload(c, *g, s->tri_normal, (size_t)s->num_triangles * 9 * 4, &p))) return rc;
        D.tri_n = (const float *)p;
    }
    if (s->tri_tangent) {
        if ((rc = upload(c, *g, s->tri_tangent, (size_t)s->num_triangles * 9 * 4, &p))) return rc;
        D.tri_g = (const float *)p;
    }
    if (s->tri_texcoord) {
        if ((rc = upload(c, *g, s->tri_texcoord, (size_t)s->num_triangles * 6 * 4, &p))) return rc;
        D.tri_uv = (const float *)p;
    }
    if ((rc = upload(c, *g, mats.data(), mats.size() * sizeof(DevMat), &p))) return rc;
    D.mats = (const DevMat *)p;
    /* textures: texels, then the table */
    if (s->num_textures) {
        std::vector<DevTex> texs(s->num_textures);
        for (uint32_t i = 0; i < s->num_textures; i++) {
            const sray_texture_desc &td = s->textures[i];
            DevTex &t = texs[i];
            memset(&t, 0, sizeof t);
            t.w = td.width;
            t.h = td.height;
            t.filter = td.filter;
            t.wrap = td.wrap;
            memcpy(t.xf, td.xform, sizeof t.xf);
            if ((rc = upload(c, *g, td.texels, (size_t)td.width * td.height * 16, &p))) return rc;
            t.texels = (const float4 *)p;
        }
        if ((rc = upload(c, *g, texs.data(), texs.size() * sizeof(DevTex), &p))) return rc;
        D.texs = (const DevTex *)p;
        D.ntex = (int32_t)s->num_textures;
    }
    if ((rc = upload(c, *g, lights.data(), lights.size() * sizeof(DevLight), &p))) return rc;
    D.lights = (const DevLight *)p;
    D.scene_root = sroot;
    D.nlights = (int32_t)s->num_lights;
    D.nobjs = (int32_t)s->num_objects;
    D.fast_boxes = 1;
    for (size_t k = 0; k < ft.lo.size() / 4 && D.fast_boxes; k++)
        for (int a = 0; a < 3; a++) {
            const float l = ft.lo[k * 4 + a], h = ft.hi[k * 4 + a];
            if (!(std::isfinite(l) && std::isfinite(h) && l <= h && std::fabs(l) <= 0x1p120f && std::fabs(h) <= 0x1p120f))
                D.fast_boxes = 0;
        }
    D.env_ambient = make_float4(s->env_ambient[0], s->env_ambient[1], s->env_ambient[2], s->env_ambient[3]);
    D.env_color = make_float4(s->env_color[0], s->env_color[1], s->env_color[2], s->env_color[3]);
    memcpy(D.cam, s->camera.xform, sizeof D.cam);
    D.fz = (float)(1.0 / std::tan(0.5 * s->camera.vfov));
    D.max_dist = s->camera.max_dist;
    D.shutter = s->camera.shutter;
    g->nobjs = s->num_objects;
    sray::photon_scene_from_desc(s, g->photon);
    g->camera = s->camera;
    g->camera.time_xforms = nullptr; /* the caller's rows: only D.cam_tab is kept */
    HIPCHK(c, hipStreamSynchronize(c->stream));
    uint32_t id = c->next_scene++;
    c->scenes[id] = std::move(g);
    *scene_id = id;
    return SRAY_OK;
}

/* per-level open segments of the balanced tree over n photons (host) */
void pm_levels(uint32_t n, std::vector<std::vector<uint32_t>> &lo, std::vector<std::vector<uint32_t>> &hi)
{
    lo.clear();
    hi.clear();
    std::vector<uint32_t> cl = {0}, ch = {n};
    while (!cl.empty()) {
        std::vector<uint32_t> l2, h2, nl, nh;
        for (size_t k = 0; k < cl.size(); k++) {
            uint32_t a = cl[k], b = ch[k];
            if (b - a >= 2) {
                l2.push_back(a);
                h2.push_back(b);
            }
            uint32_t m = (a + b) >> 1;
            if (m > a + 1) {  /* left child with >= 2 elements still needs sorting */
                nl.push_back(a);
                nh.push_back(m);
            }
            if (b > m + 2) {
                nl.push_back(m + 1);
                nh.push_back(b);
            }
        }
        if (!l2.empty()) {
            lo.push_back(l2);
            hi.push_back(h2);
        }
        cl.swap(nl);
        ch.swap(nh);
    }
}

/* photon map from n photons in storage order (host or device memory): the
 * array is kept as given (for download) and a kd-tree copy is built */
int do_pmap_build(sray_gpu_ctx *c, uint32_t scene_id, const sray_photon *ph, bool on_device, uint32_t n, int kind,
                  uint32_t *pmap_id)
{
    if (kind != SRAY_PMAP_CAUSTIC && kind != SRAY_PMAP_GLOBAL) return fail(c, SRAY_ERR_ARG, "bad photon map kind");
    if (scene_id && !find_scene(c, scene_id)) return fail(c, SRAY_ERR_NOTFOUND, "unknown scene %u", scene_id);
    if (n && !ph) return fail(c, SRAY_ERR_ARG, "photons is NULL");
    if (n >= (1u << 30)) return fail(c, SRAY_ERR_LIMIT, "more than 2^30 photons");
    auto g = std::make_unique<GpuPmap>();
    g->scene_id = scene_id;
    g->kind = kind;
    int rc;
    auto alloc = [&](size_t bytes, void **p) -> int {
        DevBuf b;
        hipError_t e = hipMalloc(&b.p, bytes ? bytes : 16);
        if (e != hipSuccess) return fail(c, SRAY_ERR_NOMEM, "hipMalloc(%zu) failed", bytes);
        b.bytes = bytes;
        g->bufs.push_back(b);
        g->bytes += bytes;
        *p = b.p;
        return SRAY_OK;
    };
    void *pos, *dir, *nrm, *colb, *axis, *store;
    if ((rc = alloc((size_t)n * sizeof(sray_photon), &store)) || (rc = alloc((size_t)n * 16, &pos)) || (rc = alloc((size_t)n * 16, &dir)) || (rc = alloc((size_t)n * 16, &nrm)) ||
        (rc = alloc((size_t)n * 4, &colb)) || (rc = alloc((size_t)n, &axis)))
        return rc;
    uint32_t depth = 0;
    if (n) {
        /* scratch for the build */
        DevBuf pos0, ord[2], key[2], seg, bb, slo, shi, sax;
        sray_photon *dph = (sray_photon *)store;
        auto cleanup = [&]() {
            DevBuf *bs[] = {&pos0, &ord[0], &ord[1], &key[0], &key[1], &seg, &bb, &slo, &shi, &sax};
            for (DevBuf *b : bs) release(*b);
        };
        std::vector<std::vector<uint32_t>> L, H;
        pm_levels(n, L, H);
        depth = (uint32_t)L.size() + 1;
        size_t maxseg = 1;
        for (auto &v : L) maxseg = std::max(maxseg, v.size());
        if ((rc = ensure(c, pos0, (size_t)n * 16)) ||
            (rc = ensure(c, ord[0], (size_t)n * 4)) || (rc = ensure(c, ord[1], (size_t)n * 4)) ||
            (rc = ensure(c, key[0], (size_t)n * 8)) || (rc = ensure(c, key[1], (size_t)n * 8)) ||
            (rc = ensure(c, seg, (size_t)n * 4)) || (rc = ensure(c, bb, maxseg * 24)) ||
            (rc = ensure(c, slo, maxseg * 4)) || (rc = ensure(c, shi, maxseg * 4)) || (rc = ensure(c, sax, maxseg))) {
            cleanup();
            return rc;
        }
        hipError_t e = hipMemcpyAsync(dph, ph, (size_t)n * sizeof(sray_photon),
                                      on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, c->stream);
        if (e != hipSuccess) {
            cleanup();
            return fail(c, SRAY_ERR_HIP, "photon upload failed");
        }
        (void)hipMemsetAsync(axis, 0, n, c->stream);
        k_pm_init<<<blocks(n, 256), 256, 0, c->stream>>>(dph, n, dptr<float4>(pos0),
                                                          dptr<uint32_t>(ord[0]));
        int endbit = 32;
        while ((1ull << (endbit - 32)) <= n) endbit++;
        int cur = 0;
        for (size_t l = 0; l < L.size(); l++) {
            uint32_t ns = (uint32_t)L[l].size();
            (void)hipMemcpyAsync(slo.p, L[l].data(), ns * 4, hipMemcpyHostToDevice, c->stream);
            (void)hipMemcpyAsync(shi.p, H[l].data(), ns * 4, hipMemcpyHostToDevice, c->stream);
            k_pm_seg<<<blocks(n, 256), 256, 0, c->stream>>>(dptr<uint32_t>(slo), dptr<uint32_t>(shi), ns, n,
                                                             dptr<int32_t>(seg));
            k_pm_bbinit<<<blocks((size_t)ns * 6, 256), 256, 0, c->stream>>>(dptr<uint32_t>(bb), ns);
            k_pm_bbox<<<blocks(n, 256), 256, 0, c->stream>>>(dptr<uint32_t>(ord[cur]), dptr<float4>(pos0),
                                                              dptr<int32_t>(seg), n, dptr<uint32_t>(bb));
            k_pm_axis<<<blocks(ns, 256), 256, 0, c->stream>>>(dptr<uint32_t>(bb), ns, dptr<uint8_t>(sax));
            k_pm_keys<<<blocks(n, 256), 256, 0, c->stream>>>(dptr<uint32_t>(ord[cur]), dptr<float4>(pos0),
                                                              dptr<int32_t>(seg), dptr<uint32_t>(slo),
                                                              dptr<uint8_t>(sax), n,
                                                              dptr<unsigned long long>(key[0]));
            if ((rc = radix_sort_pairs(c, dptr<unsigned long long>(key[0]), dptr<unsigned long long>(key[1]),
                                       dptr<uint32_t>(ord[cur]), dptr<uint32_t>(ord[cur ^ 1]), n, endbit))) {
                cleanup();
                return rc;
            }
            cur ^= 1;
            k_pm_nodes<<<blocks(ns, 256), 256, 0, c->stream>>>(dptr<uint32_t>(slo), dptr<uint32_t>(shi),
                                                                dptr<uint8_t>(sax), ns, (uint8_t *)axis);
        }
        k_pm_final<<<blocks(n, 256), 256, 0, c->stream>>>(dph, dptr<uint32_t>(ord[cur]), n,
                                                           (float4 *)pos, (float4 *)dir, (float4 *)nrm, (float *)colb);
        e = hipStreamSynchronize(c->stream);
        cleanup();
        if (e != hipSuccess) return fail(c, SRAY_ERR_HIP, "kd-tree build failed: %s", hipGetErrorString(e));
        if (depth + 1 > (uint32_t)PMAP_STACK) return fail(c, SRAY_ERR_LIMIT, "kd-tree deeper than the gather stack");
    }
    PmapTree &T = g->tree;
    T.pos = (const float4 *)pos;
    T.dir = (const float4 *)dir;
    T.nrm = (const float4 *)nrm;
    T.colb = (const float *)colb;
    T.axis = (const uint8_t *)axis;
    T.n = n;
    T.depth = depth;
    g->store = (sray_photon *)store;
    uint32_t id = c->next_pmap++;
    c->pmaps[id] = std::move(g);
    if (scene_id) find_scene(c, scene_id)->pmap[kind] = id; /* the scene's caustic/global map */
    *pmap_id = id;
    return SRAY_OK;
}

int do_pmap_gather(sray_gpu_ctx *c, uint32_t pmap_id, const sray_gather_query *q, uint32_t n, float radius, int k,
                   int mode, float *rgb)
{
    auto it = c->pmaps.find(pmap_id);
    if (it == c->pmaps.end()) return fail(c, SRAY_ERR_NOTFOUND, "unknown photon map %u", pmap_id);
    bool grid = (mode & SRAY_GATHER_GRID) != 0;
    mode &= ~SRAY_GATHER_GRID;
    if (mode != SRAY_GATHER_IRRADIANCE && mode != SRAY_GATHER_RADIANCE) return fail(c, SRAY_ERR_ARG, "bad gather mode");
    if (k < 0) return fail(c, SRAY_ERR_ARG, "k < 0");
    if (!(radius > 0.0f)) return fail(c, SRAY_ERR_ARG, "radius must be > 0");
    if (n == 0) return SRAY_OK;
    const PmapTree &T = it->second->tree;
    int rc;
    if (grid) {
        PmapGrid G;
        if ((rc = pmap_grid_get(c, pmap_id, radius, c->tune.gather_cell, &G)) ||
            (rc = ensure(c, c->gq, (size_t)n * sizeof(sray_gather_query))) || (rc = ensure(c, c->grgb, (size_t)n * 12)))
            return rc;
        HIPCHK(c, hipMemcpyAsync(c->gq.p, q, (size_t)n * sizeof(sray_gather_query), hipMemcpyHostToDevice, c->stream));
        {
            ProfScope ps(c, "k_pm_gather_grid");
            k_pm_gather_grid<<<blocks((size_t)n * 16, 256), 256, 0, c->stream>>>(
                G, dptr<sray_gather_query>(c->gq), n, radius, mode, k, dptr<float>(c->grgb));
        }
        HIPCHK(c, hipGetLastError());
        HIPCHK(c, hipMemcpyAsync(rgb, c->grgb.p, (size_t)n * 12, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        prof_resolve(c);
        return SRAY_OK;
    }
    if ((rc = ensure(c, c->gq, (size_t)n * sizeof(sray_gather_query))) || (rc = ensure(c, c->grgb, (size_t)n * 12)) ||
        (rc = ensure(c, c->gcnt, (size_t)n * 4)) || (rc = ensure(c, c->gneed, n)) ||
        (rc = ensure(c, c->giota, (size_t)n * 4)) || (rc = ensure(c, c->gfq, (size_t)n * 4)) ||
        (rc = ensure(c, c->goff, (size_t)n * 4 + 4)) || (rc = ensure(c, c->nsel, 4)))
        return rc;
    HIPCHK(c, hipMemcpyAsync(c->gq.p, q, (size_t)n * sizeof(sray_gather_query), hipMemcpyHostToDevice, c->stream));
    {
        ProfScope ps(c, "k_pm_gather");
        k_pm_gather<<<blocks(n, 128), 128, 0, c->stream>>>(T, dptr<sray_gather_query>(c->gq), n, radius, mode, k,
                                                          dptr<float>(c->grgb), dptr<uint32_t>(c->gcnt),
                                                          dptr<uint8_t>(c->gneed));
    }
    if (k > 0) {
        k_iota<<<blocks(n, 256), 256, 0, c->stream>>>(dptr<uint32_t>(c->giota), n);
        uint32_t nf = 0;
        if ((rc = select_flagged(c, dptr<uint32_t>(c->giota), dptr<uint8_t>(c->gneed), dptr<uint32_t>(c->gfq), n, &nf)))
            return rc;
        if (nf) {
            /* candidate offsets: exclusive scan of the flagged queries' counts */
            std::vector<uint32_t> fq(nf), cnt(n);
            HIPCHK(c, hipMemcpy(fq.data(), c->gfq.p, (size_t)nf * 4, hipMemcpyDeviceToHost));
            HIPCHK(c, hipMemcpy(cnt.data(), c->gcnt.p, (size_t)n * 4, hipMemcpyDeviceToHost));
            std::vector<uint32_t> off(nf);
            uint64_t tot = 0;
            for (uint32_t j = 0; j < nf; j++) {
                off[j] = (uint32_t)tot;
                tot += cnt[fq[j]];
            }
            if (tot >= (1ull << 32)) return fail(c, SRAY_ERR_LIMIT, "too many k-NN candidates");
            if ((rc = ensure(c, c->gcand, (size_t)tot * sizeof(Cand)))) return rc;
            HIPCHK(c, hipMemcpyAsync(c->goff.p, off.data(), (size_t)nf * 4, hipMemcpyHostToDevice, c->stream));
            k_pm_collect<<<blocks(nf, 128), 128, 0, c->stream>>>(T, dptr<sray_gather_query>(c->gq),
                                                                dptr<uint32_t>(c->gfq), nf, dptr<uint32_t>(c->goff),
                                                                radius, dptr<Cand>(c->gcand));
            k_pm_select<<<blocks(nf, 64), 64, 0, c->stream>>>(T, dptr<sray_gather_query>(c->gq), dptr<uint32_t>(c->gfq),
                                                              nf, dptr<uint32_t>(c->goff), dptr<uint32_t>(c->gcnt),
                                                              dptr<Cand>(c->gcand), k, radius, mode,
                                                              dptr<float>(c->grgb));
        }
    }
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(rgb, c->grgb.p, (size_t)n * 12, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    prof_resolve(c);
    return SRAY_OK;
}

/* shoot_photons (src/render.cc:67-122): projection maps on the host, one
 * device lane per emitted photon, stable compaction in emission order, then
 * the kd-tree build; the maps are bound to the scene for its renders */
int do_shoot(sray_gpu_ctx *c, uint32_t scene_id, const sray_photon_params *pp, uint32_t *ids, sray_photon_stats *st)
{
    GpuScene *g = find_scene(c, scene_id);
    if (!g) return fail(c, SRAY_ERR_NOTFOUND, "unknown scene %u", scene_id);
    if (pp->caust_photons < 0 || pp->gi_photons < 0) return fail(c, SRAY_ERR_ARG, "negative photon count");
    if (pp->t1 < pp->t0) return fail(c, SRAY_ERR_ARG, "photon interval t1 < t0");
    if (g->timed && (pp->t0 < g->frame_t0 || pp->t1 > g->frame_t1))
        return fail(c, SRAY_ERR_ARG, "photon interval %d..%d msec leaves the scene's keyframed interval %d..%d",
                    pp->t0, pp->t1, g->frame_t0, g->frame_t1);
    const sray::PhotonScene &ps = g->photon;
    size_t nl = ps.lights.size();
    if (st) memset(st, 0, sizeof *st);
    ids[0] = ids[1] = 0;
    if (nl == 0 || (pp->caust_photons == 0 && pp->gi_photons == 0)) return SRAY_OK;
    std::vector<sray::ProjMap> all(nl), spec(nl);
    for (size_t li = 0; li < nl; li++) sray::build_projmaps(ps, (int)li, pp->t0, pp->t1, pp->seed, all[li], spec[li]);
    PhotonJob J;
    memset(&J, 0, sizeof J);
    sray::cell_tables(J.tab.cv, J.tab.cu);
    J.nlights = (int32_t)nl;
    J.t0 = pp->t0;
    J.t1 = pp->t1;
    J.iter = pp->iter;
    J.seed = pp->seed;
    int rc;
    DevBuf dl, dc, dout, dkeep, dsel, dseg;
    auto cleanup = [&]() {
        DevBuf *bs[] = {&dl, &dc, &dout, &dkeep, &dsel, &dseg};
        for (DevBuf *b : bs) release(*b);
    };
    if ((rc = ensure(c, dl, nl * sizeof(PhotonLightDev))) || (rc = ensure(c, dc, nl * SRAY_PROJ_U * SRAY_PROJ_V)) ||
        (rc = ensure(c, dseg, sizeof(unsigned long long))) || (rc = ensure(c, c->nsel, 4))) {
        cleanup();
        return rc;
    }
    hipEvent_t e0 = c->ev0, e1 = c->ev1;
    double trace_ms = 0, build_ms = 0;
    for (int kind = 0; kind < 2; kind++) {
        int num = kind == SRAY_PMAP_CAUSTIC ? pp->caust_photons : pp->gi_photons;
        if (num <= 0) continue;
        std::vector<uint32_t> nphot;
        sray::photon_counts(ps, num, nphot);
        std::vector<PhotonLightDev> L(nl);
        std::vector<uint8_t> cells(nl * SRAY_PROJ_U * SRAY_PROJ_V);
        uint64_t total = 0;
        for (size_t li = 0; li < nl; li++) {
            /* caustic photons aim at the specular objects' map, global ones at all objects */
            const sray::ProjMap &m = kind == SRAY_PMAP_CAUSTIC ? spec[li] : all[li];
            bool use = m.coverage != 0.0;
            double e = (double)pp->photon_energy * (use ? m.coverage : 1.0);
            for (int k = 0; k < 3; k++) L[li].col0[k] = (float)((double)ps.lights[li].color[k] * e);
            L[li].use_map = use ? 1 : 0;
            L[li].first = (uint32_t)total;
            L[li].count = nphot[li];
            L[li].scale = nphot[li] ? 1.0 / (double)nphot[li] : 0.0;
            memcpy(&cells[li * SRAY_PROJ_U * SRAY_PROJ_V], m.cells, SRAY_PROJ_U * SRAY_PROJ_V);
            total += nphot[li];
        }
        if (total >= (1ull << 30)) {
            cleanup();
            return fail(c, SRAY_ERR_LIMIT, "more than 2^30 photons");
        }
        if (st) st->shot[kind] = total;
        if (total == 0) continue;
        uint32_t n = (uint32_t)total;
        if ((rc = ensure(c, dout, (size_t)n * sizeof(sray_photon))) || (rc = ensure(c, dkeep, n)) ||
            (rc = ensure(c, dsel, (size_t)n * sizeof(sray_photon)))) {
            cleanup();
            return rc;
        }
        HIPCHK(c, hipMemcpyAsync(dl.p, L.data(), nl * sizeof(PhotonLightDev), hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(dc.p, cells.data(), cells.size(), hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipMemsetAsync(dseg.p, 0, sizeof(unsigned long long), c->stream));
        J.lights = dptr<PhotonLightDev>(dl);
        J.cells = dptr<uint8_t>(dc);
        J.kind = kind;
        J.total = n;
        HIPCHK(c, hipEventRecord(e0, c->stream));
        {
            ProfScope ps_(c, "k_photon");
            k_photon<true><<<blocks(n, TRACE_BLOCK), TRACE_BLOCK, 0, c->stream>>>(
                g->dev, J, dptr<sray_photon>(dout), dptr<uint8_t>(dkeep), dptr<unsigned long long>(dseg));
        }
        HIPCHK(c, hipGetLastError());
        uint32_t stored = 0;
        if ((rc = select_flagged(c, dptr<sray_photon>(dout), dptr<uint8_t>(dkeep), dptr<sray_photon>(dsel), n,
                                 &stored))) {
            cleanup();
            return rc;
        }
        HIPCHK(c, hipEventRecord(e1, c->stream));
        HIPCHK(c, hipEventSynchronize(e1));
        float ms = 0;
        HIPCHK(c, hipEventElapsedTime(&ms, e0, e1));
        trace_ms += ms;
        unsigned long long segs = 0;
        HIPCHK(c, hipMemcpy(&segs, dseg.p, sizeof segs, hipMemcpyDeviceToHost));
        if (st) {
            st->stored[kind] = stored;
            st->segments += segs;
        }
        /* the previous map of this kind stays alive until released by its id */
        auto t0 = std::chrono::steady_clock::now();
        if ((rc = do_pmap_build(c, scene_id, dptr<sray_photon>(dsel), true, stored, kind, &ids[kind]))) {
            cleanup();
            return rc;
        }
        build_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    cleanup();
    prof_resolve(c);
    if (st) {
        st->trace_ms = trace_ms;
        st->build_ms = build_ms;
    }
    return SRAY_OK;
}

int do_pmap_download(sray_gpu_ctx *c, uint32_t pmap_id, sray_photon *out, uint32_t cap, uint32_t *n)
{
    auto it = c->pmaps.find(pmap_id);
    if (it == c->pmaps.end()) return fail(c, SRAY_ERR_NOTFOUND, "unknown photon map %u", pmap_id);
    uint32_t cnt = it->second->tree.n;
    *n = cnt;
    if (!out || cnt == 0) return SRAY_OK;
    if (cap < cnt) return fail(c, SRAY_ERR_ARG, "photon buffer holds %u, map has %u", cap, cnt);
    HIPCHK(c, hipMemcpyAsync(out, it->second->store, (size_t)cnt * sizeof(sray_photon), hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return SRAY_OK;
}

void free_pmap(GpuPmap &m)
{
    for (auto &b : m.bufs) release(b);
    for (auto &g : m.grids) {
        DevBuf *bs[] = {&g.second->pos, &g.second->dir, &g.second->nrm, &g.second->colb, &g.second->cell};
        for (DevBuf *d : bs) release(*d);
    }
    m.grids.clear();
}

/* the shading-gather grid of a photon map for one radius: cells of the
 * radius (larger if the photons' box would need more than 2^24 cells),
 * photons sorted by cell with a device radix sort */
int pmap_grid_get(sray_gpu_ctx *c, uint32_t pmap_id, float radius, float cellr, PmapGrid *out)
{
    auto it = c->pmaps.find(pmap_id);
    if (it == c->pmaps.end()) return fail(c, SRAY_ERR_NOTFOUND, "unknown photon map %u", pmap_id);
    GpuPmap &pm = *it->second;
    if (!(cellr > 0.0f)) cellr = 1.0f;
    uint32_t kr, kc;
    memcpy(&kr, &radius, 4);
    memcpy(&kc, &cellr, 4);
    uint64_t key = ((uint64_t)kr << 32) | kc;
    auto g = pm.grids.find(key);
    if (g != pm.grids.end()) {
        *out = g->second->g;
        return SRAY_OK;
    }
    if (!(radius > 0.0f)) return fail(c, SRAY_ERR_ARG, "gather radius must be > 0");
    auto G = std::make_unique<GpuPmap::Grid>();
    const PmapTree &T = pm.tree;
    uint32_t n = T.n;
    PmapGrid &pg = G->g;
    memset(&pg, 0, sizeof pg);
    int rc;
    if (n == 0) {
        if ((rc = ensure(c, G->cell, 8))) return rc;
        HIPCHK(c, hipMemsetAsync(G->cell.p, 0, 8, c->stream));
        pg.cell_start = dptr<uint32_t>(G->cell);
        pg.dx = pg.dy = pg.dz = 1;
        pg.inv_cs = 1.0f;
        *out = pg;
        pm.grids[key] = std::move(G);
        return SRAY_OK;
    }
    DevBuf bb, keys[2], vals[2];
    auto cleanup = [&]() {
        DevBuf *bs[] = {&bb, &keys[0], &keys[1], &vals[0], &vals[1]};
        for (DevBuf *b : bs) release(*b);
    };
    if ((rc = ensure(c, bb, 24))) return rc;
    k_pm_bbinit<<<1, 64, 0, c->stream>>>(dptr<uint32_t>(bb), 1);
    k_grid_bbox<<<blocks(n, 256), 256, 0, c->stream>>>(T.pos, n, dptr<uint32_t>(bb));
    uint32_t hb[6];
    if (hipMemcpyAsync(hb, bb.p, 24, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess) {
        cleanup();
        return fail(c, SRAY_ERR_HIP, "photon bounds failed");
    }
    float lo[3], hi[3];
    for (int k = 0; k < 3; k++) {
        lo[k] = host_unkey(hb[k]);
        hi[k] = host_unkey(hb[3 + k]);
    }
    /* cell size: cellr radii, or larger to keep the cell count bounded */
    const double MAXC = (double)(1u << 24);
    double cs = (double)radius * (double)cellr;
    auto cells = [&](double s, int64_t d[3]) {
        double t = 1;
        for (int k = 0; k < 3; k++) {
            d[k] = (int64_t)std::floor(((double)hi[k] - (double)lo[k]) / s) + 1;
            t *= (double)d[k];
        }
        return t;
    };
    int64_t dd[3];
    while (cells(cs, dd) > MAXC) cs *= 1.25;
    pg.ox = lo[0];
    pg.oy = lo[1];
    pg.oz = lo[2];
    pg.inv_cs = (float)(1.0 / cs);
    pg.dx = (int32_t)dd[0];
    pg.dy = (int32_t)dd[1];
    pg.dz = (int32_t)dd[2];
    uint32_t ncell = (uint32_t)(dd[0] * dd[1] * dd[2]);
    if ((rc = ensure(c, keys[0], (size_t)n * 4)) || (rc = ensure(c, keys[1], (size_t)n * 4)) ||
        (rc = ensure(c, vals[0], (size_t)n * 4)) || (rc = ensure(c, vals[1], (size_t)n * 4)) ||
        (rc = ensure(c, G->cell, ((size_t)ncell + 1) * 4)) || (rc = ensure(c, G->pos, (size_t)n * 16)) ||
        (rc = ensure(c, G->dir, (size_t)n * 16)) || (rc = ensure(c, G->nrm, (size_t)n * 16)) ||
        (rc = ensure(c, G->colb, (size_t)n * 4))) {
        cleanup();
        return rc;
    }
    k_grid_key<<<blocks(n, 256), 256, 0, c->stream>>>(T.pos, n, pg, dptr<uint32_t>(keys[0]), dptr<uint32_t>(vals[0]));
    int endbit = 1;
    while (endbit < 32 && (1ull << endbit) <= ncell) endbit++;
    if ((rc = radix_sort_pairs(c, dptr<uint32_t>(keys[0]), dptr<uint32_t>(keys[1]), dptr<uint32_t>(vals[0]),
                               dptr<uint32_t>(vals[1]), n, endbit))) {
        cleanup();
        return rc;
    }
    hipError_t e;
    k_grid_start<<<blocks((size_t)ncell + 1, 256), 256, 0, c->stream>>>(dptr<uint32_t>(keys[1]), n, ncell,
                                                                        dptr<uint32_t>(G->cell));
    k_grid_final<<<blocks(n, 256), 256, 0, c->stream>>>(T, dptr<uint32_t>(vals[1]), n, dptr<float4>(G->pos),
                                                        dptr<float4>(G->dir), dptr<float4>(G->nrm),
                                                        dptr<float>(G->colb));
    e = hipStreamSynchronize(c->stream);
    cleanup();
    if (e != hipSuccess) return fail(c, SRAY_ERR_HIP, "grid build failed: %s", hipGetErrorString(e));
    pg.v.pos = dptr<float4>(G->pos);
    pg.v.dir = dptr<float4>(G->dir);
    pg.v.nrm = dptr<float4>(G->nrm);
    pg.v.colb = dptr<float>(G->colb);
    pg.v.axis = nullptr;
    pg.v.n = n;
    pg.v.depth = 0;
    pg.cell_start = dptr<uint32_t>(G->cell);
    *out = pg;
    pm.grids[key] = std::move(G);
    return SRAY_OK;
}

/* do_scene_upload with host memory exhaustion (octree build, staging)
 * reported as an error instead of escaping the C ABI */
int scene_upload_guarded(sray_gpu_ctx *c, const sray_scene_desc *s, uint32_t *scene_id)
{
    try {
        return do_scene_upload(c, s, scene_id);
    } catch (const std::bad_alloc &) {
        return fail(c, SRAY_ERR_NOMEM, "host memory exhausted while building the scene's octrees");
    }
}

int do_trace_device(sray_gpu_ctx *c, uint32_t scene_id, const sray_ray *d_rays, uint32_t n, sray_hit *d_hits,
                    uint32_t flags, unsigned long long *d_counts)
{
    GpuScene *g = find_scene(c, scene_id);
    if (!g) return fail(c, SRAY_ERR_NOTFOUND, "unknown scene %u", scene_id);
    if (n == 0) return SRAY_OK;
    bool any = flags & SRAY_TRACE_ANY, count = (flags & SRAY_TRACE_COUNT) && d_counts;
    uint32_t nb = blocks(n, TRACE_BLOCK);
    if (flags & SRAY_TRACE_WAVE) {
        if (count) return fail(c, SRAY_ERR_ARG, "SRAY_TRACE_WAVE takes no visit counters");
        ProfScope ps(c, "k_trace_rays_wave");
        const bool sp = flags & SRAY_TRACE_SPARSE;
        auto kern = any ? (sp ? k_trace_rays_wave<true, true> : k_trace_rays_wave<true, false>)
                        : (sp ? k_trace_rays_wave<false, true> : k_trace_rays_wave<false, false>);
        kern<<<nb, TRACE_BLOCK, 0, c->stream>>>(g->dev, d_rays, n, d_hits);
        HIPCHK(c, hipGetLastError());
        return SRAY_OK;
    }
    ProfScope ps(c, any ? "k_trace_rays_any" : "k_trace_rays");
    if (any && count) k_trace_rays<true, true><<<nb, TRACE_BLOCK, 0, c->stream>>>(g->dev, d_rays, n, d_hits, d_counts);
    else if (any) k_trace_rays<true, false><<<nb, TRACE_BLOCK, 0, c->stream>>>(g->dev, d_rays, n, d_hits, d_counts);
    else if (count) k_trace_rays<false, true><<<nb, TRACE_BLOCK, 0, c->stream>>>(g->dev, d_rays, n, d_hits, d_counts);
    else k_trace_rays<false, false><<<nb, TRACE_BLOCK, 0, c->stream>>>(g->dev, d_rays, n, d_hits, d_counts);
    HIPCHK(c, hipGetLastError());
    return SRAY_OK;
}

} // namespace

/* ================================================================== */
/* the gather that allocates nothing                                    */
/* ================================================================== */
/* Per-query counters (4 B count + 1 B selection flag) of
 * sray_gpu_pmap_gather_device_noalloc.  Each GPU holds its own copy, at
 * its own address, until the process ends.  In host code `scratch` names
 * a host-side stand-in, not the device array: its device address comes
 * from hipGetSymbolAddress, once per device, with that device current. */
static __device__ __align__(256) uint8_t scratch[(size_t)SRAY_NOALLOC_MAX_QUERIES * 5];

namespace {

constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
uint8_t *g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

/* the device address of `scratch` on the current device (dev), looked up
 * once; hipGetSymbolAddress allocates nothing */
int scratch_init(int dev)
{
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return SRAY_ERR_LIMIT;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[dev]) return SRAY_OK;
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return SRAY_ERR_HIP;
    g_scratch_ptr[dev] = (uint8_t *)p;
    return SRAY_OK;
}

uint8_t *scratch_ptr(int dev)
{
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return g_scratch_ptr[dev];
}

} // namespace

/* ================================================================== */
/* C ABI                                                                */
/* ================================================================== */
extern "C" {

int sray_gpu_create(int device, uint32_t flags, sray_gpu_ctx **out)
{
    (void)flags;
    if (!out) return SRAY_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return SRAY_ERR_NODEVICE;
    if (device < 0 || device >= ndev) return SRAY_ERR_ARG;
    auto *c = new sray_gpu_ctx;
    c->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess ||
        scratch_init(device) != SRAY_OK) {
        if (c->stream) (void)hipStreamDestroy(c->stream);
        if (c->ev0) (void)hipEventDestroy(c->ev0);
        if (c->ev1) (void)hipEventDestroy(c->ev1);
        delete c;
        return SRAY_ERR_HIP;
    }
    *out = c;
    return SRAY_OK;
}

void sray_gpu_destroy(sray_gpu_ctx *c)
{
    if (!c) return;
    engine_destroy(c);
    {
        ApiLock lk(c);
        (void)hipSetDevice(c->device);
        (void)hipStreamSynchronize(c->stream);
        for (auto &kv : c->scenes)
            for (auto &b : kv.second->bufs) release(b);
        c->scenes.clear();
        for (auto &kv : c->pmaps) free_pmap(*kv.second);
        c->pmaps.clear();
        DevBuf *bs[] = {&c->seltmp, &c->sorttmp, &c->nsel, &c->rays, &c->hitbuf, &c->counts, &c->gq, &c->grgb, &c->gcnt,
                        &c->gneed, &c->gfq, &c->goff, &c->gcand, &c->giota};
        for (DevBuf *b : bs) release(*b);
        prof_resolve(c);
        for (hipEvent_t e : c->evpool) (void)hipEventDestroy(e);
        (void)hipEventDestroy(c->ev0);
        (void)hipEventDestroy(c->ev1);
        (void)hipStreamDestroy(c->stream);
    }
    delete c;
}

const char *sray_gpu_last_error(sray_gpu_ctx *c) { return c ? c->err.c_str() : "no context"; }

int sray_gpu_scene_upload(sray_gpu_ctx *c, const sray_scene_desc *s, uint32_t *scene_id)
{
    if (!c || !scene_id) return SRAY_ERR_ARG;
    ApiLock lk(c);
    if (hipSetDevice(c->device) != hipSuccess) return fail(c, SRAY_ERR_HIP, "hipSetDevice failed");
    return scene_upload_guarded(c, s, scene_id);
}

int sray_gpu_autotune_octree(sray_gpu_ctx *c, const sray_scene_desc *s, const sray_render_params *p, int rounds,
                             const int32_t *items, const int32_t *depths, int ncand, int *best, double *best_ms)
{
    if (!c || !s || !p || !items || !depths || ncand <= 0) return SRAY_ERR_ARG;
    ApiLock lk(c);
    if (hipSetDevice(c->device) != hipSuccess) return fail(c, SRAY_ERR_HIP, "hipSetDevice failed");
    if (rounds < 1) rounds = 1;
    size_t bytes = (size_t)(p->x1 - p->x0) * (size_t)(p->y1 - p->y0) * 16;
    DevBuf out;
    int rc = ensure(c, out, bytes);
    if (rc) return rc;
    int bi = -1;
    double bms = 0;
    for (int k = 0; k < ncand; k++) {
        sray_scene_desc d = *s;
        d.mesh_oct_items = items[k];
        d.mesh_oct_depth = depths[k];
        uint32_t sid = 0;
        double ms = 0;
        if ((rc = scene_upload_guarded(c, &d, &sid)) == SRAY_OK) {
            rc = time_render(c, sid, p, rounds, dptr<float>(out), &ms);
            auto it = c->scenes.find(sid);
            (void)hipStreamSynchronize(c->stream);
            for (auto &b : it->second->bufs) release(b);
            c->scenes.erase(it);
        }
        if (rc == SRAY_ERR_LIMIT) continue; /* this depth does not fit the traversal stack */
        if (rc) {
            release(out);
            return rc;
        }
        if (bi < 0 || ms < bms) {
            bi = k;
            bms = ms;
        }
    }
    release(out);
    if (bi < 0) return fail(c, SRAY_ERR_LIMIT, "no octree candidate fits the traversal limits");
    if (best) *best = bi;
    if (best_ms) *best_ms = bms;
    return SRAY_OK;
}

int sray_gpu_scene_release(sray_gpu_ctx *c, uint32_t scene_id)
{
    if (!c) return SRAY_ERR_ARG;
    ApiLock lk(c);
    auto it = c->scenes.find(scene_id);
    if (it == c->scenes.end()) return fail(c, SRAY_ERR_NOTFOUND, "unknown scene %u", scene_id);
    int rc = engine_scene_busy(c, scene_id);
    if (rc) return rc;
    (void)hipStreamSynchronize(c->stream);
    for (auto &b : it->second->bufs) release(b);
    c->scenes.erase(it);
    return SRAY_OK;
}

int sray_gpu_scene_bytes(sray_gpu_ctx *c, uint32_t scene_id, uint64_t *bytes)
{
    if (!c || !bytes) return SRAY_ERR_ARG;
    ApiLock lk(c);
    if (scene_id == 0) {
        /* everything the context holds on the device */
        uint64_t t = engine_bytes(c);
        for (auto &kv : c->scenes) t += kv.second->bytes;
        for (auto &kv : c->pmaps) {
            t += kv.second->bytes;
            for (auto &g : kv.second->grids)
                t += g.second->pos.bytes + g.second->dir.bytes + g.second->nrm.bytes + g.second->colb.bytes +
                     g.second->cell.bytes;
        }
        DevBuf *bs[] = {&c->seltmp, &c->sorttmp, &c->nsel, &c->rays, &c->hitbuf, &c->counts, &c->gq, &c->grgb, &c->gcnt,
                        &c->gneed, &c->gfq, &c->goff, &c->gcand, &c->giota};
        for (DevBuf *d : bs) t += d->bytes;
        *bytes = t;
        return SRAY_OK;
    }
    GpuScene *g = find_scene(c, scene_id);
    if (!g) return fail(c, SRAY_ERR_NOTFOUND, "unknown scene %u", scene_id);
    *bytes = g->bytes;
    return SRAY_OK;
}

int sray_gpu_trace_device(sray_gpu_ctx *c, uint32_t scene_id, const sray_ray *d_rays, uint32_t n, sray_hit *d_hits,
                          uint32_t flags)
{
    if (!c || (n && (!d_rays || !d_hits))) return SRAY_ERR_ARG;
    ApiLock lk(c);
    return do_trace_device(c, scene_id, d_rays, n, d_hits, flags & ~SRAY_TRACE_COUNT, nullptr);
}

int sray_gpu_trace(sray_gpu_ctx *c, uint32_t scene_id, const sray_ray *rays, uint32_t n, sray_hit *hits,
                   uint32_t flags, sray_trace_counts *counts)
{
    if (!c || (n && (!rays || !hits))) return SRAY_ERR_ARG;
    ApiLock lk(c);
    if (hipSetDevice(c->device) != hipSuccess) return fail(c, SRAY_ERR_HIP, "hipSetDevice failed");
    if (counts) memset(counts, 0, sizeof *counts);
    if (!find_scene(c, scene_id)) return fail(c, SRAY_ERR_NOTFOUND, "unknown scene %u", scene_id);
    if (n == 0) return SRAY_OK;
    int rc;
    if ((rc = ensure(c, c->rays, (size_t)n * sizeof(sray_ray)))) return rc;
    if ((rc = ensure(c, c->hitbuf, (size_t)n * sizeof(sray_hit)))) return rc;
    if ((rc = ensure(c, c->counts, 8 * sizeof(unsigned long long)))) return rc;
    HIPCHK(c, hipMemcpyAsync(c->rays.p, rays, (size_t)n * sizeof(sray_ray), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemsetAsync(c->counts.p, 0, 8 * sizeof(unsigned long long), c->stream));
    rc = do_trace_device(c, scene_id, dptr<sray_ray>(c->rays), n, dptr<sray_hit>(c->hitbuf),
                         counts ? (flags | SRAY_TRACE_COUNT) : flags, dptr<unsigned long long>(c->counts));
    if (rc) return rc;
    HIPCHK(c, hipMemcpyAsync(hits, c->hitbuf.p, (size_t)n * sizeof(sray_hit), hipMemcpyDeviceToHost, c->stream));
    unsigned long long hc[8] = {0};
    HIPCHK(c, hipMemcpyAsync(hc, c->counts.p, sizeof hc, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    prof_resolve(c);
    if (counts) {
        counts->nodes = hc[0];
        counts->spheres = hc[1];
        counts->cylinders = hc[2];
        counts->triangles = hc[3];
        counts->item_refs = hc[4];
        counts->mesh_enters = hc[5];
        counts->item_boxes = hc[6];
    }
    return SRAY_OK;
}

int sray_gpu_pmap_upload(sray_gpu_ctx *c, uint32_t scene_id, const sray_photon *p, uint32_t n, int kind,
                         uint32_t *pmap_id)
{
    if (!c || !pmap_id) return SRAY_ERR_ARG;
    ApiLock lk(c);
    if (hipSetDevice(c->device) != hipSuccess) return fail(c, SRAY_ERR_HIP, "hipSetDevice failed");
    return do_pmap_build(c, scene_id, p, false, n, kind, pmap_id);
}

int sray_gpu_pmap_release(sray_gpu_ctx *c, uint32_t pmap_id)
{
    if (!c) return SRAY_ERR_ARG;
    ApiLock lk(c);
    auto it = c->pmaps.find(pmap_id);
    if (it == c->pmaps.end()) return fail(c, SRAY_ERR_NOTFOUND, "unknown photon map %u", pmap_id);
    int rc = engine_pmap_busy(c, pmap_id);
    if (rc) return rc;
    (void)hipStreamSynchronize(c->stream);
    free_pmap(*it->second);
    c->pmaps.erase(it);
    for (auto &kv : c->scenes)
        for (uint32_t &b : kv.second->pmap)
            if (b == pmap_id) b = 0;
    return SRAY_OK;
}

int sray_gpu_pmap_gather(sray_gpu_ctx *c, uint32_t pmap_id, const sray_gather_query *q, uint32_t n, float radius, int k,
                         int mode, float *rgb)
{
    if (!c || (n && (!q || !rgb))) return SRAY_ERR_ARG;
    ApiLock lk(c);
    if (hipSetDevice(c->device) != hipSuccess) return fail(c, SRAY_ERR_HIP, "hipSetDevice failed");
    return do_pmap_gather(c, pmap_id, q, n, radius, k, mode, rgb);
}

int sray_gpu_pmap_gather_device_noalloc(sray_gpu_ctx *c, uint32_t pmap_id, const sray_gather_query *d_q, uint32_t n,
                                        float radius, int mode, float *d_rgb)
{
    if (!c || (n && (!d_q || !d_rgb))) return SRAY_ERR_ARG;
    ApiLock lk(c);
    if (n > SRAY_NOALLOC_MAX_QUERIES)
        return fail(c, SRAY_ERR_LIMIT, "%u queries: the gather that allocates nothing takes at most %u", n,
                    SRAY_NOALLOC_MAX_QUERIES);
    if (mode != SRAY_GATHER_IRRADIANCE && mode != SRAY_GATHER_RADIANCE)
        return fail(c, SRAY_ERR_ARG, "bad gather mode (irradiance or radiance, kd-tree only)");
    if (!(radius > 0.0f)) return fail(c, SRAY_ERR_ARG, "radius must be > 0");
    auto it = c->pmaps.find(pmap_id);
    if (it == c->pmaps.end()) return fail(c, SRAY_ERR_NOTFOUND, "unknown photon map %u", pmap_id);
    if (n == 0) return SRAY_OK;
    if (hipSetDevice(c->device) != hipSuccess) return fail(c, SRAY_ERR_HIP, "hipSetDevice failed");
    uint8_t *scr = scratch_ptr(c->device);
    if (!scr) return fail(c, SRAY_ERR_HIP, "no scratch address for device %d", c->device);
    /* counts in the first 4 * max bytes, selection flags after them; k = 0
     * writes both for every query, so nothing needs clearing first */
    uint32_t *cnt = (uint32_t *)scr;
    uint8_t *need = scr + (size_t)SRAY_NOALLOC_MAX_QUERIES * 4;
    k_pm_gather<<<blocks(n, 128), 128, 0, c->stream>>>(it->second->tree, d_q, n, radius, mode, 0, d_rgb, cnt, need);
    HIPCHK(c, hipGetLastError());
    return SRAY_OK;
}

int sray_gpu_shoot_photons(sray_gpu_ctx *c, uint32_t scene_id, const sray_photon_params *pp, uint32_t *caustic,
                           uint32_t *global, sray_photon_stats *st)
{
    if (!c || !pp) return SRAY_ERR_ARG;
    ApiLock lk(c);
    uint32_t ids[2] = {0, 0};
    int rc = do_shoot(c, scene_id, pp, ids, st);
    if (caustic) *caustic = ids[0];
    if (global) *global = ids[1];
    return rc;
}

int sray_gpu_pmap_download(sray_gpu_ctx *c, uint32_t pmap_id, sray_photon *out, uint32_t cap, uint32_t *n)
{
    if (!c || !n) return SRAY_ERR_ARG;
    ApiLock lk(c);
    if (hipSetDevice(c->device) != hipSuccess) return fail(c, SRAY_ERR_HIP, "hipSetDevice failed");
    return do_pmap_download(c, pmap_id, out, cap, n);
}

int sray_gpu_malloc(sray_gpu_ctx *c, size_t bytes, void **dp)
{
    if (!c || !dp) return SRAY_ERR_ARG;
    ApiLock lk(c);
    if (hipSetDevice(c->device) != hipSuccess) return fail(c, SRAY_ERR_HIP, "hipSetDevice failed");
    hipError_t e = hipMalloc(dp, bytes ? bytes : 16);
    if (e != hipSuccess) {
        *dp = nullptr;
        return fail(c, SRAY_ERR_NOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    }
    return SRAY_OK;
}

int sray_gpu_free(sray_gpu_ctx *c, void *dp)
{
    if (!c) return SRAY_ERR_ARG;
    ApiLock lk(c);
    if (!dp) return SRAY_OK;
    (void)hipStreamSynchronize(c->stream);
    HIPCHK(c, hipFree(dp));
    return SRAY_OK;
}

int sray_gpu_memcpy(sray_gpu_ctx *c, void *dst, const void *src, size_t bytes, int kind)
{
    if (!c || (bytes && (!dst || !src)) || kind < 0 || kind > 2) return SRAY_ERR_ARG;
    ApiLock lk(c);
    hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice : kind == 1 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
    HIPCHK(c, hipMemcpyAsync(dst, src, bytes, k, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return SRAY_OK;
}

int sray_gpu_synchronize(sray_gpu_ctx *c)
{
    if (!c) return SRAY_ERR_ARG;
    ApiLock lk(c);
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipDeviceSynchronize());
    return SRAY_OK;
}

int sray_gpu_set_status_fd(sray_gpu_ctx *c, int fd)
{
    if (!c) return SRAY_ERR_ARG;
    c->status_fd.store(fd < 0 ? -1 : fd);
    return SRAY_OK;
}

int sray_gpu_set_profiling(sray_gpu_ctx *c, int on)
{
    if (!c) return SRAY_ERR_ARG;
    ApiLock lk(c);
    c->profiling = on != 0;
    return SRAY_OK;
}

int sray_gpu_kernel_stats(sray_gpu_ctx *c, sray_kernel_stat *out, int cap, int *n, int reset)
{
    if (!c || !n) return SRAY_ERR_ARG;
    ApiLock lk(c);
    (void)hipStreamSynchronize(c->stream);
    prof_resolve(c);
    *n = (int)c->kstats.size();
    int k = 0;
    for (auto &kv : c->kstats) {
        if (!out || k >= cap) break;
        memset(&out[k], 0, sizeof out[k]);
        snprintf(out[k].name, sizeof out[k].name, "%s", kv.first.c_str());
        out[k].launches = kv.second.launches;
        out[k].total_ms = kv.second.ms;
        k++;
    }
    if (reset) c->kstats.clear();
    return SRAY_OK;
}

} /* extern "C" */
