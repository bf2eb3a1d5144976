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

} // namespace srd

#endif
