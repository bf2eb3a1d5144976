/*
 * surface.h — what happens at a hit on gfx950: the surface point, normal and
 * texture coordinates of the winning item (SurfPoint, src/sphere.cc:107-129,
 * src/cylinder.cc:166-182, src/mesh.cc:147-163, 293-306, 321-445) and the
 * material helpers shared by the shading and photon kernels (fresnel,
 * src/shader.cc:285-291; refraction and the ior stack of vmath's Ray).
 */
#ifndef SRAY_SURFACE_H_
#define SRAY_SURFACE_H_

#include "traverse.h"

namespace srd {

/* interp_face, src/mesh.cc:321-445 (inverse bilinear u,v; lerped attributes) */
__device__ inline void interp_face(const float4 *tri, const float *tn, const float *tg, const float *tuv, uint32_t ti, V3 ispos,
                            V3 normal, V3 &on, V3 &og, float &ou, float &ov)
{
    float4 A = tri[(size_t)ti * 3], B = tri[(size_t)ti * 3 + 1], C = tri[(size_t)ti * 3 + 2];
    V3 p00 = v3(A.x, A.y, A.z), p10 = v3(B.x, B.y, B.z), p01 = v3(C.x, C.y, C.z), p11 = p01;
    V3 pa = add(sub(p00, p10), sub(p11, p01));
    V3 pb = sub(p10, p00);
    V3 pc = sub(p01, p00);
    V3 pd = p00;
    V3 na = cross(pa, normal), nb = cross(pb, normal), nc = cross(pc, normal);
    float du0 = dot(nc, pd);
    float du1 = dot(na, pd) + dot(nc, pb);
    float du2 = dot(na, pb);
    float dv0 = dot(nb, pd);
    float dv1 = dot(na, pd) + dot(nb, pc);
    float dv2 = dot(na, pc);
    float ru, rv;
    if (fabsf(du2) <= XSMALL) {
        float divisor = du1 - dot(na, ispos);
        ru = fabsf(divisor) > XSMALL ? ((dot(nc, ispos) - du0) / divisor) : 0.0f;
    } else {
        V3 qux = dvs(na, 2.0f * du2);
        V3 quy = dvs(neg(nc), du2);
        float dux = -du1 / (2.0f * du2);
        float duy = du0 / du2;
        float ka = dux + dot(qux, ispos);
        float kb = duy + dot(quy, ispos);
        float u0 = ka - sqrtf(ka * ka - kb);
        float u1 = ka + sqrtf(ka * ka - kb);
        ru = (u0 < 0.0f || u0 > 1.0f) ? u1 : u0;
    }
    if (fabsf(dv2) <= XSMALL) {
        float divisor = dv1 - dot(na, ispos);
        rv = fabsf(divisor) > XSMALL ? ((dot(nb, ispos) - dv0) / divisor) : 0.0f;
    } else {
        V3 qvx = dvs(na, 2.0f * dv2);
        V3 qvy = dvs(neg(nb), dv2);
        float dvx = -dv1 / (2.0f * dv2);
        float dvy = dv0 / dv2;
        float ka = dvx + dot(qvx, ispos);
        float kb = dvy + dot(qvy, ispos);
        float v0 = ka - sqrtf(ka * ka - kb);
        float v1 = ka + sqrtf(ka * ka - kb);
        rv = (v0 < 0.0f || v0 > 1.0f) ? v1 : v0;
    }
    auto lerp = [](float a, float b, float t) { return a + (b - a) * t; };
    auto l2 = [&](float a00, float a10, float a01) {
        float s0 = lerp(a00, a01, rv);
        float s1 = lerp(a10, a01, rv);
        return lerp(s0, s1, ru);
    };
    if (tuv) {
        const float *t = tuv + (size_t)ti * 6;
        ou = l2(t[0], t[2], t[4]);
        ov = l2(t[1], t[3], t[5]);
    } else {
        ou = ov = 0.0f;
    }
    if (tn) {
        const float *n = tn + (size_t)ti * 9;
        on = v3(l2(n[0], n[3], n[6]), l2(n[1], n[4], n[7]), l2(n[2], n[5], n[8]));
    } else on = v3(0.0f, 0.0f, 0.0f);
    if (tg) {
        const float *g = tg + (size_t)ti * 9;
        og = v3(l2(g[0], g[3], g[6]), l2(g[1], g[4], g[7]), l2(g[2], g[5], g[8]));
    } else og = v3(0.0f, 0.0f, 0.0f);
}

__device__ __forceinline__ float fresnel(float r, float cosa)
{
    float inv = 1.0f - cosa;
    float inv_sq = inv * inv;
    return r + (1.0f - r) * inv_sq * inv_sq * inv;
}

/* refraction of the unit incident d about n; total internal reflection
 * returns the mirror direction (libvmath Vector3::refraction, fixed here) */
__device__ __forceinline__ V3 refract(V3 d, V3 n, float ior)
{
    float cos_inc = -dot(d, n);
    float radical = 1.0f + ior * ior * (cos_inc * cos_inc - 1.0f);
    if (radical < 0.0f) return reflect(d, n);
    float beta = ior * cos_inc - sqrtf(radical);
    return v3(d.x * ior + n.x * beta, d.y * ior + n.y * beta, d.z * ior + n.z * beta);
}

/* per-ray ior stack (vmath Ray::enter/leave/calc_ior): the current ior and
 * the iors of up to D enclosing media.  Entering a medium with a full stack
 * drops the outer ior and sets overflow. */
template <int D> struct IorStackT {
    float cur;
    float s[D];
    uint32_t n;
    bool overflow;
    __device__ void enter(float x)
    {
        if (n < (uint32_t)D) s[n++] = cur;
        else overflow = true;
        cur = x;
    }
    __device__ void leave() { cur = n > 0 ? s[--n] : 1.0f; }
    __device__ float calc(bool entering, float mat_ior) const
    {
        if (entering) return cur / mat_ior;
        float to = n > 0 ? s[n - 1] : 1.0f;
        return cur / to;
    }
};
using IorStack = IorStackT<IOR_DEPTH>;

/* world-space surface point Pw and normal N of the hit (ob, t, ti) of the
 * world ray (ro, rd) at time; N is normalised for triangles only, as
 * SurfPoint::normal in the reference.  want_uv: texture coordinates of
 * spheres and cylinders (meshes get them from interp_face when they carry
 * attributes). */
__device__ __forceinline__ void hit_surface(const SceneDev &S, const DevObj &ob, int time, V3 ro, V3 rd, float t,
                                            uint32_t ti, bool want_uv, V3 &Pw, V3 &N, float &tu, float &tv)
{
    float I[12], M[12];
    obj_inv_at(ob, time, I);
    obj_xf_at(ob, time, M);
    V3 lo = xf_point(I, ro), ld = xf_vec(I, rd);
    V3 p = add(lo, mul(ld, t));
    if (ob.type != SRAY_OBJ_MESH) {
        /* surface point of a quadric from the double-precision root */
        double td;
        bool ok = ob.type == SRAY_OBJ_SPHERE ? sphere_td((double)ob.rad, lo, ld, td)
                                             : cylinder_td((double)ob.rad, (double)ob.start, (double)ob.end, lo, ld, td);
        if (ok)
            p = v3((float)((double)lo.x + (double)ld.x * td), (float)((double)lo.y + (double)ld.y * td),
                   (float)((double)lo.z + (double)ld.z * td));
    }
    tu = 0.0f;
    tv = 0.0f;
    if (ob.type == SRAY_OBJ_SPHERE) {
        V3 nrm = dvs(p, ob.rad);
        N = xf_normal(I, nrm);
        if (want_uv) {
            float r = length(p);
            tu = (atan2f(p.z, p.x) + 3.14159265358979f) / 6.28318530717959f;
            tv = r > 0.0f ? acosf(fmaxf(-1.0f, fminf(1.0f, p.y / r))) / 3.14159265358979f : 0.0f;
        }
    } else if (ob.type == SRAY_OBJ_CYLINDER) {
        V3 nrm = dvs(v3(p.x, 0.0f * p.y, p.z), ob.rad);
        N = xf_normal(I, nrm);
        if (want_uv) {
            tu = (atan2f(p.z, p.x) + 3.14159265358979f) / 6.28318530717959f;
            tv = p.y;
        }
    } else {
        float4 A = S.tri[(size_t)ti * 3], B = S.tri[(size_t)ti * 3 + 1], C = S.tri[(size_t)ti * 3 + 2];
        V3 e1 = v3(B.x - A.x, B.y - A.y, B.z - A.z), e2 = v3(C.x - A.x, C.y - A.y, C.z - A.z);
        V3 gn = cross(e1, e2);
        V3 nrm = gn;
        if (ob.has_attr) {
            V3 in, ig;
            interp_face(S.tri, (ob.has_attr & 1) ? S.tri_n : nullptr, (ob.has_attr & 4) ? S.tri_g : nullptr,
                        (ob.has_attr & 2) ? S.tri_uv : nullptr, ti, p, gn, in, ig, tu, tv);
            nrm = dot(in, in) < XSMALL ? gn : in;
        }
        N = normalized(xf_normal(I, nrm));
    }
    Pw = xf_point(M, p);
}

} // namespace srd

#endif
