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

/* interp_face, src/mesh.cc:321-445 (inverse bilinear u,v of the quad v0 v1
 * v2 v2; lerped attributes).  Evaluated in double: for this degenerate quad
 * du2 is zero up to rounding, and float rounding noise would pass the
 * XSMALL test into the quadratic branch; the reference computes in double. */
struct D3 {
    double x, y, z;
};
__device__ __forceinline__ D3 d3(float4 a) { return D3{(double)a.x, (double)a.y, (double)a.z}; }
__device__ __forceinline__ D3 d3(V3 a) { return D3{(double)a.x, (double)a.y, (double)a.z}; }
__device__ __forceinline__ D3 dsub(D3 a, D3 b) { return D3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ D3 dadd(D3 a, D3 b) { return D3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ D3 dneg(D3 a) { return D3{-a.x, -a.y, -a.z}; }
__device__ __forceinline__ D3 ddiv(D3 a, double s) { return D3{a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ double ddot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ D3 dcross(D3 a, D3 b)
{
    return D3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

__device__ inline void interp_face(const float4 *tri, const float *tn, const float *tg, const float *tuv, uint32_t ti,
                                   V3 ispos_f, V3 normal_f, V3 &on, V3 &og, float &ou, float &ov)
{
    const double XS = SRAY_XSMALL_NUMBER;
    D3 p00 = d3(tri[(size_t)ti * 3]), p10 = d3(tri[(size_t)ti * 3 + 1]), p01 = d3(tri[(size_t)ti * 3 + 2]), p11 = p01;
    D3 ispos = d3(ispos_f), normal = d3(normal_f);
    D3 pa = dadd(dsub(p00, p10), dsub(p11, p01));
    D3 pb = dsub(p10, p00);
    D3 pc = dsub(p01, p00);
    D3 pd = p00;
    D3 na = dcross(pa, normal), nb = dcross(pb, normal), nc = dcross(pc, normal);
    double du0 = ddot(nc, pd);
    double du1 = ddot(na, pd) + ddot(nc, pb);
    double du2 = ddot(na, pb);
    double dv0 = ddot(nb, pd);
    double dv1 = ddot(na, pd) + ddot(nb, pc);
    double dv2 = ddot(na, pc);
    double ru, rv;
    if (fabs(du2) <= XS) {
        double divisor = du1 - ddot(na, ispos);
        ru = fabs(divisor) > XS ? ((ddot(nc, ispos) - du0) / divisor) : 0.0;
    } else {
        D3 qux = ddiv(na, 2.0 * du2);
        D3 quy = ddiv(dneg(nc), du2);
        double dux = -du1 / (2.0 * du2);
        double duy = du0 / du2;
        double ka = dux + ddot(qux, ispos);
        double kb = duy + ddot(quy, ispos);
        double u0 = ka - sqrt(ka * ka - kb);
        double u1 = ka + sqrt(ka * ka - kb);
        ru = (u0 < 0.0 || u0 > 1.0) ? u1 : u0;
    }
    if (fabs(dv2) <= XS) {
        double divisor = dv1 - ddot(na, ispos);
        rv = fabs(divisor) > XS ? ((ddot(nb, ispos) - dv0) / divisor) : 0.0;
    } else {
        D3 qvx = ddiv(na, 2.0 * dv2);
        D3 qvy = ddiv(dneg(nb), dv2);
        double dvx = -dv1 / (2.0 * dv2);
        double dvy = dv0 / dv2;
        double ka = dvx + ddot(qvx, ispos);
        double kb = dvy + ddot(qvy, ispos);
        double v0 = ka - sqrt(ka * ka - kb);
        double v1 = ka + sqrt(ka * ka - kb);
        rv = (v0 < 0.0 || v0 > 1.0) ? v1 : v0;
    }
    auto lerp = [](double a, double b, double t) { return a + (b - a) * t; };
    auto l2 = [&](float a00, float a10, float a01) {
        double s0 = lerp((double)a00, (double)a01, rv);
        double s1 = lerp((double)a10, (double)a01, rv);
        return (float)lerp(s0, s1, ru);
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
    /* entries are read and written with constant indices only (selects over
     * the D slots), so the stack stays in registers */
    __device__ float at(uint32_t k) const
    {
        float v = s[0];
#pragma unroll
        for (int j = 1; j < D; j++) v = k == (uint32_t)j ? s[j] : v;
        return v;
    }
    __device__ void enter(float x)
    {
        if (n < (uint32_t)D) {
#pragma unroll
            for (int j = 0; j < D; j++) s[j] = n == (uint32_t)j ? cur : s[j];
            n++;
        } else {
            overflow = true;
        }
        cur = x;
    }
    __device__ void leave()
    {
        if (n > 0) {
            n--;
            cur = at(n);
        } else {
            cur = 1.0f;
        }
    }
    __device__ float calc(bool entering, float mat_ior) const
    {
        if (entering) return cur / mat_ior;
        float to = n > 0 ? at(n - 1) : 1.0f;
        return cur / to;
    }
};
using IorStack = IorStackT<IOR_DEPTH>;

/* world-space surface point Pw and normal N of the hit (ob, t, ti) of the
 * world ray (ro, rd) at time; N is normalised for triangles only, as
 * SurfPoint::normal in the reference.  want_uv: texture coordinates of
 * spheres and cylinders (meshes get them from interp_face when they carry
 * attributes; ATTR = false leaves that path out, for scenes without such
 * meshes, which keeps its registers out of the caller).  G (want_tan):
 * SurfPoint::tangent: (0,1,0) x the local normal of a quadric, transformed
 * as a normal (src/sphere.cc:121,128, src/cylinder.cc:173,181); a mesh's
 * interpolated tangent, transformed and normalised (src/mesh.cc:298-304;
 * a zero tangent stays zero). */
template <bool ATTR = true>
__device__ __forceinline__ void hit_surface(const SceneDev &S, const DevObj &ob, int time, V3 ro, V3 rd, float t,
                                            uint32_t ti, bool want_uv, V3 &Pw, V3 &N, float &tu, float &tv,
                                            bool want_tan, V3 &G)
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
    G = v3(0.0f, 0.0f, 0.0f);
    if (ob.type == SRAY_OBJ_SPHERE) {
        V3 nrm = dvs(p, ob.rad);
        N = xf_normal(I, nrm);
        if (ATTR && want_tan) G = xf_normal(I, cross(v3(0.0f, 1.0f, 0.0f), nrm));
        if (want_uv) {
            float r = length(p);
            tu = (atan2f(p.z, p.x) + 3.14159265358979f) / 6.28318530717959f;
            tv = r > 0.0f ? acosf(fmaxf(-1.0f, fminf(1.0f, p.y / r))) / 3.14159265358979f : 0.0f;
        }
    } else if (ob.type == SRAY_OBJ_CYLINDER) {
        V3 nrm = dvs(v3(p.x, 0.0f * p.y, p.z), ob.rad);
        N = xf_normal(I, nrm);
        if (ATTR && want_tan) G = xf_normal(I, cross(v3(0.0f, 1.0f, 0.0f), nrm));
        if (want_uv) {
            tu = (atan2f(p.z, p.x) + 3.14159265358979f) / 6.28318530717959f;
            tv = p.y;
        }
    } else {
        float4 A = S.tri[(size_t)ti * 3], B = S.tri[(size_t)ti * 3 + 1], C = S.tri[(size_t)ti * 3 + 2];
        V3 gn = v3(A.w, B.w, C.w); /* (v1-v0) x (v2-v0), stored at upload */
        V3 nrm = gn;
        if (ATTR && ob.has_attr) {
            V3 in, ig;
            interp_face(S.tri, (ob.has_attr & 1) ? S.tri_n : nullptr, (ob.has_attr & 4) ? S.tri_g : nullptr,
                        (ob.has_attr & 2) ? S.tri_uv : nullptr, ti, p, gn, in, ig, tu, tv);
            nrm = dot(in, in) < XSMALL ? gn : in;
            if (want_tan) {
                V3 tg = xf_normal(I, ig);
                float tl = length(tg);
                G = tl > 0.0f ? dvs(tg, tl) : tg;
            }
        }
        N = normalized(xf_normal(I, nrm));
    }
    Pw = xf_point(M, p);
}

/* Texture::lookup (src/texture.cc:60-83) with get_texel (:85-101): the
 * coordinates through the texture matrix, truncated to texels, then one
 * texel (nearest) or the average of the 5-texel cross (linear), each
 * wrapped (repeat) or clamped to [0, size] — the reference's clamp lets
 * the index reach the size itself, one texel past the row; the row-major
 * index is kept, and one past the last texel reads the last texel. */
__device__ __forceinline__ float4 tex_texel(const DevTex &T, int tx, int ty)
{
    if (T.wrap == SRAY_TEX_WRAP_CLAMP) {
        tx = tx < 0 ? 0 : (tx > T.w ? T.w : tx);
        ty = ty < 0 ? 0 : (ty > T.h ? T.h : ty);
    } else {
        tx %= T.w;
        ty %= T.h;
        if (tx < 0) tx += T.w;
        if (ty < 0) ty += T.h;
    }
    long long k = (long long)ty * T.w + tx;
    const long long last = (long long)T.w * T.h - 1;
    return T.texels[k > last ? last : k];
}
__device__ __forceinline__ int tex_coord(float c, int size)
{
    /* (int)(c * size), C truncation; coordinates beyond +-2^30 texels are
     * held there (a float-to-int conversion out of range is undefined) */
    float f = c * (float)size;
    f = fmaxf(fminf(f, 1073741824.0f), -1073741824.0f);
    return (int)f;
}
__device__ __forceinline__ float4 tex_lookup(const DevTex &T, float u, float v)
{
    const float x = T.xf[0] * u + T.xf[1] * v + T.xf[2] * 0.0f + T.xf[3];
    const float y = T.xf[4] * u + T.xf[5] * v + T.xf[6] * 0.0f + T.xf[7];
    const int tx = tex_coord(x, T.w), ty = tex_coord(y, T.h);
    if (T.filter == SRAY_TEX_FILTER_NEAREST) return tex_texel(T, tx, ty);
    const float4 a = tex_texel(T, tx, ty), b = tex_texel(T, tx + 1, ty), c = tex_texel(T, tx - 1, ty),
                 d = tex_texel(T, tx, ty + 1), e = tex_texel(T, tx, ty - 1);
    return make_float4(((((a.x + b.x) + c.x) + d.x) + e.x) / 5.0f, ((((a.y + b.y) + c.y) + d.y) + e.y) / 5.0f,
                       ((((a.z + b.z) + c.z) + d.z) + e.z) / 5.0f, ((((a.w + b.w) + c.w) + d.w) + e.w) / 5.0f);
}

/* MatAttrib::get_color (src/material.inl:23-28): texture x colour, or the
 * colour alone (map 0); get_value is its x component */
__device__ __forceinline__ float4 mat_attr(const SceneDev &S, int map, float4 col, float u, float v)
{
    if (map <= 0) return col;
    const float4 t = tex_lookup(S.texs[map - 1], u, v);
    return make_float4(t.x * col.x, t.y * col.y, t.z * col.z, t.w * col.w);
}
__device__ __forceinline__ float mat_value(const SceneDev &S, int map, float val, float u, float v)
{
    if (map <= 0) return val;
    return tex_lookup(S.texs[map - 1], u, v).x * val;
}

/* get_bump_normal (src/shader.cc:293-316): the normal map's colour * 2 - 1,
 * x and y scaled by the attribute's value, normalised, then taken from the
 * tangent frame (tangent, normal x tangent, normal) of the unflipped
 * surface normal */
__device__ __forceinline__ V3 bump_normal(const SceneDev &S, int map, float scale, V3 N, V3 G, float u, float v)
{
    const float4 c = tex_lookup(S.texs[map - 1], u, v);
    V3 n = v3(c.x * 2.0f - 1.0f, c.y * 2.0f - 1.0f, c.z * 2.0f - 1.0f);
    n.x = n.x * scale;
    n.y = n.y * scale;
    n = normalized(n);
    const V3 b = cross(N, G);
    return v3(G.x * n.x + b.x * n.y + N.x * n.z, G.y * n.x + b.y * n.y + N.y * n.z, G.z * n.x + b.z * n.y + N.z * n.z);
}

} // namespace srd

#endif
