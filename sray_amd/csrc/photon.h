/*
 * photon.h — photon tracing on gfx950, one photon per lane.
 *
 * Replaces Light::gen_photon (src/light.cc:111-148), the emission loop of
 * Scene::build_caustics_map / build_global_map (src/scene.cc:445-549) and
 * the recursive Scene::trace_caustics_photon / trace_global_photon
 * (src/scene.cc:571-760), which here become one bounded loop per lane with
 * a mode switch: a global photon that meets a specular surface continues as
 * a caustic photon.  Each bounce is a closest-hit traversal (traverse.h)
 * and the surface point of shading (surface.h), so photons land exactly
 * where the render's rays would.
 *
 * Lane g shoots emission j = g - first(light) of its light; its random
 * draws are keyed by (j, kind << 24 | light, bounce), so the result does
 * not depend on the launch shape.  A lane writes its photon (or nothing) to
 * slot g; a stable compaction afterwards keeps emission order, which is the
 * order add_photon sees photons in the reference.
 */
#ifndef SRAY_PHOTON_H_
#define SRAY_PHOTON_H_

#include "surface.h"

namespace srd {

constexpr int PHOTON_IOR_DEPTH = 16; /* enclosing media remembered by a photon */

/* per light of one shoot (host-prepared) */
struct alignas(8) PhotonLightDev {
    float col0[3];   /* initial power: colour * energy * coverage */
    int32_t use_map; /* coverage != 0: emission directions restricted to full cells */
    uint32_t first;  /* first lane of this light */
    uint32_t count;  /* photons shot from this light */
    double scale;    /* 1 / count (PhotonMap::scale_photon_power, src/pmap.cc:90-100) */
};

struct CellTab {
    float cv[SRAY_PROJ_V + 1];
    float cu[SRAY_PROJ_U + 1];
};

struct PhotonJob {
    const PhotonLightDev *lights;
    const uint8_t *cells; /* per light: SRAY_PROJ_U * SRAY_PROJ_V bytes */
    int32_t nlights;
    int32_t kind;         /* SRAY_PMAP_CAUSTIC / SRAY_PMAP_GLOBAL */
    uint32_t total;
    int32_t t0, t1, iter;
    uint64_t seed;
    CellTab tab;
};

__device__ __forceinline__ void cell_of(const CellTab &T, float x, float y, float z, int &uc, int &vc)
{
    float r = sqrtf(x * x + y * y + z * z);
    float cy = y / r;
    int v = 0;
    for (int k = 1; k < SRAY_PROJ_V; k++) v += (cy <= T.cv[k]) ? 1 : 0;
    float rho = sqrtf(x * x + z * z);
    int u;
    if (!(rho > 0.0f)) {
        u = x < 0.0f ? SRAY_PROJ_U / 2 : 0;
    } else {
        float cx = x / rho;
        if (!(z < 0.0f)) {
            u = 0;
            for (int k = 1; k <= SRAY_PROJ_U / 2; k++) u += (cx <= T.cu[k]) ? 1 : 0;
        } else {
            u = SRAY_PROJ_U / 2;
            for (int k = SRAY_PROJ_U / 2 + 1; k < SRAY_PROJ_U; k++) u += (cx >= T.cu[k]) ? 1 : 0;
        }
    }
    uc = min(u, SRAY_PROJ_U - 1);
    vc = min(v, SRAY_PROJ_V - 1);
}

/* light sample point for photon work (Light::get_point, src/light.cc:106-109,
 * 272-275, 346-354), same draws as the host's photon_light_point */
__device__ __forceinline__ V3 photon_light_point(const DevLight &lt, uint64_t seed, uint32_t a, uint32_t b, uint32_t c,
                                                 uint32_t purpose, int msec)
{
    if (lt.type == SRAY_LIGHT_BOX) {
        float f0 = prand(seed, a, b, c, purpose, 0, 0);
        float f1 = prand(seed, a, b, c, purpose, 0, 1);
        float f2 = prand(seed, a, b, c, purpose, 0, 2);
        V3 v = v3(f0 * lt.dim[0] - 0.5f * lt.dim[0], f1 * lt.dim[1] - 0.5f * lt.dim[1], f2 * lt.dim[2] - 0.5f * lt.dim[2]);
        return xf_point(lt.xf, v);
    }
    float p[3] = {lt.pos[0], lt.pos[1], lt.pos[2]};
    if (lt.tab) {
        const float *r = time_row(lt.tab, lt.tab_t0, lt.tab_n, msec, 3);
        p[0] = r[0], p[1] = r[1], p[2] = r[2];
    }
    if (lt.type == SRAY_LIGHT_SPHERE) {
        for (uint32_t at = 0; at < PHOTON_SPHERE_ATTEMPTS; at++) {
            float d[3];
            if (marsaglia(prand(seed, a, b, c, purpose, at, 0), prand(seed, a, b, c, purpose, at, 1), d))
                return v3(p[0] + d[0] * lt.radius, p[1] + d[1] * lt.radius, p[2] + d[2] * lt.radius);
        }
    }
    return v3(p[0], p[1], p[2]);
}

template <bool COUNT>
__global__ __launch_bounds__(TRACE_BLOCK) void k_photon(SceneDev S, PhotonJob J, sray_photon *__restrict__ out,
                                                        uint8_t *__restrict__ keep,
                                                        unsigned long long *__restrict__ segs)
{
    __shared__ uint32_t stk[STACK * TRACE_BLOCK];
    __shared__ unsigned long long bseg;
    if (threadIdx.x == 0) bseg = 0;
    __syncthreads();
    uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t nseg = 0;
    if (g < J.total) {
        int li = 0;
        while (li + 1 < J.nlights && g >= J.lights[li].first + J.lights[li].count) li++;
        const PhotonLightDev &PL = J.lights[li];
        const DevLight &lt = S.lights[li];
        uint32_t a = g - PL.first;
        uint32_t b = ((uint32_t)J.kind << 24) | (uint32_t)li;
        uint64_t seed = J.seed;
        int msec = J.t0 == J.t1 ? J.t0
                                : (int)(rng_u32(seed, a, b, 0, rslot(K_PHOTON, P_TIME, 0, 0, 0)) %
                                        (uint32_t)(J.t1 - J.t0)) + J.t0;
        V3 ro = photon_light_point(lt, seed, a, b, 0, P_LPOINT, msec);
        float col[3] = {PL.col0[0], PL.col0[1], PL.col0[2]};
        /* emission direction: uniform, rejected until it falls in a full
         * cell of the projection map (ProjMap, src/projmap.cc:194-208) */
        float d[3] = {0.0f, 0.0f, 1.0f};
        const uint8_t *cells = J.cells + (size_t)li * (SRAY_PROJ_U * SRAY_PROJ_V);
        for (uint32_t at = 0; at < PHOTON_ATTEMPTS; at++) {
            if (!marsaglia(prand(seed, a, b, 0, P_EMIT, at, 0), prand(seed, a, b, 0, P_EMIT, at, 1), d)) continue;
            if (!PL.use_map) break;
            int uc, vc;
            cell_of(J.tab, d[0], d[1], d[2], uc, vc);
            if (cells[vc * SRAY_PROJ_U + uc]) break;
        }
        const float MAG = (float)SRAY_RAY_MAGNITUDE;
        V3 rd = v3(d[0] * MAG, d[1] * MAG, d[2] * MAG);
        int iter = J.iter;
        IorStackT<PHOTON_IOR_DEPTH> st;
        st.cur = 1.0f;
        st.n = 0;
        st.overflow = false;
        bool caustic_mode = J.kind == SRAY_PMAP_CAUSTIC;
        int type = 0; /* 0 direct, 1 caustic, 2 diffusely bounced */
        uint32_t bounce = 1;
        bool stored = false;
        V3 pdir = v3(0.0f, 0.0f, 0.0f), ppos = pdir, pnrm = pdir;
        int guard = 0;
        TraceCounts cnt = {0, 0, 0, 0, 0, 0, 0};
        while (true) {
            if (guard >= PHOTON_MAX_BOUNCES) break;
            guard++;
            if (iter <= 0) break;
            V3 incident = caustic_mode ? rd : normalized(rd);
            HitRec h = traverse<false, false>(S, ro, rd, msec, &stk[threadIdx.x], TRACE_BLOCK, cnt);
            nseg++;
            if (h.obj < 0) break;
            const DevObj &ob = S.objs[h.obj];
            const DevMat &m = S.mats[ob.mat];
            const bool mapped = (m.flags & MAT_MAPPED) != 0;
            const bool has_normal = (m.flags & SRAY_MAT_HAS_NORMAL) != 0;
            const bool bump = mapped && has_normal && m.map[SRAY_MAP_NORMAL] > 0;
            V3 Pw, N, G;
            float tu, tv;
            hit_surface(S, ob, msec, ro, rd, h.t, h.tri, mapped, Pw, N, tu, tv, caustic_mode && bump, G);
            float ray_mag = length(rd);
            bool entering;
            V3 normal;
            if (dot(rd, N) > 0.0f) {
                normal = neg(N);
                entering = false;
            } else {
                normal = N;
                entering = true;
            }
            /* a missing attribute reads as 1 (Material::def_attr, src/material.cc:77);
             * mapped ones at the hit's texture coordinates (src/scene.cc:583-586, 677-711) */
            float4 spc = (m.flags & SRAY_MAT_HAS_SPECULAR) ? m.ks : make_float4(1.0f, 1.0f, 1.0f, 1.0f);
            float refl = (m.flags & SRAY_MAT_HAS_REFLECT) ? m.reflect : 1.0f;
            float refr = (m.flags & SRAY_MAT_HAS_REFRACT) ? m.refract : 1.0f;
            float mat_ior = m.ior;
            float4 kd = m.kd;
            if (mapped) {
                if (m.flags & SRAY_MAT_HAS_SPECULAR) spc = mat_attr(S, m.map[SRAY_MAP_SPECULAR], spc, tu, tv);
                if (m.flags & SRAY_MAT_HAS_REFLECT) refl = mat_value(S, m.map[SRAY_MAP_REFLECT], refl, tu, tv);
                if (m.flags & SRAY_MAT_HAS_REFRACT) refr = mat_value(S, m.map[SRAY_MAP_REFRACT], refr, tu, tv);
                mat_ior = mat_value(S, m.map[SRAY_MAP_IOR], mat_ior, tu, tv);
                kd = mat_attr(S, m.map[SRAY_MAP_DIFFUSE], kd, tu, tv);
            }
            float rnum = prand(seed, a, b, bounce, P_RR, 0, 0);
            uint32_t here = bounce++;
            int act; /* 0 store, 1 reflect, 2 refract, 3 diffuse, 4 absorb */
            if (caustic_mode) {
                float ior = st.calc(entering, mat_ior);
                /* trace_caustics_photon takes the "normal" attribute's
                 * normal (src/scene.cc:603-606); trace_global_photon does not */
                if (has_normal) normal = bump ? bump_normal(S, m.map[SRAY_MAP_NORMAL], m.nscale, N, G, tu, tv) : N;
                float ray_dot_n = dot(dvs(neg(rd), ray_mag), normal);
                float sf0 = refr > 0.0f ? (ior - 1.0f) / (ior + 1.0f) : 1.0f;
                float fres = fresnel(sf0 * sf0, ray_dot_n);
                refl = fres * refl;
                refr = (1.0f - fres) * refr;
                act = rnum < refl ? 1 : rnum < refr ? 2 : 0;
            } else {
                float diff_avg = (kd.x + kd.y + kd.z) / 3.0f;
                float spec_avg = (spc.x + spc.y + spc.z) / 3.0f;
                if (rnum < diff_avg) {
                    act = 3;
                    float dd[3] = {0.0f, 0.0f, 1.0f};
                    for (uint32_t at = 0; at < PHOTON_SPHERE_ATTEMPTS; at++)
                        if (marsaglia(prand(seed, a, b, here, P_DIFF, at, 0), prand(seed, a, b, here, P_DIFF, at, 1), dd))
                            break;
                    V3 dir = v3(dd[0], dd[1], dd[2]);
                    if (dot(dir, normal) < 0.0f) dir = neg(dir);
                    float ndotl = dot(neg(incident), normal);
                    rd = mul(dir, MAG);
                    iter--;
                    ro = Pw;
                    type = 2;
                    col[0] = col[0] * (ndotl * (kd.x / diff_avg));
                    col[1] = col[1] * (ndotl * (kd.y / diff_avg));
                    col[2] = col[2] * (ndotl * (kd.z / diff_avg));
                    continue;
                } else if (rnum < diff_avg + spec_avg) {
                    float ior = st.calc(entering, mat_ior);
                    float ray_dot_n = dot(dvs(neg(rd), ray_mag), normal);
                    float sf0 = refr > 0.0f ? (ior - 1.0f) / (ior + 1.0f) : 1.0f;
                    float fres = fresnel(sf0 * sf0, ray_dot_n);
                    refl = fres * refl;
                    refr = (1.0f - fres) * refr;
                    float rn2 = prand(seed, a, b, here, P_RR, 0, 1);
                    act = rn2 < refl ? 1 : rn2 < refr ? 2 : 4;
                    if (act != 4) {
                        caustic_mode = true; /* continues as trace_caustics_photon */
                        guard = 0;
                    }
                } else {
                    act = 0;
                }
            }
            if (act == 4) break;
            if (act == 0) {
                pdir = caustic_mode ? normalized(rd) : incident;
                ppos = Pw;
                pnrm = normal;
                stored = true;
                break;
            }
            if (act == 1) {
                rd = reflect(rd, normal);
                iter--;
            } else {
                /* the refracted photon keeps its bounce budget (as the reference) */
                float ior = st.calc(entering, mat_ior);
                IorStackT<PHOTON_IOR_DEPTH> ts = st;
                if (entering) ts.enter(mat_ior);
                else ts.leave();
                V3 td = mul(refract(dvs(rd, ray_mag), normal, ior), ray_mag);
                if (dot(td, normal) > 0.0f) {
                    if (entering) ts.leave();
                    else ts.enter(mat_ior);
                }
                st = ts;
                rd = td;
            }
            ro = Pw;
            type = 1;
            col[0] = col[0] * spc.x;
            col[1] = col[1] * spc.y;
            col[2] = col[2] * spc.z;
        }
        bool ok = stored && (J.kind != SRAY_PMAP_CAUSTIC || type == 1);
        keep[g] = ok ? 1 : 0;
        if (ok) {
            sray_photon p;
            p.pos[0] = ppos.x, p.pos[1] = ppos.y, p.pos[2] = ppos.z;
            p.dir[0] = pdir.x, p.dir[1] = pdir.y, p.dir[2] = pdir.z;
            p.norm[0] = pnrm.x, p.norm[1] = pnrm.y, p.norm[2] = pnrm.z;
            for (int k = 0; k < 3; k++) p.col[k] = (float)((double)col[k] * PL.scale);
            out[g] = p;
        }
    }
    if (COUNT && nseg) atomicAdd(&bseg, (unsigned long long)nseg);
    if (COUNT) {
        __syncthreads();
        if (threadIdx.x == 0 && bseg) atomicAdd(segs, bseg);
    }
}

} // namespace srd

#endif
