/*
 * engine.hip — the wavefront renderer of the S-ray hot path on gfx950, for
 * one request (sray_gpu_render) or many in flight (sray_gpu_submit/poll).
 *
 * Replaces render_frame / start_frame / render_block (src/render.cc:124-244),
 * Scene::trace_ray (src/scene.cc:551-569) and the shade_phong recursion
 * (src/shader.cc:83-283).  A request is a pixel rectangle of one resident
 * scene seen through a camera; its pixels live in a pixel arena (per-pixel
 * framebuffer sum, sample count, request slot).  A pass takes one sample of
 * every active pixel of every request in flight: primary rays, then bounce
 * levels of closest-hit trace -> shade -> shadow trace -> accumulation ->
 * stable compaction of the children, depth first over sub-chunks.  After a
 * pass the pixels that stop (sample budget or variance) leave the active
 * list; a request whose pixels have all stopped is written out and its
 * slot and arena ranges are reused.  New requests join at pass boundaries
 * (continuous batching).  Every random draw is keyed by the request's seed,
 * pixel, sample and path, and every pixel's contributions are summed by one
 * thread in queue order, so a request's image does not depend on what else
 * is in flight.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <deque>

#include "runtime.h"
#include "surface.h"

using namespace srd;
using namespace srt;

namespace {

/* Camera::get_primary_ray, src/camera.cc:119-142 */
__device__ __forceinline__ void primary_ray(const CamDev &C, const RenderDev &P, int x, int y, int sample, V3 &o,
                                            V3 &d, int &time)
{
    float ysz = 2.0f;
    float xsz = P.aspect * ysz;
    float px = ((float)x / (float)P.width) * xsz - xsz / 2.0f;
    float py = 1.0f - ((float)y / (float)P.height) * ysz;
    uint32_t pix = (uint32_t)y * (uint32_t)P.width + (uint32_t)x;
    if (sample > 0) {
        float f0 = rng_unit(P.seed, pix, (uint32_t)sample, 0u, rslot(K_CAMERA, 0, 0, 0, 0));
        float f1 = rng_unit(P.seed, pix, (uint32_t)sample, 0u, rslot(K_CAMERA, 0, 0, 0, 1));
        px = px + xsz * (f0 - 0.5f) / (float)P.width;
        py = py + ysz * (f1 - 0.5f) / (float)P.height;
    }
    V3 dl = v3(px * C.max_dist, py * C.max_dist, C.fz * C.max_dist);
    time = P.frame_time;
    if (P.mblur && sample > 0 && C.shutter > 0) {
        uint32_t u = rng_u32(P.seed, pix, (uint32_t)sample, 0u, rslot(K_CAMERA, 0, 0, 0, 2));
        time += (int)(u % (uint32_t)C.shutter) - C.shutter / 2;
    }
    /* ray.transformed(get_matrix(ray.time)) */
    const float *cm = C.tab ? time_row(C.tab, C.t0, C.n, time, 12) : C.cam;
    o = xf_point(cm, v3(0.0f, 0.0f, 0.0f));
    d = xf_vec(cm, dl);
}

__global__ void k_primary(CamDev C, RenderDev P, int sample, sray_ray *out, uint32_t n)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int x = P.x0 + (int)(i % (uint32_t)P.W), y = P.y0 + (int)(i / (uint32_t)P.W);
    V3 o, d;
    int time;
    primary_ray(C, P, x, y, sample, o, d, time);
    sray_ray r;
    r.origin[0] = o.x;
    r.origin[1] = o.y;
    r.origin[2] = o.z;
    r.dir[0] = d.x;
    r.dir[1] = d.y;
    r.dir[2] = d.z;
    r.time = time;
    r.pad_ = 0;
    out[i] = r;
}

/* one primary ray per active pixel of the chunk [c0, c0+m), at the
 * pixel's next sample index */
template <bool UNI>
__global__ void k_raygen(Tabs T, const uint32_t *__restrict__ active, uint32_t c0, uint32_t m, QRay *__restrict__ q,
                         uint32_t *__restrict__ count)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) *count = m; /* level 0's count for the level kernels */
    if (i >= m) return;
    const uint32_t e = active[c0 + i], pr = ent_pix(e);
    uint32_t rq = UNI ? 0u : T.preq[pr];
    const ReqDev &Q = T.reqs[rq];
    const RenderDev &P = Q.P;
    uint32_t lp = pr - Q.pix_base;
    int x = P.x0 + (int)(lp % (uint32_t)P.W), y = P.y0 + (int)(lp / (uint32_t)P.W);
    int sample = T.nsamp[pr] + (int)ent_off(e);
    V3 o, d;
    int time;
    primary_ray(Q.C, P, x, y, sample, o, d, time);
    QRay r;
    r.o = make_float4(o.x, o.y, o.z, __int_as_float(time));
    r.d = make_float4(d.x, d.y, d.z, 1.0f);
    r.w = make_float4(1.0f, 1.0f, 1.0f, 1.0f);
    r.m = make_uint4(i, 1u, (uint32_t)P.iter, rq << SLOT_SHIFT);
    r.stk = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    q[i] = r;
}

/* The traversal kernels walk the step in runs of rays, one run per wave
 * at a time, wave w taking runs w, w + W, w + 2W, ... of the W waves of
 * the grid: with a grid as large as the step (runs of 64, one per wave)
 * the hardware dispatcher balances the load; with a persistent grid (as
 * many waves as stay resident) each wave loops, and a run holds the
 * smallest power of two of rays in [RUN_MIN, 64] that covers the step in
 * one run per wave.  A deep bounce level of a few tens of thousands of
 * rays then spreads over every SIMD in waves of 16 or 32 rays: more waves
 * hide each other's load latency, and a wave lasts as long as the slowest
 * of fewer rays (the other lanes help in the wave-wide leaf tests).  Either
 * way the launch shape does not depend on the step's count, which the
 * kernels read from the device, and every ray keeps its own walk. */
constexpr uint32_t RUN_MIN = 16;
template <bool PERSIST> __device__ __forceinline__ uint32_t run_len(uint32_t n)
{
    if (!PERSIST) return 64u;
    const uint32_t waves = gridDim.x * blockDim.x / 64u;
    uint32_t r = RUN_MIN;
    while (r < 64u && (uint64_t)r * waves < (uint64_t)n) r <<= 1;
    return r;
}
__device__ __forceinline__ uint32_t wave_index() { return (blockIdx.x * blockDim.x + threadIdx.x) / 64u; }
__device__ __forceinline__ uint32_t grid_waves() { return gridDim.x * blockDim.x / 64u; }

/* closest hit of every queued ray (Scene::cast_ray with a SurfPoint) */
/* No occupancy hints: with item boxes in the walk, the dense instance
 * spilled under a 6- or 7-wave hint (80 / 72 VGPRs). */
template <bool COUNT, bool UNI, int BS, bool PERSIST, bool SPARSE>
__global__ __launch_bounds__(BS) void k_trace_q(Tabs T, const QRay *__restrict__ q, StepN sn,
                                                int4 *__restrict__ hits, unsigned long long *counts)
{
    __shared__ uint32_t stk[STACK * BS];
    const uint32_t n = step_n(sn);
    TraceCounts cnt = {0, 0, 0, 0, 0, 0, 0};
    /* without PERSIST the grid has one run per wave: no loop state is kept */
    const uint32_t run = run_len<PERSIST>(n), lane = __lane_id();
    for (uint32_t base = wave_index() * run; base < n; base = PERSIST ? base + grid_waves() * run : n) {
        /* every lane of the wave walks (the leaf tests of sparse waves are
         * shared); a lane past the run or the end has no ray */
        const uint32_t i = base + lane;
        const bool live = lane < run && i < n;
        auto walk = [&](const auto &src) {
            if (UNI) {
                HitRec h = traverse_src<false, COUNT, true, SPARSE>(trav_scene_uniform(T.scenes[0]), src, live, &stk[threadIdx.x], BS, cnt);
                if (live) hits[i] = make_int4(__float_as_int(h.t), h.obj, (int)h.tri, (int)h.rank);
            } else {
                /* one walk per distinct scene of the wave (a wave's rays
                 * nearly always come from one request), each with the
                 * scene's descriptor wave-uniform */
                const uint32_t sidx = live ? T.reqs[q[i].m.w >> SLOT_SHIFT].scene : 0u;
                unsigned long long todo = __ballot(live);
                while (todo) {
                    const uint32_t s0 = read_lane_u(sidx, __ffsll(todo) - 1);
                    const bool mine = live && sidx == s0;
                    todo &= ~__ballot(mine);
                    HitRec h = traverse_src<false, COUNT, true, SPARSE>(trav_scene_uniform(T.scenes[s0]), src, mine, &stk[threadIdx.x], BS, cnt);
                    if (mine) hits[i] = make_int4(__float_as_int(h.t), h.obj, (int)h.tri, (int)h.rank);
                }
            }
        };
        if (SPARSE) {
            /* sparse levels hold few waves per SIMD, so registers do not
             * cost occupancy: the world ray stays in registers (no reload
             * on entering or leaving a mesh) */
            float4 o = make_float4(0.0f, 0.0f, 0.0f, 0.0f), d = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
            if (live) {
                o = q[i].o;
                d = q[i].d;
            }
            walk(RayRegs{v3(o.x, o.y, o.z), v3(d.x, d.y, d.z), __float_as_int(o.w)});
        } else {
            /* the world ray is re-read from the queue when needed instead
             * of held in registers */
            walk(RayRecord<QRay>{q, i});
        }
    }
    flush_counts<COUNT>(cnt, counts);
}

/* occlusion of every shadow ray (any hit: cast_ray(ray) != NULL,
 * src/shader.cc:266), L slots per ray of the step */
template <bool COUNT, bool UNI, int BS, bool PERSIST, bool SPARSE>
__global__ __launch_bounds__(BS) void k_trace_shadow(Tabs T, const SRay *__restrict__ sq,
                                                     const float4 *__restrict__ scon, StepN sn, int L,
                                                     uint8_t *__restrict__ occ, unsigned long long *counts)
{
    __shared__ uint32_t stk[STACK * BS];
    const uint32_t n = step_n(sn) * (uint32_t)L;
    TraceCounts cnt = {0, 0, 0, 0, 0, 0, 0};
    const uint32_t run = run_len<PERSIST>(n), lane = __lane_id();
    for (uint32_t base = wave_index() * run; base < n; base = PERSIST ? base + grid_waves() * run : n) {
        /* every lane of the wave walks; past the run or the end, or in an
         * unused slot, it has no ray.  own: the slot is this lane's */
        const uint32_t i = base + lane;
        const bool own = lane < run && i < n;
        bool live = own;
        float4 o = make_float4(0.0f, 0.0f, 0.0f, 0.0f), d = o;
        if (live) {
            d = sq[i].d;
            o = sq[i].o;
            live = ((const float *)scon)[(size_t)i * 4 + 3] != 0.0f;
        }
        /* the world ray stays in registers here: this kernel is at full
         * occupancy already, and re-reading it would only add loads */
        const RayRegs src{v3(o.x, o.y, o.z), v3(d.x, d.y, d.z), __float_as_int(o.w)};
        if (UNI) {
            HitRec h = traverse_src<true, COUNT, true, SPARSE>(trav_scene_uniform(T.scenes[0]), src, live, &stk[threadIdx.x], BS, cnt);
            if (own) occ[i] = h.obj >= 0 ? 1 : 0;
        } else {
            /* one walk per distinct scene of the wave, as in k_trace_q */
            const uint32_t sidx = live ? T.reqs[__float_as_uint(d.w)].scene : 0u;
            if (own && !live) occ[i] = 0;
            unsigned long long todo = __ballot(live);
            while (todo) {
                const uint32_t s0 = read_lane_u(sidx, __ffsll(todo) - 1);
                const bool mine = live && sidx == s0;
                todo &= ~__ballot(mine);
                HitRec h = traverse_src<true, COUNT, true, SPARSE>(trav_scene_uniform(T.scenes[s0]), src, mine, &stk[threadIdx.x], BS, cnt);
                if (mine) occ[i] = h.obj >= 0 ? 1 : 0;
            }
        }
    }
    flush_counts<COUNT>(cnt, counts);
}

/* light sample point: Light::get_point src/light.cc:106-109, SphLight
 * :272-275, BoxLight :346-354 (matrix at time 0) */
__device__ __forceinline__ V3 sph_rand(uint64_t seed, uint32_t pix, uint32_t sample, uint32_t path, uint32_t kind,
                                       uint32_t i1, uint32_t i2, float r)
{
    /* Marsaglia (1972): uniform on the sphere from a rejection pair */
    for (uint32_t a = 0; a < 64; a++) {
        float u1 = rng_unit(seed, pix, sample, path, rslot(kind, i1, i2, a, 0));
        float u2 = rng_unit(seed, pix, sample, path, rslot(kind, i1, i2, a, 1));
        float x1 = 2.0f * u1 - 1.0f, x2 = 2.0f * u2 - 1.0f;
        float s = x1 * x1 + x2 * x2;
        if (s < 1.0f && s > 0.0f) {
            float qq = 2.0f * sqrtf(1.0f - s);
            V3 d = v3(x1 * qq, x2 * qq, 1.0f - 2.0f * s);
            return mul(d, r);
        }
    }
    return v3(0.0f, 0.0f, r);
}

__device__ __forceinline__ V3 light_point(const DevLight &lt, uint64_t seed, uint32_t pix, uint32_t sample,
                                          uint32_t path, uint32_t li, uint32_t si, int time)
{
    if (lt.type == SRAY_LIGHT_BOX) {
        float fx = rng_unit(seed, pix, sample, path, rslot(K_LIGHT, li, si, 0, 0));
        float fy = rng_unit(seed, pix, sample, path, rslot(K_LIGHT, li, si, 0, 1));
        float fz = rng_unit(seed, pix, sample, path, rslot(K_LIGHT, li, si, 0, 2));
        V3 v = v3(fx * lt.dim[0] - 0.5f * lt.dim[0], fy * lt.dim[1] - 0.5f * lt.dim[1],
                  fz * lt.dim[2] - 0.5f * lt.dim[2]);
        return xf_point(lt.xf, v);
    }
    /* get_position(msec) of a point / sphere light */
    V3 p = v3(lt.pos[0], lt.pos[1], lt.pos[2]);
    if (lt.tab) {
        const float *r = time_row(lt.tab, lt.tab_t0, lt.tab_n, time, 3);
        p = v3(r[0], r[1], r[2]);
    }
    if (lt.type == SRAY_LIGHT_SPHERE) {
        V3 s = sph_rand(seed, pix, sample, path, K_LIGHT, li, si, lt.radius);
        return add(p, s);
    }
    return p;
}

/* Shading of one queued ray and its hit (shade_phong, src/shader.cc:83-247,
 * calc_lighting :250-283, shade_undef :36-71; the miss colour of
 * Scene::trace_ray, src/scene.cc:568): the direct term, the shadow rays
 * with their contributions, and the children (reflect, refract, diffuse),
 * written through the sink (the level arrays of the wavefront, or a
 * lane's registers on the one-child chain).  Counts of spawned rays by
 * kind go to c_sh, c_rl, c_rr, c_df. */
template <bool PHOTONS, bool UNI, bool SCN1, bool ATTR, class Sink>
__device__ __forceinline__ void shade_ray(const Tabs &T, const uint32_t *__restrict__ active, uint32_t c0,
                                          const QRay &R, int4 h, int L, int K, Sink &sink, uint32_t &c_sh,
                                          uint32_t &c_rl, uint32_t &c_rr, uint32_t &c_df)
{
    uint32_t slot = R.m.x, path = R.m.y, flg = (R.m.w >> 8) & 0xffu;
    uint32_t rq = UNI ? 0u : R.m.w >> SLOT_SHIFT;
    const ReqDev &Q = T.reqs[rq];
    const RenderDev &P = Q.P;
    /* SCN1: every request in flight renders scene 0 */
    const SceneDev S = T.scenes[(UNI || SCN1) ? 0u : Q.scene];
    const uint32_t e = active[c0 + slot], pr = ent_pix(e);
    uint32_t lp = pr - Q.pix_base;
    int sample = T.nsamp[pr] + (int)ent_off(e);
    int iter = (int)R.m.z;
    V3 W = v3(R.w.x, R.w.y, R.w.z);
    float energy = R.d.w;
    int time = __float_as_int(R.o.w);
    V3 ro = v3(R.o.x, R.o.y, R.o.z), rd = v3(R.d.x, R.d.y, R.d.z);
    float4 dir = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    int used = 0;        /* shadow slots written */
    uint32_t spawn = 0;  /* reflect (bit 0) / refract (bit 1) child written */
    int dwritten = 0;    /* diffuse child flags written */
    sink.slot(slot);
    if (PHOTONS) sink.gather_clear();
    if (h.y < 0) {
        /* miss: env_color (src/scene.cc:568) */
        dir.x = W.x * S.env_color.x;
        dir.y = W.y * S.env_color.y;
        dir.z = W.z * S.env_color.z;
        if (path == 1u) dir.w = S.env_color.w;
        sink.set_direct(dir);
    } else {
        float t = __int_as_float(h.x);
        const DevObj &ob = S.objs[h.y];
        const DevMat &mat = S.mats[ob.mat];
        bool phong = (mat.flags & SRAY_MAT_PHONG) != 0;
        /* texture maps: compiled into the ATTR instances only (scenes with
         * mapped materials count as scenes with attributes) */
        const bool mapped = ATTR && phong && (mat.flags & MAT_MAPPED);
        const bool has_normal = phong && (mat.flags & SRAY_MAT_HAS_NORMAL);
        const bool bump = mapped && has_normal && mat.map[SRAY_MAP_NORMAL] > 0;
        V3 Pw, N, G;
        float tu, tv;
        hit_surface<ATTR>(S, ob, time, ro, rd, t, (uint32_t)h.z, !phong || mapped, Pw, N, tu, tv, bump, G);
        if (path == 1u) dir.w = 1.0f;
        uint32_t pix = (uint32_t)(P.y0 + (int)(lp / (uint32_t)P.W)) * (uint32_t)P.width + (uint32_t)(P.x0 + (int)(lp % (uint32_t)P.W));
        int nl = S.nlights;
        if (!phong) {
            /* shade_undef: checker texture from texcoords */
            int cx = (int)(tu * 20.0f), cy = (int)(tv * 10.0f);
            V3 tex = ((cx % 2) == (cy % 2)) ? v3(1.0f, 0.3f, 0.2f) : v3(0.2f, 0.3f, 1.0f);
            dir.x = W.x * (0.1f * tex.x);
            dir.y = W.y * (0.1f * tex.y);
            dir.z = W.z * (0.1f * tex.z);
            for (int l = 0; l < nl; l++) {
                const DevLight &lt = S.lights[l];
                V3 lpos = light_point(lt, P.seed, pix, (uint32_t)sample, path, (uint32_t)l, 0u, time);
                V3 ldir = sub(lpos, Pw);
                float ldist = length(ldir);
                V3 sd = ldir;
                ldir = dvs(ldir, ldist);
                V3 vdir = neg(normalized(rd));
                V3 hdir = normalized(add(ldir, vdir));
                float ndotl = fmaxf(dot(N, ldir), 0.0f);
                float ndoth = fmaxf(dot(N, hdir), 0.0f);
                float sp = powf(ndoth, 50.0f);
                SRay s;
                s.o = make_float4(Pw.x, Pw.y, Pw.z, __int_as_float(time));
                s.d = make_float4(sd.x, sd.y, sd.z, __uint_as_float(rq));
                sink.shadow(l, s,
                            make_float4(W.x * (tex.x * (lt.color[0] * ndotl) + lt.color[0] * sp),
                                        W.y * (tex.y * (lt.color[1] * ndotl) + lt.color[1] * sp),
                                        W.z * (tex.z * (lt.color[2] * ndotl) + lt.color[2] * sp), 1.0f));
                c_sh++;
            }
            used = nl;
            sink.set_direct(dir);
        } else {
            bool entering;
            V3 normal;
            if (dot(rd, N) > 0.0f) {
                normal = neg(N);
                entering = false;
            } else {
                normal = N;
                entering = true;
            }
            float ray_mag = length(rd);
            V3 incident = dvs(rd, ray_mag);
            float4 kd = mat.kd;
            float4 ks = (mat.flags & SRAY_MAT_HAS_SPECULAR) ? mat.ks : make_float4(0.0f, 0.0f, 0.0f, 1.0f);
            float shininess = mat.shininess;
            if (mapped) {
                kd = mat_attr(S, mat.map[SRAY_MAP_DIFFUSE], kd, tu, tv);
                if (mat.flags & SRAY_MAT_HAS_SPECULAR) ks = mat_attr(S, mat.map[SRAY_MAP_SPECULAR], ks, tu, tv);
                shininess = mat_value(S, mat.map[SRAY_MAP_SHININESS], shininess, tu, tv);
            }
            if (flg & QF_DIFFUSE) {
                /* diffuse GI ray: radiance estimate from the global map
                 * (src/shader.cc:119-123) */
                if (PHOTONS) {
                    GatherReq g;
                    g.pos = make_float4(Pw.x, Pw.y, Pw.z, __uint_as_float(2u));
                    g.nrm = make_float4(normal.x, normal.y, normal.z, 0.0f);
                    g.w = make_float4(W.x, W.y, W.z, 0.0f);
                    g.kd = kd;
                    sink.gather(g);
                } else {
                    dir.x = W.x * (0.0f * kd.x);
                    dir.y = W.y * (0.0f * kd.y);
                    dir.z = W.z * (0.0f * kd.z);
                }
                sink.set_direct(dir);
            } else {
                float refl_fact = (mat.flags & SRAY_MAT_HAS_REFLECT) ? mat.reflect : 0.0f;
                float trans_fact = (mat.flags & SRAY_MAT_HAS_REFRACT) ? mat.refract : 0.0f;
                float mat_ior = mat.ior;
                if (mapped) {
                    if (mat.flags & SRAY_MAT_HAS_REFLECT)
                        refl_fact = mat_value(S, mat.map[SRAY_MAP_REFLECT], refl_fact, tu, tv);
                    if (mat.flags & SRAY_MAT_HAS_REFRACT)
                        trans_fact = mat_value(S, mat.map[SRAY_MAP_REFRACT], trans_fact, tu, tv);
                    mat_ior = mat_value(S, mat.map[SRAY_MAP_IOR], mat_ior, tu, tv);
                }
                IorStack st;
                st.cur = R.w.w;
                st.s[0] = R.stk.x;
                st.s[1] = R.stk.y;
                st.s[2] = R.stk.z;
                st.s[3] = R.stk.w;
                st.n = R.m.w & 0xffu;
                st.overflow = false;
                float ior = st.calc(entering, mat_ior);
                /* a "normal" attribute replaces the flipped normal, src/shader.cc:137-138:
                 * by the mapped normal, or without a map by the unflipped one */
                if (has_normal) normal = bump ? bump_normal(S, mat.map[SRAY_MAP_NORMAL], mat.nscale, N, G, tu, tv) : N;
                /* ambient: (env_ambient * kd) * kd, src/shader.cc:142,246 */
                dir.x = W.x * ((S.env_ambient.x * kd.x) * kd.x);
                dir.y = W.y * ((S.env_ambient.y * kd.y) * kd.y);
                dir.z = W.z * ((S.env_ambient.z * kd.z) * kd.z);
                int slotk = 0;
                V3 vref = reflect(incident, normal);
                for (int l = 0; l < nl; l++) {
                    const DevLight &lt = S.lights[l];
                    int ns = lt.type != SRAY_LIGHT_POINT ? P.shadow_samples : 1;
                    float scale = ns > 1 ? 1.0f / (float)ns : 1.0f;
                    for (int si = 0; si < ns; si++) {
                        V3 lpos = light_point(lt, P.seed, pix, (uint32_t)sample, path, (uint32_t)l, (uint32_t)si,
                                              time);
                        V3 ldir = sub(lpos, Pw);
                        float ldist = length(ldir);
                        V3 sd = ldir;
                        ldir = dvs(ldir, ldist);
                        float ndotl = fmaxf(dot(normal, ldir), 0.0f);
                        float rdotl = fmaxf(dot(vref, ldir), 0.0f);
                        float sp = powf(rdotl, shininess);
                        SRay s;
                        s.o = make_float4(Pw.x, Pw.y, Pw.z, __int_as_float(time));
                        s.d = make_float4(sd.x, sd.y, sd.z, __uint_as_float(rq));
                        sink.shadow(slotk, s,
                                    make_float4(W.x * ((kd.x * (lt.color[0] * ndotl) + ks.x * (lt.color[0] * sp)) * scale),
                                                W.y * ((kd.y * (lt.color[1] * ndotl) + ks.y * (lt.color[1] * sp)) * scale),
                                                W.z * ((kd.z * (lt.color[2] * ndotl) + ks.z * (lt.color[2] * sp)) * scale),
                                                1.0f));
                        slotk++;
                        c_sh++;
                    }
                }
                used = slotk;
                if (PHOTONS && P.gi) {
                    /* diffuse hemisphere rays (src/shader.cc:158-185); a
                     * direction flipped into the hemisphere keeps its
                     * negative ndotl and so spawns no ray */
                    float dsc = P.diffuse_samples > 1 ? 1.0f / (float)P.diffuse_samples : 1.0f;
                    for (int di = 0; di < P.diffuse_samples; di++) {
                        V3 gd = sph_rand(P.seed, pix, (uint32_t)sample, path, K_DIFFUSE, (uint32_t)di, 0u, 1.0f);
                        float ndotl = dot(gd, normal);
                        if (ndotl < 0.0f) gd = neg(gd);
                        float diff_energy = ndotl * energy;
                        if (diff_energy > P.min_energy) {
                            QRay cr;
                            V3 dd = mul(gd, ray_mag);
                            cr.o = make_float4(Pw.x, Pw.y, Pw.z, R.o.w);
                            cr.d = make_float4(dd.x, dd.y, dd.z, diff_energy);
                            float wf = ndotl * dsc;
                            cr.w = make_float4(W.x * (kd.x * wf), W.y * (kd.y * wf), W.z * (kd.z * wf), R.w.w);
                            cr.m = make_uint4(slot, path | 0x40000000u | ((uint32_t)di << 20), 0x7fffffffu,
                                              (R.m.w & 0xffu) | (QF_DIFFUSE << 8) | (rq << SLOT_SHIFT));
                            cr.stk = R.stk;
                            sink.child(2 + di, cr);
                            c_df++;
                        }
                        sink.flag(2 + di, diff_energy > P.min_energy ? 1 : 0);
                    }
                    dwritten = P.diffuse_samples;
                }
                if (PHOTONS && P.caust) {
                    /* caustics: irradiance from the caustic map (src/shader.cc:231-236),
                     * added by k_gather */
                    GatherReq g;
                    g.pos = make_float4(Pw.x, Pw.y, Pw.z, __uint_as_float(1u));
                    g.nrm = make_float4(normal.x, normal.y, normal.z, 0.0f);
                    g.w = make_float4(W.x, W.y, W.z, 0.0f);
                    g.kd = kd;
                    sink.gather(g);
                }
                V3 nrd = dvs(neg(rd), ray_mag);
                float ray_dot_n = dot(nrd, normal);
                float sf0 = trans_fact > 0.0f ? (ior - 1.0f) / (ior + 1.0f) : 1.0f;
                float fres = fresnel(sf0 * sf0, ray_dot_n);
                float refl_energy = refl_fact * energy;
                if (refl_energy > P.min_energy && iter > 0) {
                    V3 nd = reflect(rd, normal);
                    QRay c;
                    c.o = make_float4(Pw.x, Pw.y, Pw.z, R.o.w);
                    c.d = make_float4(nd.x, nd.y, nd.z, refl_energy);
                    float wf = fres * refl_fact;
                    c.w = make_float4(W.x * (ks.x * wf), W.y * (ks.y * wf), W.z * (ks.z * wf), R.w.w);
                    c.m = make_uint4(slot, path * 2u, (uint32_t)(iter - 1), R.m.w);
                    c.stk = R.stk;
                    sink.child(0, c);
                    spawn |= 1u;
                    c_rl++;
                }
                float trans_energy = trans_fact * energy;
                if (trans_energy > P.min_energy && iter > 0) {
                    IorStack ts = st;
                    if (entering) ts.enter(mat_ior);
                    else ts.leave();
                    V3 td = mul(refract(dvs(rd, ray_mag), normal, ior), ray_mag);
                    if (dot(td, normal) > 0.0f) {
                        if (entering) ts.leave();
                        else ts.enter(mat_ior);
                    }
                    if (ts.overflow) sink.ior_overflow();
                    QRay c;
                    c.o = make_float4(Pw.x, Pw.y, Pw.z, R.o.w);
                    c.d = make_float4(td.x, td.y, td.z, trans_energy);
                    float wf = (1.0f - fres) * trans_fact;
                    c.w = make_float4(W.x * (ks.x * wf), W.y * (ks.y * wf), W.z * (ks.z * wf), ts.cur);
                    c.m = make_uint4(slot, path * 2u + 1u, (uint32_t)(iter - 1), ts.n | (flg << 8) | (rq << SLOT_SHIFT));
                    c.stk = make_float4(ts.s[0], ts.s[1], ts.s[2], ts.s[3]);
                    sink.child(1, c);
                    spawn |= 2u;
                    c_rr++;
                }
                sink.set_direct(dir);
            }
        }
    }
    /* flags of the child slots and the unused shadow slots, written once */
    sink.flag(0, (uint8_t)(spawn & 1u));
    sink.flag(1, (uint8_t)((spawn >> 1) & 1u));
    for (int k = 2 + dwritten; k < K; k++) sink.flag(k, 0);
    for (int l = used; l < L; l++) sink.shadow_unused(l);
}

/* the level arrays of a step: ray i's outputs at its own offsets */
struct ArraySink {
    uint32_t i;
    int L, K;
    float4 *direct;
    SRay *sq;
    float4 *scon;
    uint32_t *rslot;
    QRay *kids;
    uint8_t *kflag;
    GatherReq *greq;
    uint32_t *err;
    __device__ void slot(uint32_t s) { rslot[i] = s; }
    __device__ void gather_clear() { greq[i].pos = make_float4(0.0f, 0.0f, 0.0f, __uint_as_float(0u)); }
    __device__ void gather(const GatherReq &g) { greq[i] = g; }
    __device__ void set_direct(float4 d) { direct[i] = d; }
    __device__ void shadow(int l, const SRay &s, float4 c)
    {
        sq[(size_t)i * L + l] = s;
        scon[(size_t)i * L + l] = c;
    }
    __device__ void shadow_unused(int l) { scon[(size_t)i * L + l] = make_float4(0.0f, 0.0f, 0.0f, 0.0f); }
    __device__ void child(int k, const QRay &c) { kids[(size_t)i * K + k] = c; }
    __device__ void flag(int k, uint8_t f) { kflag[(size_t)i * K + k] = f; }
    __device__ void ior_overflow() { atomicOr(err, 1u); }
};

/* shade_phong (src/shader.cc:83-247) / shade_undef (:36-71) for every hit
 * of a level step, emitting direct terms, shadow rays, children and
 * photon-gather requests into the level arrays */
template <bool PHOTONS, bool UNI, bool SCN1>
__global__ __launch_bounds__(256) void k_shade(Tabs T, const uint32_t *__restrict__ active, uint32_t c0,
                                               const QRay *__restrict__ q, const int4 *__restrict__ hits, StepN sn,
                                               int L, int K, float4 *__restrict__ direct, SRay *__restrict__ sq,
                                               float4 *__restrict__ scon, uint32_t *__restrict__ rslot,
                                               QRay *__restrict__ kids, uint8_t *__restrict__ kflag,
                                               GatherReq *__restrict__ greq, unsigned long long *__restrict__ stats,
                                               unsigned long long *__restrict__ slot_rays, uint32_t *__restrict__ err)
{
    __shared__ uint32_t bcnt[5];
    /* per-request ray counts of the block (many requests: one global atomic
     * per request and block instead of one per request and wave) */
    __shared__ uint32_t scnt[UNI ? 1 : MAX_SLOTS];
    if (threadIdx.x < 5) bcnt[threadIdx.x] = 0;
    if (!UNI)
        for (uint32_t k = threadIdx.x; k < MAX_SLOTS; k += blockDim.x) scnt[k] = 0;
    __syncthreads();
    const uint32_t n = step_n(sn);
    uint32_t c_sh = 0, c_rl = 0, c_rr = 0, c_df = 0;
    /* grid-stride over the step, block-uniform bases (wave operations below) */
    for (uint32_t gbase = blockIdx.x * blockDim.x; gbase < n; gbase += gridDim.x * blockDim.x) {
        const uint32_t i = gbase + threadIdx.x;
        const uint32_t before = c_sh + c_rl + c_rr + c_df;
        if (i < n) {
            ArraySink sink{i, L, K, direct, sq, scon, rslot, kids, kflag, greq, err};
            shade_ray<PHOTONS, UNI, SCN1, true>(T, active, c0, q[i], hits[i], L, K, sink, c_sh, c_rl, c_rr, c_df);
        }
        if (!UNI)
            wave_keyed_add(scnt, i < n ? q[i].m.w >> SLOT_SHIFT : 0u, c_sh + c_rl + c_rr + c_df - before, i < n);
    }
    if (c_sh) atomicAdd(&bcnt[0], c_sh);
    if (c_rl) atomicAdd(&bcnt[1], c_rl);
    if (c_rr) atomicAdd(&bcnt[2], c_rr);
    if (c_df) atomicAdd(&bcnt[3], c_df);
    __syncthreads();
    if (threadIdx.x < 4 && bcnt[threadIdx.x]) atomicAdd(&stats[1 + threadIdx.x], (unsigned long long)bcnt[threadIdx.x]);
    if (!UNI)
        for (uint32_t k = threadIdx.x; k < MAX_SLOTS; k += blockDim.x)
            if (scnt[k]) atomicAdd(&slot_rays[k], (unsigned long long)scnt[k]);
}

/* A lane's registers as the shading sink of the one-child chain: the direct
 * term, up to LMAX shadow rays with their contributions, and the one child
 * (every request in flight spawns at most one child per ray and has no
 * photon maps, so there are no gather requests and no diffuse children). */
/* one sample c of arena pixel pr added to its sum f in float
 * (src/render.cc:226-229), its record kept for the variance stop test
 * (:232-237, 261-270); false when the pixel stops after it */
__device__ __forceinline__ bool add_sample(const ReqDev &Q, uint32_t pr, int sample, float4 c, float4 &f,
                                           float4 *__restrict__ subpix)
{
    const RenderDev &P = Q.P;
    const int maxs = P.max_samples;
    f.x = f.x + c.x;
    f.y = f.y + c.y;
    f.z = f.z + c.z;
    f.w = f.w + c.w;
    int ns = sample + 1;
    bool k = ns < maxs;
    if (maxs > 1) {
        float4 *sp = subpix + Q.sub_base + (size_t)(pr - Q.pix_base) * (size_t)maxs;
        sp[sample] = c;
        if (ns >= P.min_samples && ns > 1) {
            float rcp = 1.0f / (float)ns;
            float4 mean = make_float4(f.x * rcp, f.y * rcp, f.z * rcp, f.w * rcp);
            float sq = 0.0f;
            for (int j = 0; j < ns; j++) {
                float4 s = sp[j];
                float dx = mean.x - s.x, dy = mean.y - s.y, dz = mean.z - s.z, dw = mean.w - s.w;
                float l2 = dx * dx + dy * dy + dz * dz + dw * dw;
                sq = sq + l2;
            }
            float v = sq * rcp;
            if (v < P.max_var) k = false;
        }
    }
    return k;
}

template <int LMAX> struct ChainSink {
    float4 *acc;   /* the sample's sum: the direct term is added when set */
    SRay sh[LMAX];
    float4 con[LMAX];
    QRay *kid;     /* the child, in LDS (it would hold 20 registers through the shading) */
    uint32_t kids; /* bit k: child k spawned */
    uint32_t *err;
    __device__ void slot(uint32_t) {}
    __device__ void gather_clear() {}
    __device__ void gather(const GatherReq &) {}
    __device__ void set_direct(float4 d)
    {
        acc->x = acc->x + d.x;
        acc->y = acc->y + d.y;
        acc->z = acc->z + d.z;
        acc->w = acc->w + d.w;
    }
    /* one slot: its index is 0 (L <= 1), which keeps the arrays in registers */
    __device__ void shadow(int l, const SRay &s, float4 c)
    {
        sh[LMAX == 1 ? 0 : l] = s;
        con[LMAX == 1 ? 0 : l] = c;
    }
    __device__ void shadow_unused(int l) { con[LMAX == 1 ? 0 : l] = make_float4(0.0f, 0.0f, 0.0f, 0.0f); }
    __device__ void child(int, const QRay &c) { *kid = c; }
    __device__ void flag(int k, uint8_t f) { kids |= f ? 1u << k : 0u; }
    __device__ void ior_overflow() { atomicOr(err, 1u); }
};

/* The one-child chain: when no ray of the requests in flight can spawn more
 * than one child and none has photon maps, a sample's rays form a single
 * line, and one lane follows it from the primary ray to its end: closest
 * hit, shading, the shadow rays, and the sample's sum.  Each ray's
 * arithmetic is that of the level kernels (k_raygen, k_trace_q, k_shade,
 * k_trace_shadow, k_accum), and with one ray per sample per level the sum
 * is added in the same order, so the image is the same; there are no level
 * queues, compactions or per-level launches.  acc[i] receives sample i's
 * sum for k_pass_finish. */
template <bool UNI, bool SCN1, int LMAX, bool ATTR, int BS>
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(ATTR || LMAX > 1 ? 1 : 4))) void k_chain(Tabs T, const uint32_t *__restrict__ active, uint32_t c0, uint32_t m,
                                              int L, int K, float4 *__restrict__ acc,
                                              unsigned long long *__restrict__ stats,
                                              unsigned long long *__restrict__ slot_rays, uint32_t *__restrict__ err)
{
    /* 20 KB a block of two waves: four waves per SIMD */
    __shared__ uint32_t stk[STACK * BS];
    __shared__ QRay kids[BS];
    uint32_t c_sh = 0, c_rl = 0, c_rr = 0, c_df = 0;
    TraceCounts cnt = {0, 0, 0, 0, 0, 0, 0};
    /* this lane's rays of its last request, added to the block's counts at
     * the end (the grid has a lane per sample, so that is all of them) */
    uint32_t last_rq = 0, last_cnt = 0;
    for (uint32_t gbase = blockIdx.x * (uint32_t)BS; gbase < m; gbase += gridDim.x * (uint32_t)BS) {
        const uint32_t i = gbase + threadIdx.x;
        const uint32_t before = c_sh + c_rl + c_rr + c_df;
        uint32_t rq = 0;
        if (i < m) {
            /* k_raygen */
            const uint32_t e = active[c0 + i], pr = ent_pix(e);
            rq = UNI ? 0u : T.preq[pr];
            const ReqDev &Q = T.reqs[rq];
            const RenderDev &P = Q.P;
            const uint32_t lp = pr - Q.pix_base;
            const int x = P.x0 + (int)(lp % (uint32_t)P.W), y = P.y0 + (int)(lp / (uint32_t)P.W);
            V3 o, d;
            int time;
            primary_ray(Q.C, P, x, y, T.nsamp[pr] + (int)ent_off(e), o, d, time);
            QRay R;
            R.o = make_float4(o.x, o.y, o.z, __int_as_float(time));
            R.d = make_float4(d.x, d.y, d.z, 1.0f);
            R.w = make_float4(1.0f, 1.0f, 1.0f, 1.0f);
            R.m = make_uint4(i, 1u, (uint32_t)P.iter, rq << SLOT_SHIFT);
            R.stk = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            const SceneDev S = T.scenes[(UNI || SCN1) ? 0u : Q.scene];
            float4 a = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            for (;;) {
                /* k_trace_q */
                const HitRec hr = traverse<false, false>(S, v3(R.o.x, R.o.y, R.o.z), v3(R.d.x, R.d.y, R.d.z),
                                                         __float_as_int(R.o.w), &stk[threadIdx.x], BS, cnt);
                const int4 h = make_int4(__float_as_int(hr.t), hr.obj, (int)hr.tri, (int)hr.rank);
                /* k_shade */
                ChainSink<LMAX> sk;
                sk.acc = &a;
                sk.kid = &kids[threadIdx.x];
                sk.kids = 0u;
                sk.err = err;
                /* k_shade, with k_accum's first addition (the direct term) */
                shade_ray<false, UNI, SCN1, ATTR>(T, active, c0, R, h, L, K, sk, c_sh, c_rl, c_rr, c_df);
                /* k_trace_shadow and the rest of k_accum (accum_run's order) */
#pragma unroll
                for (int l = 0; l < LMAX; l++) {
                    const float4 cc = sk.con[l];
                    if (l < L && cc.w != 0.0f) {
                        const SRay &sr = sk.sh[l];
                        const HitRec oc = traverse<true, false>(S, v3(sr.o.x, sr.o.y, sr.o.z), v3(sr.d.x, sr.d.y, sr.d.z),
                                                                __float_as_int(sr.o.w), &stk[threadIdx.x], BS, cnt);
                        if (oc.obj < 0) {
                            a.x = a.x + cc.x;
                            a.y = a.y + cc.y;
                            a.z = a.z + cc.z;
                        }
                    }
                }
                if (!sk.kids) break;
                R = kids[threadIdx.x];
            }
            acc[i] = a;
        }
        if (!UNI && i < m) {
            const uint32_t add = c_sh + c_rl + c_rr + c_df - before;
            if (last_cnt && rq != last_rq) atomicAdd(&slot_rays[last_rq], (unsigned long long)last_cnt);
            if (rq != last_rq) last_cnt = 0;
            last_rq = rq;
            last_cnt += add;
        }
    }
    /* block sums of the spawn counts and of the per-request counts, in the
     * stack's LDS (no walk is left) */
    __syncthreads();
    uint32_t *bcnt = stk, *scnt = stk + 4;
    static_assert(STACK * BS >= 4 + MAX_SLOTS, "the block sums fit in the stack's LDS");
    for (uint32_t k = threadIdx.x; k < (UNI ? 4u : 4u + MAX_SLOTS); k += BS) stk[k] = 0;
    __syncthreads();
    if (c_sh) atomicAdd(&bcnt[0], c_sh);
    if (c_rl) atomicAdd(&bcnt[1], c_rl);
    if (c_rr) atomicAdd(&bcnt[2], c_rr);
    if (c_df) atomicAdd(&bcnt[3], c_df);
    if (!UNI) wave_keyed_add(scnt, last_rq, last_cnt, last_cnt != 0u);
    __syncthreads();
    if (threadIdx.x < 4 && bcnt[threadIdx.x]) atomicAdd(&stats[1 + threadIdx.x], (unsigned long long)bcnt[threadIdx.x]);
    if (!UNI)
        for (uint32_t k = threadIdx.x; k < MAX_SLOTS; k += BS)
            if (scnt[k]) atomicAdd(&slot_rays[k], (unsigned long long)scnt[k]);
}

/* photon gathers requested by k_shade: caustic irradiance (>= 8 photons,
 * src/pmap.cc:109-139) added to the direct term, or the radiance estimate
 * (src/pmap.cc:141-178) of a diffuse GI ray times kd */
/* the estimate of request i into its direct term: mode 1 adds
 * W * caustic irradiance (no estimate from fewer than 8 photons), mode 2
 * sets W * (GI radiance * kd) */
__device__ __forceinline__ void gather_finish(const GatherReq *__restrict__ greq, uint32_t i, uint32_t mode,
                                              const ShadePmaps &PM, const GatherAcc &a, float4 *__restrict__ direct)
{
    float4 w = greq[i].w;
    float4 d = direct[i];
    if (mode == 1u) {
        V3 irr = a.count < 8 ? v3(0.0f, 0.0f, 0.0f)
                             : v3((float)(a.r * PM.scale), (float)(a.g * PM.scale), (float)(a.b * PM.scale));
        d.x = d.x + w.x * irr.x;
        d.y = d.y + w.y * irr.y;
        d.z = d.z + w.z * irr.z;
    } else {
        V3 rad = v3((float)(a.r * PM.scale), (float)(a.g * PM.scale), (float)(a.b * PM.scale));
        float4 kd = greq[i].kd;
        d.x = w.x * (rad.x * kd.x);
        d.y = w.y * (rad.y * kd.y);
        d.z = w.z * (rad.z * kd.z);
    }
    direct[i] = d;
}

/* KNN: a request with gather_k > 0 whose range holds more than k photons
 * is listed in knn (count at knn[0], then (index, photons in range)
 * pairs) for k_gather_knn, which writes its estimate */
template <bool UNI, int GATHER_LANES, bool KNN>
__device__ __forceinline__ void gather_one(Tabs T, const QRay *__restrict__ q, const GatherReq *__restrict__ greq,
                                           uint32_t i, int lane, float4 *__restrict__ direct, uint32_t *__restrict__ knn)
{
    float4 pos = greq[i].pos;
    uint32_t mode = __float_as_uint(pos.w);
    if (mode == 0u) return;
    uint32_t rq = UNI ? 0u : q[i].m.w >> SLOT_SHIFT;
    const ShadePmaps &PM = T.reqs[rq].PM;
    float4 nn = greq[i].nrm;
    GatherAcc a;
    gather_grid_coop<GATHER_LANES>(mode == 1u ? PM.caust : PM.glob, pos.x, pos.y, pos.z, PM.radius,
                                   mode == 1u ? SRAY_GATHER_IRRADIANCE : SRAY_GATHER_RADIANCE, (double)nn.x,
                                   (double)nn.y, (double)nn.z, lane, a);
    if (lane != 0) return;
    if (KNN && PM.k > 0 && a.count > (uint32_t)PM.k) {
        const uint32_t j = atomicAdd(knn, 1u);
        knn[1u + 2u * j] = i;
        knn[2u + 2u * j] = a.count;
        return;
    }
    gather_finish(greq, i, mode, PM, a, direct);
}

/* KNN: the instance for passes where a request gathers the k nearest */
template <bool UNI, int GATHER_LANES, bool KNN>
__global__ __launch_bounds__(256) void k_gather(Tabs T, const QRay *__restrict__ q, const GatherReq *__restrict__ greq,
                                                StepN sn, float4 *__restrict__ direct, uint32_t *__restrict__ knn)
{
    const uint32_t n = step_n(sn);
    const uint32_t groups = gridDim.x * (blockDim.x / GATHER_LANES);
    const int lane = (int)(threadIdx.x % GATHER_LANES);
    /* grid-stride over the requests, one group of lanes each (whole groups
     * move and leave together) */
    for (uint32_t i = (blockIdx.x * blockDim.x + threadIdx.x) / GATHER_LANES; i < n; i += groups)
        gather_one<UNI, GATHER_LANES, KNN>(T, q, greq, i, lane, direct, knn);
}

/* the requests k_gather listed: the estimate over the k nearest photons
 * (gather_grid_kcap), a group of lanes each.  The k nearest are looked
 * for first within r sqrt(1.5 k / n), n the photons within r (the radius
 * that holds 1.5 k photons where their density is even, photons lying on
 * surfaces), then, if that holds fewer than k, within r. */
template <bool UNI, int GATHER_LANES>
__global__ __launch_bounds__(256) void k_gather_knn(Tabs T, const QRay *__restrict__ q,
                                                    const GatherReq *__restrict__ greq,
                                                    const uint32_t *__restrict__ knn, float4 *__restrict__ direct)
{
    __shared__ KcapLds<GATHER_LANES> lds[256 / GATHER_LANES];
    const uint32_t n = knn[0];
    const uint32_t groups = gridDim.x * (blockDim.x / GATHER_LANES);
    const int lane = (int)(threadIdx.x % GATHER_LANES);
    KcapLds<GATHER_LANES> &L = lds[threadIdx.x / GATHER_LANES];
    for (uint32_t j = (blockIdx.x * blockDim.x + threadIdx.x) / GATHER_LANES; j < n; j += groups) {
        const uint32_t i = knn[1u + 2u * j], inr = knn[2u + 2u * j];
        float4 pos = greq[i].pos;
        uint32_t mode = __float_as_uint(pos.w);
        uint32_t rq = UNI ? 0u : q[i].m.w >> SLOT_SHIFT;
        const ShadePmaps &PM = T.reqs[rq].PM;
        float4 nn = greq[i].nrm;
        const PmapGrid &Gd = mode == 1u ? PM.caust : PM.glob;
        const int gm = mode == 1u ? SRAY_GATHER_IRRADIANCE : SRAY_GATHER_RADIANCE;
        GatherAcc a;
        const float f = 1.5f * (float)PM.k / (float)inr;
        bool done = false;
        if (f < 0.8f)
            done = gather_grid_kcap<GATHER_LANES, false>(Gd, pos.x, pos.y, pos.z, PM.radius * sqrtf(f), gm,
                                                         (double)nn.x, (double)nn.y, (double)nn.z, lane,
                                                         (uint32_t)PM.k, L, a);
        if (!done)
            gather_grid_kcap<GATHER_LANES, false>(Gd, pos.x, pos.y, pos.z, PM.radius, gm, (double)nn.x, (double)nn.y,
                                                  (double)nn.z, lane, (uint32_t)PM.k, L, a);
        if (lane == 0) gather_finish(greq, i, mode, PM, a, direct);
    }
}

/* the sum of the run of equal sample slots starting at i, if i starts one */
__device__ __forceinline__ void accum_run(const uint32_t *__restrict__ rslot, const float4 *__restrict__ direct,
                                          const float4 *__restrict__ scon, const uint8_t *__restrict__ occ,
                                          uint32_t n, int L, float4 *__restrict__ acc, uint32_t i)
{
    uint32_t slot = rslot[i];
    if (i > 0 && rslot[i - 1] == slot) return;
    float4 s = acc[slot];
    for (uint32_t j = i; j < n && rslot[j] == slot; j++) {
        float4 d = direct[j];
        s.x = s.x + d.x;
        s.y = s.y + d.y;
        s.z = s.z + d.z;
        s.w = s.w + d.w;
        for (int l = 0; l < L; l++) {
            size_t k = (size_t)j * L + l;
            float4 c = scon[k];
            if (c.w != 0.0f && !occ[k]) {
                s.x = s.x + c.x;
                s.y = s.y + c.y;
                s.z = s.z + c.z;
            }
        }
    }
    acc[slot] = s;
}

/* per-sample accumulation: the first ray of each run of equal sample slots
 * sums the run (queues are ordered by slot) */
__global__ void k_accum(const uint32_t *__restrict__ rslot, const float4 *__restrict__ direct,
                        const float4 *__restrict__ scon, const uint8_t *__restrict__ occ, StepN sn, int L,
                        float4 *__restrict__ acc, uint32_t *__restrict__ knn)
{
    /* the step's k-nearest list is spent: empty it for the next step */
    if (knn && blockIdx.x == 0 && threadIdx.x == 0) knn[0] = 0u;
    const uint32_t n = step_n(sn);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        accum_run(rslot, direct, scon, occ, n, L, acc, i);
}

/* ------------------------------------------------------------------ */
/* stable compaction of a level step's flagged children into the next   */
/* level's queue, with the count read and written on the device         */
/* ------------------------------------------------------------------ */
/* pass 1: flagged children per block (devsort.h) */
__global__ __launch_bounds__(CMP_THREADS) void k_cmp_count(const uint8_t *__restrict__ flags, StepN sn, int K,
                                                           uint32_t *__restrict__ bcount)
{
    cmp_block_count(flags, step_n(sn) * (uint32_t)K, bcount);
}

/* pass 2: each block writes its flagged children in order from its
 * offset; the last block writes the total, the next level's count */
__global__ __launch_bounds__(CMP_THREADS) void k_cmp_scatter(const QRay *__restrict__ in,
                                                             const uint8_t *__restrict__ flags, StepN sn, int K,
                                                             const uint32_t *__restrict__ bcount,
                                                             uint32_t *__restrict__ total, QRay *__restrict__ out)
{
    cmp_block_scatter(in, flags, step_n(sn) * (uint32_t)K, bcount, total, out);
}

__global__ void k_fill_f4(float4 *p, size_t n, float4 v)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_iota(uint32_t *p, uint32_t n, uint32_t base)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = base + i;
}

/* a request's admission in one launch: its arena pixels cleared (sum,
 * samples taken, request slot) and its active-list entries, the pixels of
 * a W x H rectangle in tile order: tiles of tw x th pixels (tw * th = 64,
 * one wave), rows of tiles top to bottom, tiles left to right, row-major
 * inside a tile (tw = 64: plain row order), run entries per pixel (sample
 * offsets 0..run-1) */
__global__ void k_admit(uint32_t *__restrict__ act, float4 *__restrict__ fb, uint32_t *__restrict__ nsamp,
                        uint32_t *__restrict__ preq, uint32_t npix, uint32_t base, uint32_t W, uint32_t H,
                        uint32_t tw, uint32_t run, uint32_t slot)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < npix) {
        fb[t] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        nsamp[t] = 0u;
        preq[t] = slot;
    }
    if (t >= npix * run) return;
    const uint32_t i = t / run, off = t - i * run;
    act += (size_t)i * (run - 1u) + off;
    uint32_t th = 64u / tw;
    uint32_t ty = i / (th * W);
    uint32_t o = i - ty * th * W;
    uint32_t hr = min(th, H - th * ty);
    uint32_t tx = o / (tw * hr);
    uint32_t wt = min(tw, W - tw * tx);
    uint32_t j = o - tx * tw * hr;
    uint32_t y = th * ty + j / wt, x = tw * tx + j % wt;
    act[i] = (base + y * W + x) | (off << ENT_SHIFT);
}

__global__ void k_fill_u32(uint32_t *p, size_t n, uint32_t v)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

/* active pixels per request slot over a grid-stride loop: counted per
 * block in LDS, then one global atomic per request and block */
__device__ __forceinline__ void slot_count(const uint32_t *__restrict__ active, uint32_t n,
                                           const uint32_t *__restrict__ preq, uint32_t *__restrict__ cnt)
{
    __shared__ uint32_t scnt[MAX_SLOTS];
    for (uint32_t k = threadIdx.x; k < MAX_SLOTS; k += blockDim.x) scnt[k] = 0;
    __syncthreads();
    for (uint32_t b = blockIdx.x * blockDim.x; b < n; b += gridDim.x * blockDim.x) {
        const uint32_t i = b + threadIdx.x;
        wave_keyed_add(scnt, i < n ? preq[ent_pix(active[i])] : 0u, 1u, i < n);
    }
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < MAX_SLOTS; k += blockDim.x)
        if (scnt[k]) atomicAdd(&cnt[k], scnt[k]);
}

/* active pixels per request slot, the list's length read on the device */
__global__ void k_slot_count_dev(const uint32_t *__restrict__ active, const uint32_t *__restrict__ np,
                                 const uint32_t *__restrict__ preq, uint32_t *__restrict__ cnt)
{
    slot_count(active, *np, preq, cnt);
}


/* end of one sample pass: each active pixel's sample added (add_sample).
 * A pixel whose request takes several samples in its first pass has them
 * in consecutive entries of the chunk: its first entry adds them one by
 * one, in sample order, with the stop test after each (the others only
 * leave the list), so the sums and the stop are those of one sample per
 * pass. */
template <bool UNI>
__global__ void k_pass_finish(Tabs T, const uint32_t *__restrict__ active, uint32_t c0, uint32_t m,
                              const float4 *__restrict__ acc, float4 *__restrict__ fb, float4 *__restrict__ subpix,
                              int *__restrict__ nsamp, uint8_t *__restrict__ keep)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const uint32_t e = active[c0 + i], pr = ent_pix(e);
    if (ent_off(e) != 0u) {
        keep[c0 + i] = 0;
        return;
    }
    uint32_t rq = UNI ? 0u : T.preq[pr];
    const ReqDev &Q = T.reqs[rq];
    int sample = nsamp[pr];
    float4 f = fb[pr];
    bool k = true;
    /* the first pass's run (every sample of it is taken, see admit) */
    const uint32_t run = sample == 0 ? Q.run : 1u;
    for (uint32_t o = 0; o < run && k; o++) {
        k = add_sample(Q, pr, sample, acc[i + o], f, subpix);
        sample++;
    }
    fb[pr] = f;
    nsamp[pr] = sample;
    keep[c0 + i] = k ? 1 : 0;
}

/* final pixel value: divide all channels by the sample count, or the
 * reference's as-written img[x] *= rcp_lut[i] (src/render.cc:240, Q1) */
__global__ void k_finalize(RenderDev P, const float4 *__restrict__ fb, const int *__restrict__ nsamp,
                           float *__restrict__ out, uint32_t npix)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npix) return;
    float4 f = fb[i];
    if (P.flags & SRAY_RENDER_REF_Q1) {
        out[(size_t)i * 4 + 0] = f.x;
        out[(size_t)i * 4 + 1] = f.y;
        out[(size_t)i * 4 + 2] = f.z;
        out[(size_t)i * 4 + 3] = f.w;
        return;
    }
    float rcp = 1.0f / (float)nsamp[i];
    out[(size_t)i * 4 + 0] = f.x * rcp;
    out[(size_t)i * 4 + 1] = f.y * rcp;
    out[(size_t)i * 4 + 2] = f.z * rcp;
    out[(size_t)i * 4 + 3] = f.w * rcp;
}

/* the images of several finished requests in one launch (k_finalize's
 * division, for requests without the Q1 pattern): job k covers pixels
 * [start_k, start_k + npix_k) of the launch.  The table is the kernel's
 * argument (FIN_MAX jobs keep it under the 4 KB argument limit), so no
 * staging buffer outlives the launch call. */
struct FinJob {
    const float4 *fb;
    const int *nsamp;
    float *dst;
    uint32_t npix, start;
};
constexpr uint32_t FIN_MAX = 112;
struct FinBatch {
    uint32_t nj, total;
    FinJob j[FIN_MAX];
};
__global__ void k_finalize_many(const FinBatch B)
{
    for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < B.total; g += gridDim.x * blockDim.x) {
        uint32_t lo = 0, hi = B.nj; /* j[lo].start <= g < j[hi].start (or the end) */
        while (hi - lo > 1u) {
            const uint32_t mid = (lo + hi) >> 1;
            if (B.j[mid].start <= g) lo = mid;
            else hi = mid;
        }
        const FinJob &J = B.j[lo];
        const uint32_t i = g - J.start;
        const float4 f = J.fb[i];
        const float rcp = 1.0f / (float)J.nsamp[i];
        J.dst[(size_t)i * 4 + 0] = f.x * rcp;
        J.dst[(size_t)i * 4 + 1] = f.y * rcp;
        J.dst[(size_t)i * 4 + 2] = f.z * rcp;
        J.dst[(size_t)i * 4 + 3] = f.w * rcp;
    }
}

__global__ void k_q1_scale(RenderDev P, const int *__restrict__ nsamp, float *__restrict__ out, uint32_t npix)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npix) return;
    int xr = (int)(i % (uint32_t)P.W), yr = (int)(i / (uint32_t)P.W);
    int ax = P.x0 + xr;
    int tx = (ax / 48) * 48;
    int org = tx > P.x0 ? tx : P.x0; /* tile origin clipped to the rectangle */
    int f = ax - org;                /* pixel offset within the tile row */
    size_t fi = ((size_t)yr * P.W + (size_t)(org - P.x0)) * 4 + (size_t)f;
    float rcp = 1.0f / (float)nsamp[i];
    out[fi] = out[fi] * rcp;
}


/* launchable instances of the tunable kernels */
/* launchable instances of the tunable kernels (the SPARSE instances hold
 * more registers: they are launched for the deep, sparse bounce levels) */
template <bool COUNT, bool UNI, bool P, bool SP>
auto trace_q_of_p(int bs) -> decltype(&k_trace_q<COUNT, UNI, 128, P, SP>)
{
    return bs == 64    ? k_trace_q<COUNT, UNI, 64, P, SP>
           : bs == 256 ? k_trace_q<COUNT, UNI, 256, P, SP>
                       : k_trace_q<COUNT, UNI, 128, P, SP>;
}
template <bool COUNT, bool UNI>
auto trace_q_of(int bs, bool persist, bool sparse) -> decltype(&k_trace_q<COUNT, UNI, 128, false, false>)
{
    if constexpr (!COUNT) {
        if (sparse) return persist ? trace_q_of_p<COUNT, UNI, true, true>(bs) : trace_q_of_p<COUNT, UNI, false, true>(bs);
    }
    return persist ? trace_q_of_p<COUNT, UNI, true, false>(bs) : trace_q_of_p<COUNT, UNI, false, false>(bs);
}
template <bool COUNT, bool UNI, bool P, bool SP>
auto trace_shadow_of_p(int bs) -> decltype(&k_trace_shadow<COUNT, UNI, 128, P, SP>)
{
    return bs == 64    ? k_trace_shadow<COUNT, UNI, 64, P, SP>
           : bs == 256 ? k_trace_shadow<COUNT, UNI, 256, P, SP>
                       : k_trace_shadow<COUNT, UNI, 128, P, SP>;
}
template <bool COUNT, bool UNI>
auto trace_shadow_of(int bs, bool persist, bool sparse) -> decltype(&k_trace_shadow<COUNT, UNI, 128, false, false>)
{
    if constexpr (!COUNT) {
        if (sparse)
            return persist ? trace_shadow_of_p<COUNT, UNI, true, true>(bs) : trace_shadow_of_p<COUNT, UNI, false, true>(bs);
    }
    return persist ? trace_shadow_of_p<COUNT, UNI, true, false>(bs) : trace_shadow_of_p<COUNT, UNI, false, false>(bs);
}
template <bool UNI> auto gather_knn_of(int lanes) -> decltype(&k_gather_knn<UNI, 16>)
{
    return lanes == 8 ? k_gather_knn<UNI, 8> : lanes == 32 ? k_gather_knn<UNI, 32> : k_gather_knn<UNI, 16>;
}
template <bool UNI, bool KNN> auto gather_of(int lanes) -> decltype(&k_gather<UNI, 16, KNN>)
{
    return lanes == 8 ? k_gather<UNI, 8, KNN> : lanes == 32 ? k_gather<UNI, 32, KNN> : k_gather<UNI, 16, KNN>;
}

} // namespace

/* ================================================================== */
/* host side                                                            */
/* ================================================================== */
namespace srt {

constexpr uint32_t MAX_CHUNK = 1u << 21; /* largest rays per level sub-chunk */
using Clock = std::chrono::steady_clock;

void fill_render_dev(const sray_render_params *p, RenderDev &R)
{
    memset(&R, 0, sizeof R);
    R.width = p->width;
    R.height = p->height;
    R.x0 = p->x0;
    R.y0 = p->y0;
    R.x1 = p->x1;
    R.y1 = p->y1;
    R.W = p->x1 - p->x0;
    R.min_samples = p->min_samples;
    R.max_samples = std::max(1, p->max_samples);
    R.max_var = p->max_var;
    R.min_energy = p->min_energy;
    R.iter = p->iter;
    R.shadow_samples = p->shadow_samples; /* 0 or less: area lights take no samples, as the reference */
    R.diffuse_samples = std::max(0, p->diffuse_samples);
    R.frame_time = p->frame_time;
    R.mblur = p->mblur;
    R.caust = p->caust_photons;
    R.gi = p->gi_photons;
    R.flags = p->flags;
    R.seed = p->seed;
    R.aspect = (float)((double)p->width / (double)p->height);
}

int validate_params(sray_gpu_ctx *c, const sray_render_params *p)
{
    if (!p) return fail(c, SRAY_ERR_ARG, "render params is NULL");
    if (p->width <= 0 || p->height <= 0) return fail(c, SRAY_ERR_ARG, "bad image size %dx%d", p->width, p->height);
    if (p->x0 < 0 || p->y0 < 0 || p->x1 > p->width || p->y1 > p->height || p->x0 >= p->x1 || p->y0 >= p->y1)
        return fail(c, SRAY_ERR_ARG, "bad pixel rectangle");
    if (p->max_samples > 256) return fail(c, SRAY_ERR_LIMIT, "max_samples > 256");
    if (p->iter < 0 || p->iter > 30) return fail(c, SRAY_ERR_LIMIT, "iter outside [0,30]");
    if (p->shadow_samples > 1024) return fail(c, SRAY_ERR_LIMIT, "shadow_samples > 1024");
    if (p->diffuse_samples > 1023 || p->diffuse_samples < 0) return fail(c, SRAY_ERR_LIMIT, "diffuse_samples outside [0,1023]");
    if ((p->gi_photons || p->caust_photons) && !(p->gather_dist > 0.0f))
        return fail(c, SRAY_ERR_ARG, "gather_dist must be > 0 with photon maps");
    return SRAY_OK;
}


/* first-fit allocator of [base, base+len) ranges of an arena */
struct RangeAlloc {
    uint64_t cap = 0;
    std::map<uint64_t, uint64_t> fr; /* free ranges: base -> len */
    bool alloc(uint64_t n, uint64_t &base)
    {
        for (auto it = fr.begin(); it != fr.end(); ++it) {
            if (it->second < n) continue;
            base = it->first;
            uint64_t len = it->second;
            fr.erase(it);
            if (len > n) fr[base + n] = len - n;
            return true;
        }
        return false;
    }
    void release(uint64_t base, uint64_t n)
    {
        if (n == 0) return;
        auto it = fr.emplace(base, n).first;
        auto nx = std::next(it);
        if (nx != fr.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            fr.erase(nx);
        }
        if (it != fr.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                fr.erase(it);
            }
        }
    }
    void grow(uint64_t ncap)
    {
        if (ncap > cap) release(cap, ncap - cap);
        cap = ncap;
    }
    /* largest free range touching the end of the arena */
    uint64_t tail_free() const
    {
        if (fr.empty()) return 0;
        auto it = std::prev(fr.end());
        return it->first + it->second == cap ? it->second : 0;
    }
};

struct Job {
    uint64_t ticket = 0;
    uint32_t scene_id = 0;
    sray_render_params p{};
    sray_camera_desc cam{};
    bool use_cam = false;
    float *host_out = nullptr; /* result to host memory (through stage), or */
    float *dev_out = nullptr;  /* to device memory */
    DevBuf stage;
    uint32_t slot = 0;
    uint64_t pix_base = 0, npix = 0, sub_base = 0, nsub = 0;
    uint64_t rays = 0, prim = 0; /* ray queries of all kinds / primary rays */
    uint32_t pmaps[2] = {0, 0};  /* photon maps read while in flight */
    int state = 0; /* 0 waiting, 1 in flight, 2 done */
    int rc = SRAY_OK;
    std::string err;
    bool limit_hit = false;
    Clock::time_point t_submit, t_admit;
    double queue_ms = 0, latency_ms = 0;
};

struct Engine {
    sray_gpu_ctx *c = nullptr;
    /* pixel arena */
    RangeAlloc pix, sub;
    DevBuf fb, nsamp, preq, keep, act[2], subpix;
    int cur = 0;
    uint32_t nact = 0;
    /* shared counters and tables */
    DevBuf stats, errf, rcounts, slotcnt, slotrays, dreqs, dscenes;
    /* lanes: a stream each, with the ray queues and per-level scratch of the
     * primary chunks it runs (chunks on different lanes overlap on the GPU,
     * so one chunk's long-ray tail at a deep level leaves room for another
     * chunk's work) */
    struct Lane {
        hipStream_t s = nullptr; /* lane 0: the context's stream; others: created here */
        hipEvent_t done = nullptr;
        /* second stream: a level's shadow rays and accumulation run there,
         * beside the next level's compaction and closest-hit trace (evA:
         * the level is shaded; evB[p]: the accumulation of the last level of
         * parity p is done) */
        hipStream_t s2 = nullptr;
        hipEvent_t evA = nullptr, evB[2] = {nullptr, nullptr};
        /* the shading outputs (direct, sq, scon, rslot, occ, greq) are held
         * twice, one set per level parity, pstride rays apart, so a level
         * can be shaded while the previous one's shadow rays are traced */
        uint64_t pstride = 0;
        std::vector<DevBuf> qlev;
        DevBuf hits, direct, sq, scon, rslot, occ, kids, kflag, acc, greq, knnq;
        DevBuf lvcnt, cmpb; /* per-level ray counts; compaction block counts / offsets */
        uint64_t bytes() const
        {
            uint64_t b = 0;
            const DevBuf *bs[] = {&hits, &direct, &sq, &scon, &rslot, &occ, &kids, &kflag, &acc, &greq, &knnq,
                                  &lvcnt, &cmpb};
            for (const DevBuf *d : bs) b += d->bytes;
            for (const DevBuf &d : qlev) b += d.bytes;
            return b;
        }
    };
    std::vector<Lane> lanes;
    hipEvent_t tables_ready = nullptr;
    int nlanes = 1; /* lanes a pass spreads its primary chunks over */
    int sparse_from = 64; /* first bounce level traced by the SPARSE kernel instances */
    bool chain = true;    /* one-child chains in one kernel when the pass allows it */
    bool sample_runs = true; /* a request's first pass may take several samples per pixel */
    bool attr = true;     /* some scene in flight has meshes with vertex attributes */
    std::map<const void *, uint32_t> pgrid; /* persistent grid per traversal kernel instance */
    /* captured descents, keyed by everything their launches depend on */
    std::map<std::vector<uint64_t>, hipGraphExec_t> graphs;
    bool use_graphs = true;
    /* traversal grids of resident size (each wave loops): 0 never, 1 at
     * every bounce level, 2 below the primary level */
    int persistent = 0;
    /* requests in flight */
    std::vector<ReqDev> hreqs = std::vector<ReqDev>(MAX_SLOTS);
    std::vector<SceneDev> hscenes;
    Job *slot_job[MAX_SLOTS] = {};
    uint32_t slot_active[MAX_SLOTS] = {};
    uint32_t live = 0, hi = 0; /* slots in use, 1 + highest slot in use */
    /* this pass's launch shape */
    int L = 0, K = 2, maxiter = 0;
    bool photons = false, uni = true, counting = false;
    bool knn = false; /* a request gathers the k nearest photons */
    bool one_scene = true; /* every request in flight renders the same scene */
    int Kb = 2;            /* most children one ray can spawn (growth bound of the levels) */
    bool more = true;      /* some request in flight can take another sample */
    /* tunables (sray_tuning), copied from the context at each pass */
    uint32_t tile_w = 64;        /* primary-ray tile width (x 64/tile_w rows) */
    uint32_t chunk = MAX_CHUNK;  /* rays per bounce-level sub-chunk */
    int trace_block = 128;       /* threads per block of the traversal kernels */
    int gather_lanes = 16;       /* lanes per photon-gather query */
    unsigned long long sec_seen = 0; /* secondary rays counted before this pass */
    /* submit / poll */
    std::thread worker;
    std::mutex qmu;
    std::condition_variable qcv;
    std::deque<Job *> pending;
    std::map<uint64_t, std::unique_ptr<Job>> jobs;
    uint64_t next_ticket = 1;
    bool stop = false, started = false;

    uint64_t live_bytes = 0; /* footprints of the requests in flight */

    /* pinned host memory: the status of a pass, read back with one copy and
     * one synchronisation, and the staging of the request / scene tables */
    struct PassStatus {
        uint32_t nsel, errf;
        unsigned long long stats[8], rcounts[2 * NCOUNT];
        uint32_t slotcnt[MAX_SLOTS];
        unsigned long long srays[MAX_SLOTS];
    };
    PassStatus *hst = nullptr;
    ReqDev *hreq_pin = nullptr;
    SceneDev *hscn_pin = nullptr;

    /* everything held on the device */
    uint64_t bytes() const { return fixed_bytes() + arena_bytes(); }
    /* ray queues and tables (sized by the largest pass so far) */
    uint64_t fixed_bytes() const
    {
        uint64_t b = 0;
        const DevBuf *bs[] = {&stats, &errf, &rcounts, &slotcnt, &slotrays, &dreqs, &dscenes};
        for (const DevBuf *d : bs) b += d->bytes;
        for (const Lane &l : lanes) b += l.bytes();
        return b;
    }
    /* the pixel and sample arenas */
    uint64_t arena_bytes() const
    {
        uint64_t b = 0;
        const DevBuf *bs[] = {&fb, &nsamp, &preq, &keep, &act[0], &act[1], &subpix};
        for (const DevBuf *d : bs) b += d->bytes;
        return b;
    }
};

/* grows b to bytes keeping its first keep bytes */
int regrow(sray_gpu_ctx *c, DevBuf &b, size_t bytes, size_t keep)
{
    if (b.bytes >= bytes) return SRAY_OK;
    DevBuf n;
    hipError_t e = hipMalloc(&n.p, bytes);
    if (e != hipSuccess) return fail(c, SRAY_ERR_NOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    n.bytes = bytes;
    if (b.p && keep) HIPCHK(c, hipMemcpyAsync(n.p, b.p, std::min(keep, b.bytes), hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    release(b);
    b = n;
    return SRAY_OK;
}

int grow_pixels(Engine &E, uint64_t need)
{
    sray_gpu_ctx *c = E.c;
    uint64_t cap = std::max<uint64_t>(std::max<uint64_t>(E.pix.cap * 2, 1u << 20), E.pix.cap + need);
    /* active-list entries hold an arena pixel in 28 bits */
    if (cap > (uint64_t)ENT_PIX + 1u) {
        if (E.pix.cap + need > (uint64_t)ENT_PIX + 1u)
            return fail(c, SRAY_ERR_LIMIT, "pixel arena beyond 2^28 pixels in flight");
        cap = (uint64_t)ENT_PIX + 1u;
    }
    uint64_t old = E.pix.cap;
    int rc;
    if ((rc = regrow(c, E.fb, cap * 16, old * 16)) || (rc = regrow(c, E.nsamp, cap * 4, old * 4)) ||
        (rc = regrow(c, E.preq, cap * 4, old * 4)) || (rc = regrow(c, E.keep, cap, 0)) ||
        (rc = regrow(c, E.act[E.cur], cap * 4, (size_t)E.nact * 4)) || (rc = regrow(c, E.act[E.cur ^ 1], cap * 4, 0)))
        return rc;
    E.pix.grow(cap);
    return SRAY_OK;
}

/* room for `need` active-list entries (a pixel may have several: one per
 * sample its request takes in a pass), keeping the current list */
int grow_list(Engine &E, uint64_t need)
{
    sray_gpu_ctx *c = E.c;
    const uint64_t have = std::min<uint64_t>(std::min(E.act[0].bytes, E.act[1].bytes) / 4, E.keep.bytes);
    if (need <= have) return SRAY_OK;
    if (need > 0xfffff000ull) return fail(c, SRAY_ERR_LIMIT, "active list beyond 2^32 entries");
    const uint64_t cap = std::min<uint64_t>(std::max(need, have * 2), 0xfffff000ull);
    int rc;
    if ((rc = regrow(c, E.act[E.cur], cap * 4, (size_t)E.nact * 4)) || (rc = regrow(c, E.act[E.cur ^ 1], cap * 4, 0)) ||
        (rc = regrow(c, E.keep, cap, 0)))
        return rc;
    return SRAY_OK;
}

int grow_sub(Engine &E, uint64_t need)
{
    uint64_t cap = std::max<uint64_t>(std::max<uint64_t>(E.sub.cap * 2, 1u << 20), E.sub.cap + need);
    int rc = regrow(E.c, E.subpix, cap * 16, E.sub.cap * 16);
    if (rc) return rc;
    E.sub.grow(cap);
    return SRAY_OK;
}

/* a request's own camera (tab null), or the scene's with its rows */
CamDev cam_dev(const sray_camera_desc &d, const float *tab = nullptr, int t0 = 0, int n = 0)
{
    CamDev C;
    memset(&C, 0, sizeof C);
    memcpy(C.cam, d.xform, sizeof C.cam);
    C.tab = tab;
    C.t0 = t0;
    C.n = n;
    C.fz = (float)(1.0 / std::tan(0.5 * d.vfov));
    C.max_dist = d.max_dist;
    C.shutter = d.shutter;
    return C;
}

/* device bytes a request needs while in flight (admission accounting) */
uint64_t job_bytes(const Job &j)
{
    uint64_t npix = (uint64_t)(j.p.x1 - j.p.x0) * (uint64_t)(j.p.y1 - j.p.y0);
    uint64_t maxs = (uint64_t)std::max(1, j.p.max_samples);
    return npix * (16 + 4 + 4 + 1 + 8 + 16) + (maxs > 1 ? npix * maxs * 16 : 0);
}

/* what admission counts as taken: resident scenes and photon maps, the
 * engines' queues, and the footprints of the requests in flight */
uint64_t admitted_bytes(sray_gpu_ctx *c)
{
    uint64_t b = 0;
    for (auto &kv : c->scenes) b += kv.second->bytes;
    for (auto &kv : c->pmaps) b += kv.second->bytes;
    if (c->plain) b += c->plain->bytes();
    if (c->serve) b += c->serve->fixed_bytes() + c->serve->live_bytes;
    return b;
}

void job_done(Engine &E, Job *j, int rc, const char *msg)
{
    std::lock_guard<std::mutex> lk(E.qmu);
    if (rc) {
        j->rc = rc;
        if (msg) j->err = msg;
    }
    j->latency_ms = std::chrono::duration<double, std::milli>(Clock::now() - j->t_submit).count();
    emit_status(E.c, 'e', j->p.x0, j->p.y0, j->p.x1 - j->p.x0, j->p.y1 - j->p.y0);
    emit_status(E.c, 'd', (int)j->latency_ms, (int)j->queue_ms, 0, 0);
    j->state = 2;
    E.qcv.notify_all();
}

void apply_tuning(Engine &E)
{
    const sray_tuning &t = E.c->tune;
    E.trace_block = t.trace_block;
    E.tile_w = (uint32_t)t.tile_w;
    E.gather_lanes = t.gather_lanes;
    E.chunk = 1u << t.chunk_log2;
    E.use_graphs = t.graphs != 0;
    E.persistent = t.persistent;
    E.nlanes = t.lanes;
    E.sparse_from = t.sparse_from;
    E.chain = t.chain != 0;
    E.sample_runs = t.sample_runs != 0;
}

/* places a request in the arena and the active list */
int admit(Engine &E, Job *j)
{
    sray_gpu_ctx *c = E.c;
    apply_tuning(E);
    int rc = validate_params(c, &j->p);
    if (rc) return rc;
    GpuScene *g = find_scene(c, j->scene_id);
    if (!g) return fail(c, SRAY_ERR_NOTFOUND, "unknown scene %u", j->scene_id);
    if (g->timed) {
        /* ray times: frame_time, or frame_time - shutter/2 + [0, shutter)
         * with motion blur (Camera::get_primary_ray, src/camera.cc:137-139) */
        const int sh = j->use_cam ? j->cam.shutter : g->camera.shutter;
        const bool blur = j->p.mblur && sh > 0;
        const int lo = blur ? j->p.frame_time - sh / 2 : j->p.frame_time;
        const int hi = blur ? lo + sh - 1 : j->p.frame_time;
        if (lo < g->frame_t0 || hi > g->frame_t1)
            return fail(c, SRAY_ERR_ARG,
                        "ray times %d..%d msec leave the scene's keyframed interval %d..%d (set the frame first)", lo,
                        hi, g->frame_t0, g->frame_t1);
    }
    uint32_t s = 0;
    while (s < MAX_SLOTS && E.slot_job[s]) s++;
    if (s == MAX_SLOTS) return fail(c, SRAY_ERR_LIMIT, "more than %u requests in flight", MAX_SLOTS);
    ReqDev Q;
    memset(&Q, 0, sizeof Q);
    fill_render_dev(&j->p, Q.P);
    Q.C = j->use_cam ? cam_dev(j->cam) : cam_dev(g->camera, g->dev.cam_tab, g->dev.cam_t0, g->dev.cam_n);
    for (int kind = 0; kind < 2; kind++) {
        bool want = kind == SRAY_PMAP_CAUSTIC ? Q.P.caust != 0 : Q.P.gi != 0;
        if (!want) continue;
        if (c->pmaps.find(g->pmap[kind]) == c->pmaps.end())
            return fail(c, SRAY_ERR_NOTFOUND, "render asks for the %s photon map but none is bound to scene %u",
                        kind == SRAY_PMAP_CAUSTIC ? "caustic" : "global", j->scene_id);
    }
    /* gather distance relative to the scene octree diagonal (src/scene.cc:384-387) */
    double gabs = g->diag * (double)j->p.gather_dist;
    Q.PM.radius = (float)gabs;
    Q.PM.k = j->p.gather_k > 0 ? j->p.gather_k : 0;
    Q.PM.scale = gabs > 0.0 ? 1.0 / (M_PI * (gabs * gabs)) : 0.0;
    for (int kind = 0; kind < 2; kind++) {
        bool want = kind == SRAY_PMAP_CAUSTIC ? Q.P.caust != 0 : Q.P.gi != 0;
        /* k-nearest gathers scan smaller boxes (their first radius is a
         * fraction of the gather radius): cells of half the tuned size */
        const float cell = c->tune.gather_cell * (Q.PM.k > 0 ? 0.5f : 1.0f);
        if (want && (rc = pmap_grid_get(c, g->pmap[kind], Q.PM.radius, cell,
                                        kind == SRAY_PMAP_CAUSTIC ? &Q.PM.caust : &Q.PM.glob)))
            return rc;
        j->pmaps[kind] = want ? g->pmap[kind] : 0u;
    }
    uint64_t npix = (uint64_t)Q.P.W * (uint64_t)(Q.P.y1 - Q.P.y0);
    uint64_t maxs = (uint64_t)Q.P.max_samples;
    uint64_t base = 0, sbase = 0;
    if (!E.pix.alloc(npix, base)) {
        if ((rc = grow_pixels(E, npix))) return rc;
        if (!E.pix.alloc(npix, base)) return fail(c, SRAY_ERR_NOMEM, "pixel arena");
    }
    if (maxs > 1 && !E.sub.alloc(npix * maxs, sbase)) {
        if ((rc = grow_sub(E, npix * maxs)) || !E.sub.alloc(npix * maxs, sbase)) {
            E.pix.release(base, npix);
            return rc ? rc : fail(c, SRAY_ERR_NOMEM, "sample arena");
        }
    }
    Q.pix_base = (uint32_t)base;
    Q.sub_base = sbase;
    /* The first pass takes the samples every pixel takes: the variance
     * test runs from max(min_samples, 2) samples on (all of max_samples
     * when min_samples >= max_samples), so that many are always taken.
     * When that count, or its largest power of two <= 16, is above one, a
     * pixel's first samples are consecutive entries of one aligned group
     * of 16, which no primary chunk (a multiple of 1024 entries) splits;
     * later passes take one sample per pixel.  The entries start at the
     * list's last multiple of 16; the fewer than 16 entries past it (one
     * sample each, as every entry of a group that a run does not fill)
     * move behind the request's. */
    uint32_t first = (uint32_t)std::min<uint64_t>(maxs, (uint64_t)std::max(Q.P.min_samples, 2));
    first = std::min(first, 16u);
    while (first & (first - 1u)) first &= first - 1u;
    const bool runs = E.sample_runs && first > 1u && (npix * first) % 16 == 0;
    Q.run = runs ? first : 1u;
    const uint64_t nent = npix * Q.run;
    const uint64_t at = runs ? (E.nact & ~15ull) : E.nact, tail = E.nact - at;
    if ((rc = grow_list(E, E.nact + nent))) {
        E.pix.release(base, npix);
        if (maxs > 1) E.sub.release(sbase, npix * maxs);
        return rc;
    }
    uint32_t *act = dptr<uint32_t>(E.act[E.cur]);
    if (tail) HIPCHK(c, hipMemcpyAsync(act + at + nent, act + at, tail * 4, hipMemcpyDeviceToDevice, c->stream));
    k_admit<<<blocks(nent, 256), 256, 0, c->stream>>>(act + at, dptr<float4>(E.fb) + base,
                                                      dptr<uint32_t>(E.nsamp) + base, dptr<uint32_t>(E.preq) + base,
                                                      (uint32_t)npix, (uint32_t)base, (uint32_t)Q.P.W,
                                                      (uint32_t)(Q.P.y1 - Q.P.y0), E.tile_w, Q.run, s);
    HIPCHK(c, hipGetLastError());
    E.nact += (uint32_t)nent;
    E.hreqs[s] = Q;
    E.slot_job[s] = j;
    E.slot_active[s] = (uint32_t)nent;
    E.live++;
    E.live_bytes += job_bytes(*j);
    E.hi = std::max(E.hi, s + 1);
    j->slot = s;
    j->pix_base = base;
    j->npix = npix;
    j->sub_base = sbase;
    j->nsub = maxs > 1 ? npix * maxs : 0;
    j->state = 1;
    j->t_admit = Clock::now();
    j->queue_ms = std::chrono::duration<double, std::milli>(j->t_admit - j->t_submit).count();
    emit_status(E.c, 's', j->p.x0, j->p.y0, j->p.x1 - j->p.x0, j->p.y1 - j->p.y0);
    return SRAY_OK;
}

void retire(Engine &E, uint32_t s)
{
    Job *j = E.slot_job[s];
    E.pix.release(j->pix_base, j->npix);
    if (j->nsub) E.sub.release(j->sub_base, j->nsub);
    E.slot_job[s] = nullptr;
    E.slot_active[s] = 0;
    E.live--;
    E.live_bytes -= job_bytes(*j);
    while (E.hi > 0 && !E.slot_job[E.hi - 1]) E.hi--;
}

/* drops every request in flight (after a device error) */
void abort_all(Engine &E)
{
    for (Engine::Lane &ln : E.lanes) {
        if (ln.s2) (void)hipStreamSynchronize(ln.s2);
        if (ln.s) (void)hipStreamSynchronize(ln.s);
    }
    (void)hipStreamSynchronize(E.c->stream);
    for (uint32_t s = 0; s < E.hi; s++)
        if (E.slot_job[s]) {
            release(E.slot_job[s]->stage);
            retire(E, s);
        }
    E.nact = 0;
}

int build_tables(Engine &E)
{
    sray_gpu_ctx *c = E.c;
    apply_tuning(E);
    E.hscenes.clear();
    std::map<uint32_t, uint32_t> idx;
    E.L = 0;
    E.K = 2;
    E.Kb = 0;
    E.attr = false;
    E.maxiter = 0;
    E.photons = false;
    E.knn = false;
    E.more = false;
    if (!E.hst) {
        if (hipHostMalloc((void **)&E.hst, sizeof(Engine::PassStatus), hipHostMallocDefault) != hipSuccess ||
            hipHostMalloc((void **)&E.hreq_pin, MAX_SLOTS * sizeof(ReqDev), hipHostMallocDefault) != hipSuccess ||
            hipHostMalloc((void **)&E.hscn_pin, MAX_SLOTS * sizeof(SceneDev), hipHostMallocDefault) != hipSuccess)
            return fail(c, SRAY_ERR_NOMEM, "pinned host memory for the pass status");
        memset(E.hst, 0, sizeof(Engine::PassStatus));
    }
    for (uint32_t s = 0; s < E.hi; s++) {
        Job *j = E.slot_job[s];
        if (!j) continue;
        GpuScene *g = find_scene(c, j->scene_id);
        if (!g) return fail(c, SRAY_ERR_NOTFOUND, "scene %u released with requests in flight", j->scene_id);
        auto it = idx.find(j->scene_id);
        if (it == idx.end()) {
            it = idx.emplace(j->scene_id, (uint32_t)E.hscenes.size()).first;
            E.hscenes.push_back(g->dev);
        }
        ReqDev &Q = E.hreqs[s];
        Q.scene = it->second;
        /* shadow slots per ray: Phong samples area lights shadow_samples
         * times and point lights once; shade_undef samples each light once */
        E.L = std::max(E.L, g->dev.nlights * std::max(1, Q.P.shadow_samples));
        E.K = std::max(E.K, 2 + (Q.P.gi ? Q.P.diffuse_samples : 0));
        E.maxiter = std::max(E.maxiter, Q.P.iter);
        E.photons = E.photons || Q.P.gi || Q.P.caust;
        E.knn = E.knn || ((Q.P.gi || Q.P.caust) && Q.PM.k > 0);
        E.more = E.more || Q.P.max_samples > 1;
        /* children a ray of this request can spawn: reflect / refract only
         * with a Phong material whose factor is > 0 (or any Phong material
         * when min_energy < 0 lets zero-energy children through), plus the
         * diffuse GI rays */
        const bool any = g->has_phong && Q.P.min_energy < 0.0f;
        const int br = ((g->pos_reflect || any) ? 1 : 0) + ((g->pos_refract || any) ? 1 : 0) +
                       (Q.P.gi ? Q.P.diffuse_samples : 0);
        E.Kb = std::max(E.Kb, br);
        E.attr = E.attr || g->attr_meshes;
    }
    E.uni = E.live == 1 && E.slot_job[0] != nullptr;
    E.one_scene = E.hscenes.size() == 1;
    int rc;
    if ((rc = ensure(c, E.dreqs, (size_t)E.hi * sizeof(ReqDev))) ||
        (rc = ensure(c, E.dscenes, E.hscenes.size() * sizeof(SceneDev))))
        return rc;
    /* through pinned staging (the previous pass's copies are complete: every
     * pass ends with a synchronisation) */
    memcpy(E.hreq_pin, E.hreqs.data(), (size_t)E.hi * sizeof(ReqDev));
    memcpy(E.hscn_pin, E.hscenes.data(), E.hscenes.size() * sizeof(SceneDev));
    HIPCHK(c, hipMemcpyAsync(E.dreqs.p, E.hreq_pin, (size_t)E.hi * sizeof(ReqDev), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(E.dscenes.p, E.hscn_pin, E.hscenes.size() * sizeof(SceneDev), hipMemcpyHostToDevice,
                             c->stream));
    return SRAY_OK;
}

struct Level {
    Engine *E;
    Engine::Lane *ln; /* the stream and scratch of this chunk */
    Tabs T;
    const uint32_t *active;
    uint32_t c0;
};

/* grid of a persistent traversal kernel: as many blocks as the device
 * keeps resident at once */
template <typename K> uint32_t persistent_grid(Engine &E, K kern, int bs)
{
    auto it = E.pgrid.find((const void *)kern);
    if (it != E.pgrid.end()) return it->second;
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, bs, 0) != hipSuccess || per_cu < 1) per_cu = 1;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, E.c->device) != hipSuccess || cus < 1)
        cus = 1;
    uint32_t g = (uint32_t)per_cu * (uint32_t)cus;
    E.pgrid[(const void *)kern] = g;
    return g;
}

/* one level step: rays [off, off + chunk) of level b's queue (count on the
 * device) through trace, shade, gather, shadow trace and accumulation,
 * and (below the deepest level) the stable compaction of their children
 * into level b + 1's queue, which also writes that level's count.
 *
 * The photon gathers, the shadow trace and the accumulation go to the
 * lane's second stream once the level is shaded, so they run beside the
 * compaction and the next step's closest-hit trace; the next step's
 * shading (which reuses their buffers) waits for them (first = no earlier
 * step to wait for). */
int launch_step(Level &LV, int b, uint32_t off, uint32_t cap, int k)
{
    Engine &E = *LV.E;
    Engine::Lane &ln = *LV.ln;
    sray_gpu_ctx *c = E.c;
    /* with per-kernel timing on, every kernel of the step goes to the lane's
     * own stream: an event-timed launch then has the device to itself */
    const hipStream_t st = ln.s, st2 = c->profiling ? ln.s : ln.s2;
    const StepN sn{dptr<uint32_t>(ln.lvcnt) + b, off, cap};
    const QRay *q = dptr<QRay>(ln.qlev[b]) + off;
    unsigned long long *rc6 = dptr<unsigned long long>(E.rcounts);
    unsigned long long *stats = dptr<unsigned long long>(E.stats);
    unsigned long long *srays = dptr<unsigned long long>(E.slotrays);
    /* this level's set of shading outputs */
    const int par = b & 1;
    const size_t po = (size_t)par * ln.pstride, pl = po * (size_t)std::max(E.L, 1);
    float4 *direct = dptr<float4>(ln.direct) + po;
    SRay *sq = dptr<SRay>(ln.sq) + pl;
    float4 *scon = dptr<float4>(ln.scon) + pl;
    uint32_t *rslot = dptr<uint32_t>(ln.rslot) + po;
    uint8_t *occ = dptr<uint8_t>(ln.occ) + pl;
    GatherReq *greq = E.photons ? dptr<GatherReq>(ln.greq) + po : nullptr;
    uint32_t *knn = E.photons && E.knn ? dptr<uint32_t>(ln.knnq) : nullptr;
    const int bs = E.trace_block;
    const uint32_t small_grid = std::min<uint32_t>(blocks(cap, 256), 2048u);
    /* traversal grid: one run of 64 rays per wave, or the resident waves */
    const bool P = E.persistent == 1 || (E.persistent == 2 && b > 0), SP = b >= E.sparse_from;
    auto tgrid = [&](auto kern, size_t items) {
        uint32_t g = blocks(items, bs);
        return P ? std::min(g, persistent_grid(E, kern, bs)) : g;
    };
    {
        ProfScope ps(c, "k_trace_q", st);
        /* the traversal reads only the scene: one scene = the uniform variant */
        auto kern = E.counting ? (E.one_scene ? trace_q_of<true, true>(bs, P, SP) : trace_q_of<true, false>(bs, P, SP))
                               : (E.one_scene ? trace_q_of<false, true>(bs, P, SP) : trace_q_of<false, false>(bs, P, SP));
        kern<<<tgrid(kern, cap), bs, 0, st>>>(LV.T, q, sn, dptr<int4>(ln.hits), rc6);
    }
    /* the set was last used two steps back (k-th step of this descent; the
     * steps before it were joined) */
    if (k >= 2) HIPCHK(c, hipStreamWaitEvent(st, ln.evB[par], 0));
    {
        ProfScope ps(c, "k_shade", st);
        auto kern = E.photons ? (E.uni ? k_shade<true, true, true>
                                       : E.one_scene ? k_shade<true, false, true> : k_shade<true, false, false>)
                              : (E.uni ? k_shade<false, true, true>
                                       : E.one_scene ? k_shade<false, false, true> : k_shade<false, false, false>);
        kern<<<small_grid, 256, 0, st>>>(LV.T, LV.active, LV.c0, q, dptr<int4>(ln.hits), sn, E.L, E.K,
                                               direct, sq, scon, rslot, dptr<QRay>(ln.kids),
                                               dptr<uint8_t>(ln.kflag), greq, stats, srays,
                                               dptr<uint32_t>(E.errf));
    }
    /* the children are compacted into the next level before the second
     * stream starts: the compaction reads only what k_shade wrote, and
     * launched beside the shadow trace its few blocks wait for free CUs
     * while the next level's closest-hit trace waits for them */
    if (b < E.maxiter) {
        ProfScope ps(c, "k_compact", st);
        uint32_t *bc = dptr<uint32_t>(ln.cmpb);
        k_cmp_count<<<CMP_BLOCKS, CMP_THREADS, 0, st>>>(dptr<uint8_t>(ln.kflag), sn, E.K, bc);
        k_cmp_scatter<<<CMP_BLOCKS, CMP_THREADS, 0, st>>>(dptr<QRay>(ln.kids), dptr<uint8_t>(ln.kflag), sn,
                                                                 E.K, bc, dptr<uint32_t>(ln.lvcnt) + b + 1,
                                                                 dptr<QRay>(ln.qlev[b + 1]));
    }
    HIPCHK(c, hipEventRecord(ln.evA, st));
    HIPCHK(c, hipStreamWaitEvent(st2, ln.evA, 0));
    if (E.photons) {
        ProfScope ps(c, "k_gather", st2);
        auto kern = E.knn ? (E.uni ? gather_of<true, true>(E.gather_lanes) : gather_of<false, true>(E.gather_lanes))
                          : (E.uni ? gather_of<true, false>(E.gather_lanes) : gather_of<false, false>(E.gather_lanes));
        const uint32_t g = std::min<uint32_t>(blocks((size_t)cap * E.gather_lanes, 256), 8192u);
        kern<<<g, 256, 0, st2>>>(LV.T, q, greq, sn, direct, knn);
    }
    if (E.photons && E.knn) {
        /* the requests with more than k photons in range (a few percent) */
        ProfScope ps(c, "k_gather_knn", st2);
        auto kern = E.uni ? gather_knn_of<true>(E.gather_lanes) : gather_knn_of<false>(E.gather_lanes);
        const uint32_t g = std::min<uint32_t>(blocks((size_t)cap * E.gather_lanes, 256), 1024u);
        kern<<<g, 256, 0, st2>>>(LV.T, q, greq, knn, direct);
    }
    if (E.L) {
        ProfScope ps(c, "k_trace_shadow", st2);
        auto kern = E.counting ? (E.one_scene ? trace_shadow_of<true, true>(bs, P, SP)
                                              : trace_shadow_of<true, false>(bs, P, SP))
                               : (E.one_scene ? trace_shadow_of<false, true>(bs, P, SP)
                                              : trace_shadow_of<false, false>(bs, P, SP));
        kern<<<tgrid(kern, (size_t)cap * E.L), bs, 0, st2>>>(LV.T, sq, scon, sn, E.L, occ, rc6 + NCOUNT);
    }
    {
        ProfScope ps(c, "k_accum", st2);
        k_accum<<<small_grid, 256, 0, st2>>>(rslot, direct, scon, occ, sn, E.L,
                                                    dptr<float4>(ln.acc), knn);
    }
    HIPCHK(c, hipEventRecord(ln.evB[par], st2));
    HIPCHK(c, hipGetLastError());
    return SRAY_OK;
}

/* the step (b, off) of at most cap rays, then the first sub-chunk of
 * every deeper level: no host round trip between levels.  A level holds
 * at most K times the rays of the one above, which bounds each step's
 * launch size (small requests launch small grids). */
int descend(Level &LV, int b, uint32_t off, uint32_t cap)
{
    Engine &E = *LV.E;
    int rc;
    uint64_t bound = cap;
    int last = b;
    for (int d = b; d <= E.maxiter && bound > 0; d++) {
        if ((rc = launch_step(LV, d, d == b ? off : 0u, (uint32_t)std::min<uint64_t>(bound, E.chunk), d - b)))
            return rc;
        last = d;
        bound = std::min<uint64_t>(bound * (uint64_t)E.Kb, E.chunk);
    }
    /* join: what follows on the lane's stream sees the last accumulation
     * (the second stream runs its work in order) */
    HIPCHK(E.c, hipStreamWaitEvent(LV.ln->s, LV.ln->evB[last & 1], 0));
    return SRAY_OK;
}

/* After a descent from level b: the sub-chunks beyond the first of levels
 * b+1 .. maxiter, deepest first, each followed by its own descent.  This
 * is exactly the depth-first order of processing every level in chunks
 * with each chunk's children before the next chunk (the queues of levels
 * <= d are untouched while level d's subtree runs). */
int finish_levels(Level &LV, int b)
{
    Engine &E = *LV.E;
    sray_gpu_ctx *c = E.c;
    if (b >= E.maxiter) return SRAY_OK;
    uint32_t n[64];
    const int nl = E.maxiter - b;
    HIPCHK(c, hipMemcpyAsync(n, dptr<uint32_t>(LV.ln->lvcnt) + b + 1, (size_t)nl * 4, hipMemcpyDeviceToHost,
                             LV.ln->s));
    HIPCHK(c, hipStreamSynchronize(LV.ln->s));
    int rc;
    for (int d = E.maxiter; d > b; d--)
        for (uint64_t off = E.chunk; off < n[d - b - 1]; off += E.chunk) {
            if ((rc = descend(LV, d, (uint32_t)off, E.chunk))) return rc;
            if ((rc = finish_levels(LV, d))) return rc;
        }
    return SRAY_OK;
}

/* whether some level of a descent from m primary rays can hold more rays
 * than one sub-chunk (a level holds at most Kb times the previous one) */
bool may_overflow(const Engine &E, uint32_t m)
{
    uint64_t worst = m;
    for (int d = 1; d <= E.maxiter; d++) {
        worst *= (uint64_t)E.Kb;
        if (worst > E.chunk) return true;
    }
    return false;
}

/* whether this pass's samples are one-child chains (k_chain) */
bool chain_pass(const Engine &E)
{
    return E.chain && !E.photons && !E.counting && E.Kb <= 1 && E.L <= 4;
}

/* the primary chunk [c0, c0 + m) as one-child chains: one kernel */
int launch_chain(Level &LV, uint32_t m)
{
    Engine &E = *LV.E;
    Engine::Lane &ln = *LV.ln;
    sray_gpu_ctx *c = E.c;
    constexpr int BS = 128;
    ProfScope ps(c, "k_chain", ln.s);
    auto pick = [&](auto u, auto o, auto n) { return E.uni ? u : E.one_scene ? o : n; };
    auto kern = E.attr ? (E.L <= 1 ? pick(k_chain<true, true, 1, true, BS>, k_chain<false, true, 1, true, BS>,
                                          k_chain<false, false, 1, true, BS>)
                                   : pick(k_chain<true, true, 4, true, BS>, k_chain<false, true, 4, true, BS>,
                                          k_chain<false, false, 4, true, BS>))
                       : (E.L <= 1 ? pick(k_chain<true, true, 1, false, BS>, k_chain<false, true, 1, false, BS>,
                                          k_chain<false, false, 1, false, BS>)
                                   : pick(k_chain<true, true, 4, false, BS>, k_chain<false, true, 4, false, BS>,
                                          k_chain<false, false, 4, false, BS>));
    kern<<<blocks(m, BS), BS, 0, ln.s>>>(LV.T, LV.active, LV.c0, m, E.L, E.K, dptr<float4>(ln.acc),
                                         dptr<unsigned long long>(E.stats), dptr<unsigned long long>(E.slotrays),
                                         dptr<uint32_t>(E.errf));
    HIPCHK(c, hipGetLastError());
    return SRAY_OK;
}

/* destroys every captured descent.  A graph launch returns before its
 * kernels run, and destroying an executable graph frees its kernel
 * argument memory, which kernels still queued (and a profiler's dispatch
 * records) read: so the streams the graphs were launched on are drained
 * first. */
void drop_graphs(Engine &E)
{
    if (E.graphs.empty()) return;
    for (Engine::Lane &ln : E.lanes) {
        if (ln.s) (void)hipStreamSynchronize(ln.s);
        if (ln.s2) (void)hipStreamSynchronize(ln.s2);
    }
    (void)hipStreamSynchronize(E.c->stream);
    for (auto &kv : E.graphs) (void)hipGraphExecDestroy(kv.second);
    E.graphs.clear();
}

/* ray generation and the descent of the primary chunk [c0, c0 + m): as
 * kernel launches, or as one captured graph replayed (same kernels, same
 * order) */
int launch_chunk(Level &LV, uint32_t m, bool one_lane)
{
    Engine &E = *LV.E;
    Engine::Lane &ln = *LV.ln;
    sray_gpu_ctx *c = E.c;
    const hipStream_t st = ln.s;
    auto body = [&]() -> int {
        k_fill_f4<<<blocks(m, 256), 256, 0, st>>>(dptr<float4>(ln.acc), m, make_float4(0, 0, 0, 0));
        {
            ProfScope ps(c, "k_raygen", st);
            auto kern = E.uni ? k_raygen<true> : k_raygen<false>;
            kern<<<blocks(m, 256), 256, 0, st>>>(LV.T, LV.active, LV.c0, m, dptr<QRay>(ln.qlev[0]),
                                                 dptr<uint32_t>(ln.lvcnt));
        }
        return descend(LV, 0, 0, m);
    };
    /* graphs only for a pass on one lane: measured, graphs replayed on two
     * lanes at once ran half again slower than the same launches made kernel
     * by kernel once a graph had been re-instantiated (4 hardware queues), and
     * 7% slower with 8 queues */
    if (!E.use_graphs || c->profiling || !one_lane) return body();
    std::vector<uint64_t> key = {LV.c0, m, (uint64_t)E.maxiter, (uint64_t)E.L, (uint64_t)E.K, E.chunk,
                                 (uint64_t)E.trace_block, (uint64_t)E.gather_lanes,
                                 (uint64_t)(E.photons | E.uni << 1 | E.one_scene << 2 | E.counting << 3 |
                                            E.persistent << 4 | E.knn << 6),
                                 (uint64_t)E.Kb,
                                 /* the first level traced by the sparse instances */
                                 (uint64_t)std::min(E.sparse_from, E.maxiter + 1),
                                 (uint64_t)(uintptr_t)LV.T.scenes, (uint64_t)(uintptr_t)LV.T.reqs,
                                 (uint64_t)(uintptr_t)LV.T.preq, (uint64_t)(uintptr_t)LV.T.nsamp,
                                 (uint64_t)(uintptr_t)LV.active};
    const DevBuf *bufs[] = {&ln.hits, &ln.direct, &ln.sq, &ln.scon, &ln.rslot, &ln.occ, &ln.kids, &ln.kflag, &ln.acc,
                            &E.stats, &E.errf, &E.rcounts, &E.slotrays, &ln.greq, &ln.knnq, &ln.lvcnt, &ln.cmpb};
    for (const DevBuf *d : bufs) key.push_back((uint64_t)(uintptr_t)d->p);
    key.push_back(ln.pstride);
    for (int b = 0; b <= E.maxiter; b++) key.push_back((uint64_t)(uintptr_t)ln.qlev[b].p);
    key.push_back((uint64_t)(uintptr_t)st);
    auto it = E.graphs.find(key);
    if (it == E.graphs.end()) {
        if (E.graphs.size() >= 64) drop_graphs(E);
        HIPCHK(c, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        int rc = body();
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(st, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return fail(c, SRAY_ERR_HIP, "graph capture failed: %s", hipGetErrorString(e));
        hipGraphExec_t x = nullptr;
        e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) return fail(c, SRAY_ERR_HIP, "graph instantiation failed: %s", hipGetErrorString(e));
        it = E.graphs.emplace(key, x).first;
    }
    HIPCHK(c, hipGraphLaunch(it->second, st));
    return SRAY_OK;
}

/* the image of the request in slot s: average (or the reference's Q1
 * scaling) into its output, through the staging buffer to host memory */
int finalize_job(Engine &E, uint32_t s)
{
    sray_gpu_ctx *c = E.c;
    Job *j = E.slot_job[s];
    const ReqDev &Q = E.hreqs[s];
    uint32_t npix = (uint32_t)j->npix;
    float *dst = j->dev_out;
    int rc;
    if (!dst) {
        if ((rc = ensure(c, j->stage, (size_t)npix * 16))) return rc;
        dst = dptr<float>(j->stage);
    }
    k_finalize<<<blocks(npix, 256), 256, 0, c->stream>>>(Q.P, dptr<float4>(E.fb) + j->pix_base,
                                                        dptr<int>(E.nsamp) + j->pix_base, dst, npix);
    if (Q.P.flags & SRAY_RENDER_REF_Q1)
        k_q1_scale<<<blocks(npix, 256), 256, 0, c->stream>>>(Q.P, dptr<int>(E.nsamp) + j->pix_base, dst, npix);
    HIPCHK(c, hipGetLastError());
    if (j->host_out)
        HIPCHK(c, hipMemcpyAsync(j->host_out, dst, (size_t)npix * 16, hipMemcpyDeviceToHost, c->stream));
    return SRAY_OK;
}

/* the images of the finished requests of a pass: those without the Q1
 * pattern FIN_MAX to a k_finalize_many launch, the others one by one;
 * then the host copies */
int finalize_jobs(Engine &E, const std::vector<uint32_t> &slots)
{
    sray_gpu_ctx *c = E.c;
    int rc;
    FinBatch B;
    B.nj = 0;
    B.total = 0;
    auto flush = [&]() -> int {
        if (!B.nj) return SRAY_OK;
        k_finalize_many<<<std::min<uint32_t>(blocks(B.total, 256), 4096u), 256, 0, c->stream>>>(B);
        HIPCHK(c, hipGetLastError());
        B.nj = 0;
        B.total = 0;
        return SRAY_OK;
    };
    for (uint32_t s : slots) {
        Job *j = E.slot_job[s];
        const ReqDev &Q = E.hreqs[s];
        if (Q.P.flags & SRAY_RENDER_REF_Q1) {
            if ((rc = finalize_job(E, s))) return rc;
            continue;
        }
        float *dst = j->dev_out;
        if (!dst) {
            if ((rc = ensure(c, j->stage, (size_t)j->npix * 16))) return rc;
            dst = dptr<float>(j->stage);
        }
        if (B.nj == FIN_MAX || (uint64_t)B.total + j->npix > 0xffffffffull)
            if ((rc = flush())) return rc;
        B.j[B.nj++] = FinJob{dptr<float4>(E.fb) + j->pix_base, dptr<int>(E.nsamp) + j->pix_base, dst,
                             (uint32_t)j->npix, B.total};
        B.total += (uint32_t)j->npix;
    }
    if ((rc = flush())) return rc;
    for (uint32_t s : slots) {
        Job *j = E.slot_job[s];
        if (j->host_out && !(E.hreqs[s].P.flags & SRAY_RENDER_REF_Q1))
            HIPCHK(c, hipMemcpyAsync(j->host_out, dptr<float>(j->stage), (size_t)j->npix * 16,
                                     hipMemcpyDeviceToHost, c->stream));
    }
    return SRAY_OK;
}

/* one sample of every active pixel; retires the requests that finish */
int run_pass(Engine &E, std::vector<Job *> &finished, bool sync_done)
{
    sray_gpu_ctx *c = E.c;
    int rc;
    if ((rc = build_tables(E))) return rc;
    if ((rc = ensure(c, E.slotcnt, MAX_SLOTS * 4)) || (rc = ensure(c, E.slotrays, MAX_SLOTS * 8)) ||
        (rc = ensure(c, c->nsel, 4)) || (rc = ensure_zero(c, E.errf, 4)) ||
        (rc = ensure_zero(c, E.stats, 8 * sizeof(unsigned long long))) ||
        (rc = ensure_zero(c, E.rcounts, 2 * NCOUNT * sizeof(unsigned long long))))
        return rc;
    if (E.sub.cap == 0 && (rc = ensure(c, E.subpix, 16))) return rc;
    /* primary chunks: with several lanes a pass is spread over them, each
     * lane running its chunks in order.  A pass of fewer pixels than a
     * chunk is split only into pieces of at least 2^20 pixels: below that
     * the extra launches and stream hand-offs cost more than the overlap of
     * the deep levels gains. */
    const uint32_t want = (uint32_t)std::max(1, E.nlanes);
    uint32_t mchunk = E.chunk;
    if (want > 1) {
        uint64_t per = ((uint64_t)E.nact + want - 1) / want;
        per = std::max<uint64_t>((per + 1023) & ~1023ull, 1u << 20);
        mchunk = (uint32_t)std::min<uint64_t>(per, E.chunk);
    }
    const uint32_t nchunks = (E.nact + mchunk - 1) / mchunk;
    const uint32_t nl = std::min(want, nchunks);
    /* per-step capacity: a step holds at most mchunk * Kb^depth rays, and at
     * most one sub-chunk */
    uint64_t cap_rays = mchunk;
    for (int d = 1; d <= E.maxiter && cap_rays < E.chunk; d++) cap_rays *= (uint64_t)std::max(E.Kb, 1);
    cap_rays = std::min<uint64_t>(cap_rays, E.chunk);
    const size_t L1 = (size_t)std::max(E.L, 1);
    while (E.lanes.size() < nl) E.lanes.emplace_back();
    if (!E.tables_ready) HIPCHK(c, hipEventCreateWithFlags(&E.tables_ready, hipEventDisableTiming));
    for (uint32_t l = 0; l < nl; l++) {
        Engine::Lane &ln = E.lanes[l];
        if (l == 0) ln.s = c->stream;
        else if (!ln.s) HIPCHK(c, hipStreamCreateWithFlags(&ln.s, hipStreamNonBlocking)); /* default priority */
        if (!ln.done) HIPCHK(c, hipEventCreateWithFlags(&ln.done, hipEventDisableTiming));
        if (!ln.s2) HIPCHK(c, hipStreamCreateWithFlags(&ln.s2, hipStreamNonBlocking)); /* default priority */
        if (!ln.evA) HIPCHK(c, hipEventCreateWithFlags(&ln.evA, hipEventDisableTiming));
        for (hipEvent_t &ev : ln.evB)
            if (!ev) HIPCHK(c, hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        if (ln.qlev.size() < (size_t)E.maxiter + 1) ln.qlev.resize((size_t)E.maxiter + 1);
        if ((rc = ensure(c, ln.qlev[0], (size_t)mchunk * sizeof(QRay)))) return rc;
        for (int b = 1; b <= E.maxiter; b++)
            if ((rc = ensure(c, ln.qlev[b], (size_t)cap_rays * E.K * sizeof(QRay)))) return rc;
        if ((rc = ensure(c, ln.hits, (size_t)cap_rays * sizeof(int4))) ||
            (rc = ensure(c, ln.direct, 2 * (size_t)cap_rays * 16)) ||
            (rc = ensure(c, ln.sq, 2 * (size_t)cap_rays * L1 * sizeof(SRay))) ||
            (rc = ensure(c, ln.scon, 2 * (size_t)cap_rays * L1 * sizeof(float4))) ||
            (rc = ensure(c, ln.rslot, 2 * (size_t)cap_rays * sizeof(uint32_t))) ||
            (rc = ensure(c, ln.occ, 2 * (size_t)cap_rays * L1)) ||
            (rc = ensure(c, ln.kids, (size_t)cap_rays * E.K * sizeof(QRay))) ||
            (rc = ensure(c, ln.kflag, (size_t)cap_rays * E.K)) || (rc = ensure(c, ln.acc, (size_t)mchunk * 16)) ||
            (E.photons && (rc = ensure(c, ln.greq, 2 * (size_t)cap_rays * sizeof(GatherReq)))) ||
            (E.photons && E.knn && (rc = ensure(c, ln.knnq, (2 * (size_t)cap_rays + 1) * 4))) ||
            (rc = ensure(c, ln.lvcnt, 64 * 4)) || (rc = ensure(c, ln.cmpb, 2 * CMP_BLOCKS * 4)))
            return rc;
        /* the k-nearest list starts empty (each step's k_accum empties it) */
        if (E.photons && E.knn) HIPCHK(c, hipMemsetAsync(ln.knnq.p, 0, 4, c->stream));
        ln.pstride = cap_rays;
    }

    HIPCHK(c, hipMemsetAsync(E.errf.p, 0, 4, c->stream));
    Tabs T{dptr<SceneDev>(E.dscenes), dptr<ReqDev>(E.dreqs), dptr<uint32_t>(E.preq), dptr<int>(E.nsamp)};
    if (!E.uni) HIPCHK(c, hipMemsetAsync(E.slotrays.p, 0, (size_t)E.hi * 8, c->stream));
    const uint32_t *A = dptr<uint32_t>(E.act[E.cur]);
    /* the other lanes start after the tables, admissions and clears above */
    if (nl > 1) HIPCHK(c, hipEventRecord(E.tables_ready, c->stream));
    for (uint32_t l = 1; l < nl; l++) HIPCHK(c, hipStreamWaitEvent(E.lanes[l].s, E.tables_ready, 0));
    auto finish = E.uni ? k_pass_finish<true> : k_pass_finish<false>;
    /* a level may outgrow one sub-chunk: then each chunk's level counts are
     * read after its descent and the remaining sub-chunks run before the
     * lane takes its next chunk (never needed when mchunk * Kb^maxiter fits) */
    const bool chain = chain_pass(E);
    const bool ovf = !chain && may_overflow(E, mchunk);
    auto finish_chunk = [&](Engine::Lane &ln, uint32_t c0, uint32_t m) {
        finish<<<blocks(m, 256), 256, 0, ln.s>>>(T, A, c0, m, dptr<float4>(ln.acc), dptr<float4>(E.fb),
                                                 dptr<float4>(E.subpix), dptr<int>(E.nsamp), dptr<uint8_t>(E.keep));
    };
    if (!ovf) {
        for (uint32_t g0 = 0; g0 < nchunks; g0 += nl) {
            const uint32_t g1 = std::min(nchunks, g0 + nl);
            for (uint32_t j = g0; j < g1; j++) {
                const uint32_t c0 = j * mchunk, m = std::min(mchunk, E.nact - c0);
                Engine::Lane &ln = E.lanes[j - g0];
                Level LV{&E, &ln, T, A, c0};
                if ((rc = chain ? launch_chain(LV, m) : launch_chunk(LV, m, nl == 1))) return rc;
                finish_chunk(ln, c0, m);
            }
            HIPCHK(c, hipGetLastError());
        }
    } else {
        /* Chunk j runs on lane j % nl.  A lane's level counts are read
         * after its descent (finish_levels), and the lane takes its next
         * chunk as soon as that chunk is finished, in the order the chunks
         * were started: no lane waits while the host reads another's. */
        std::deque<std::pair<uint32_t, uint32_t>> fifo; /* (lane, chunk) in start order */
        uint32_t next = 0;
        auto start = [&](uint32_t l) -> int {
            const uint32_t c0 = next * mchunk, m = std::min(mchunk, E.nact - c0);
            Level LV{&E, &E.lanes[l], T, A, c0};
            int r = launch_chunk(LV, m, nl == 1);
            if (r) return r;
            fifo.emplace_back(l, next++);
            return SRAY_OK;
        };
        for (uint32_t l = 0; l < nl && next < nchunks; l++)
            if ((rc = start(l))) return rc;
        while (!fifo.empty()) {
            const uint32_t l = fifo.front().first, j = fifo.front().second;
            fifo.pop_front();
            const uint32_t c0 = j * mchunk, m = std::min(mchunk, E.nact - c0);
            Engine::Lane &ln = E.lanes[l];
            Level LV{&E, &ln, T, A, c0};
            if ((rc = finish_levels(LV, 0))) return rc;
            finish_chunk(ln, c0, m);
            HIPCHK(c, hipGetLastError());
            if (next < nchunks && (rc = start(l))) return rc;
        }
    }
    /* the pass end waits for every lane */
    for (uint32_t l = 1; l < nl; l++) {
        HIPCHK(c, hipEventRecord(E.lanes[l].done, E.lanes[l].s));
        HIPCHK(c, hipStreamWaitEvent(c->stream, E.lanes[l].done, 0));
    }
    /* The pass status comes back with one copy into pinned memory and one
     * synchronisation: the next active list's length (when some request can
     * take another sample), the per-request counts, the error flag and the
     * ray counters.  When no request can take another sample every pixel
     * stops here, so the images are finalised before that synchronisation. */
    Engine::PassStatus *H = E.hst;
    unsigned long long *dstats = dptr<unsigned long long>(E.stats);
    const size_t st_off = offsetof(Engine::PassStatus, stats);
    if (E.more) {
        if ((rc = select_flagged_async(c, A, dptr<uint8_t>(E.keep), dptr<uint32_t>(E.act[E.cur ^ 1]), E.nact,
                                       dptr<uint32_t>(c->nsel))))
            return rc;
        HIPCHK(c, hipMemcpyAsync(&H->nsel, c->nsel.p, 4, hipMemcpyDeviceToHost, c->stream));
        if (!E.uni) {
            /* per-request active counts of the new list, bounded by the old length */
            k_fill_u32<<<blocks(E.hi, 256), 256, 0, c->stream>>>(dptr<uint32_t>(E.slotcnt), E.hi, 0u);
            k_slot_count_dev<<<std::min<uint32_t>(blocks(E.nact, 256), 1024u), 256, 0, c->stream>>>(
                dptr<uint32_t>(E.act[E.cur ^ 1]), dptr<uint32_t>(c->nsel), dptr<uint32_t>(E.preq),
                dptr<uint32_t>(E.slotcnt));
            HIPCHK(c, hipMemcpyAsync(H->slotcnt, E.slotcnt.p, (size_t)E.hi * 4, hipMemcpyDeviceToHost, c->stream));
        }
    } else {
        std::vector<uint32_t> all;
        for (uint32_t s = 0; s < E.hi; s++)
            if (E.slot_job[s]) all.push_back(s);
        if ((rc = finalize_jobs(E, all))) return rc;
    }
    if (!E.uni) HIPCHK(c, hipMemcpyAsync(H->srays, E.slotrays.p, (size_t)E.hi * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(&H->errf, E.errf.p, 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync((char *)H + st_off, dstats, sizeof H->stats, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(H->rcounts, E.rcounts.p, sizeof H->rcounts, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const uint32_t nn = E.more ? H->nsel : 0u;
    if (E.more) E.cur ^= 1;
    E.nact = nn;
    const unsigned long long sec = H->stats[1] + H->stats[2] + H->stats[3] + H->stats[4];
    if (E.uni) H->srays[0] = sec - E.sec_seen;
    E.sec_seen = sec;
    bool host_copies = false;
    std::vector<uint32_t> stopped;
    for (uint32_t s = 0; s < E.hi; s++) {
        Job *j = E.slot_job[s];
        if (!j) continue;
        const uint32_t cnt = !E.more ? 0u : E.uni ? nn : H->slotcnt[s];
        j->prim += E.slot_active[s];
        j->rays += E.slot_active[s] + H->srays[s];
        if (H->errf) j->limit_hit = true;
        E.slot_active[s] = cnt;
        if (!cnt) stopped.push_back(s);
    }
    /* all pixels stopped: the images (src/render.cc:238-241) */
    if (E.more && !stopped.empty() && (rc = finalize_jobs(E, stopped))) return rc;
    for (uint32_t s : stopped) {
        Job *j = E.slot_job[s];
        host_copies = host_copies || j->host_out != nullptr;
        retire(E, s);
        finished.push_back(j);
    }
    /* images complete when the request is reported done (a plain render
     * synchronises on its own end event) */
    if (!finished.empty() && (sync_done || host_copies)) HIPCHK(c, hipStreamSynchronize(c->stream));
    if (host_copies)
        for (Job *j : finished) release(j->stage);
    return SRAY_OK;
}

Engine *engine_get(sray_gpu_ctx *c, Engine *&slot)
{
    if (!slot) {
        slot = new Engine;
        slot->c = c;
    }
    return slot;
}

void engine_free(Engine *E)
{
    if (!E) return;
    drop_graphs(*E);
    for (size_t l = 0; l < E->lanes.size(); l++) {
        Engine::Lane &ln = E->lanes[l];
        if (l > 0 && ln.s) {
            (void)hipStreamSynchronize(ln.s);
            (void)hipStreamDestroy(ln.s);
        }
        if (ln.done) (void)hipEventDestroy(ln.done);
        if (ln.s2) {
            (void)hipStreamSynchronize(ln.s2);
            (void)hipStreamDestroy(ln.s2);
        }
        if (ln.evA) (void)hipEventDestroy(ln.evA);
        for (hipEvent_t ev : ln.evB)
            if (ev) (void)hipEventDestroy(ev);
        DevBuf *lb[] = {&ln.hits, &ln.direct, &ln.sq, &ln.scon, &ln.rslot, &ln.occ, &ln.kids, &ln.kflag, &ln.acc,
                        &ln.greq, &ln.knnq, &ln.lvcnt, &ln.cmpb};
        for (DevBuf *b : lb) release(*b);
        for (DevBuf &b : ln.qlev) release(b);
    }
    if (E->tables_ready) (void)hipEventDestroy(E->tables_ready);
    DevBuf *bs[] = {&E->fb, &E->nsamp, &E->preq, &E->keep, &E->act[0], &E->act[1], &E->subpix, &E->stats,
                    &E->errf, &E->rcounts, &E->slotcnt, &E->slotrays, &E->dreqs, &E->dscenes};
    for (DevBuf *b : bs) release(*b);
    if (E->hst) (void)hipHostFree(E->hst);
    if (E->hreq_pin) (void)hipHostFree(E->hreq_pin);
    if (E->hscn_pin) (void)hipHostFree(E->hscn_pin);
    delete E;
}

/* sray_gpu_render: one request through the plain engine */
int do_render(sray_gpu_ctx *c, uint32_t scene_id, const sray_render_params *p, float *d_out, sray_stats *st)
{
    int rc = validate_params(c, p);
    if (rc) return rc;
    Engine &E = *engine_get(c, c->plain);
    if ((rc = ensure(c, E.stats, 8 * sizeof(unsigned long long))) || (rc = ensure(c, E.errf, 4)) ||
        (rc = ensure(c, E.rcounts, 2 * NCOUNT * sizeof(unsigned long long))))
        return rc;
    HIPCHK(c, hipMemsetAsync(E.rcounts.p, 0, 2 * NCOUNT * sizeof(unsigned long long), c->stream));
    HIPCHK(c, hipMemsetAsync(E.stats.p, 0, 8 * sizeof(unsigned long long), c->stream));
    HIPCHK(c, hipMemsetAsync(E.errf.p, 0, 4, c->stream));
    E.sec_seen = 0;
    E.counting = (p->flags & SRAY_RENDER_COUNT) != 0;
    Job j;
    j.scene_id = scene_id;
    j.p = *p;
    j.dev_out = d_out;
    j.t_submit = Clock::now();
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    emit_status(c, 'f', 1, 1, p->x1 - p->x0, std::max(1, (int)c->tune.lanes));
    if ((rc = admit(E, &j))) return rc;
    uint64_t prim = 0, samples = 0;
    std::vector<Job *> fin;
    while (j.state != 2 && rc == SRAY_OK) {
        samples += E.nact;
        rc = run_pass(E, fin, false);
        for (Job *f : fin) f->state = 2;
        fin.clear();
    }
    if (rc) {
        abort_all(E);
        return rc;
    }
    prim = j.prim;
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipEventSynchronize(c->ev1));
    prof_resolve(c);
    /* counters as the last pass read them back (nothing after it adds to them) */
    const unsigned long long *hs = E.hst->stats, *rcn = E.hst->rcounts;
    if (j.limit_hit) return fail(c, SRAY_ERR_LIMIT, "ior stack deeper than %d media", IOR_DEPTH);
    float dev_ms = 0;
    HIPCHK(c, hipEventElapsedTime(&dev_ms, c->ev0, c->ev1));
    emit_status(c, 'e', p->x0, p->y0, p->x1 - p->x0, p->y1 - p->y0);
    emit_status(c, 'd', (int)dev_ms, 0, 0, 0);
    if (st) {
        memset(st, 0, sizeof *st);
        const float ms = dev_ms;
        st->render_ms = ms;
        st->rays_primary = prim;
        st->rays_shadow = hs[1];
        st->rays_reflect = hs[2];
        st->rays_refract = hs[3];
        st->rays_diffuse = hs[4];
        st->samples = samples;
        sray_trace_counts *tc[2] = {&st->closest, &st->shadow};
        for (int k = 0; k < 2; k++) {
            tc[k]->nodes = rcn[k * NCOUNT + 0];
            tc[k]->spheres = rcn[k * NCOUNT + 1];
            tc[k]->cylinders = rcn[k * NCOUNT + 2];
            tc[k]->triangles = rcn[k * NCOUNT + 3];
            tc[k]->item_refs = rcn[k * NCOUNT + 4];
            tc[k]->mesh_enters = rcn[k * NCOUNT + 5];
            tc[k]->item_boxes = rcn[k * NCOUNT + 6];
        }
    }
    return SRAY_OK;
}

/* ------------------------------------------------------------------ */
/* submit / poll: continuous batching on a worker thread                */
/* ------------------------------------------------------------------ */
void serve_loop(sray_gpu_ctx *c, Engine *E)
{
    (void)hipSetDevice(c->device);
    std::vector<Job *> fin;
    for (;;) {
        {
            std::unique_lock<std::mutex> lk(E->qmu);
            E->qcv.wait(lk, [&] { return E->stop || !E->pending.empty() || E->live > 0; });
            if (E->stop) break;
        }
        /* API calls waiting for the context (submit, upload, ...) go first */
        while (c->api_waiting.load() > 0) std::this_thread::yield();
        std::lock_guard<std::recursive_mutex> g(c->mu);
        /* admission: first come first served while the requests fit the cap */
        for (;;) {
            Job *j = nullptr;
            {
                std::lock_guard<std::mutex> lk(E->qmu);
                if (E->pending.empty()) break;
                j = E->pending.front();
            }
            uint64_t need = job_bytes(*j);
            uint64_t used = admitted_bytes(c);
            if (used + need > c->admission_cap && E->live > 0) break; /* wait for requests to finish */
            {
                std::lock_guard<std::mutex> lk(E->qmu);
                E->pending.pop_front();
            }
            int rc = used + need > c->admission_cap
                         ? fail(c, SRAY_ERR_NOMEM, "request needs %llu bytes beyond the admission cap",
                                (unsigned long long)need)
                         : admit(*E, j);
            if (rc) job_done(*E, j, rc, c->err.c_str());
            if (E->live == MAX_SLOTS) break;
        }
        if (E->live == 0) continue;
        int rc = run_pass(*E, fin, true);
        if (c->profiling) prof_resolve(c);
        if (rc) {
            std::string msg = c->err;
            for (uint32_t s = 0; s < E->hi; s++)
                if (E->slot_job[s]) fin.push_back(E->slot_job[s]);
            abort_all(*E);
            for (Job *j : fin) job_done(*E, j, rc, msg.c_str());
        } else {
            for (Job *j : fin)
                job_done(*E, j, j->limit_hit ? SRAY_ERR_LIMIT : SRAY_OK,
                         j->limit_hit ? "ior stack deeper than the queued rays carry" : nullptr);
        }
        fin.clear();
    }
}

/* queues n requests at once: all are validated first, then all enter the
 * queue together (the worker admits them at one pass boundary when they fit) */
int do_submit_many(sray_gpu_ctx *c, const sray_request *r, uint32_t n, float *const *rgba, uint64_t *tickets)
{
    std::vector<std::unique_ptr<Job>> jobs(n);
    int rc;
    {
        ApiLock g(c);
        for (uint32_t k = 0; k < n; k++) {
            const sray_request &q = r[k];
            if (!rgba[k]) return fail(c, SRAY_ERR_ARG, "request %u: output is NULL", k);
            if (q.flags & ~SRAY_REQUEST_DEVICE_OUT)
                return fail(c, SRAY_ERR_ARG, "request %u: unknown request flags 0x%x", k, q.flags);
            if ((rc = validate_params(c, &q.params))) return rc;
            if (!find_scene(c, q.scene_id)) return fail(c, SRAY_ERR_NOTFOUND, "unknown scene %u", q.scene_id);
            if (q.use_camera && !(q.camera.vfov > 0.0))
                return fail(c, SRAY_ERR_ARG, "request %u camera: vfov must be > 0", k);
            std::unique_ptr<Job> j(new Job);
            j->scene_id = q.scene_id;
            j->p = q.params;
            j->cam = q.camera;
            j->use_cam = q.use_camera != 0;
            if (q.flags & SRAY_REQUEST_DEVICE_OUT)
                j->dev_out = rgba[k];
            else
                j->host_out = rgba[k];
            jobs[k] = std::move(j);
        }
        Engine *E = engine_get(c, c->serve);
        if (!E->started) {
            E->started = true;
            E->worker = std::thread(serve_loop, c, E);
        }
        /* the serving engine's counters accumulate over its lifetime */
        if ((rc = ensure_zero(c, E->stats, 8 * sizeof(unsigned long long))) || (rc = ensure_zero(c, E->errf, 4)) ||
            (rc = ensure_zero(c, E->rcounts, 2 * NCOUNT * sizeof(unsigned long long))))
            return rc;
    }
    Engine *E = c->serve;
    std::lock_guard<std::mutex> lk(E->qmu);
    const Clock::time_point now = Clock::now();
    for (uint32_t k = 0; k < n; k++) {
        Job *j = jobs[k].get();
        j->ticket = E->next_ticket++;
        j->t_submit = now;
        tickets[k] = j->ticket;
        E->pending.push_back(j);
        E->jobs[j->ticket] = std::move(jobs[k]);
    }
    E->qcv.notify_all();
    return SRAY_OK;
}

int do_submit(sray_gpu_ctx *c, const sray_request *r, float *rgba, uint64_t *ticket)
{
    return do_submit_many(c, r, 1, &rgba, ticket);
}

int do_poll(sray_gpu_ctx *c, uint64_t ticket, int wait, sray_result *out)
{
    Engine *E = c->serve;
    if (!E) return fail(c, SRAY_ERR_NOTFOUND, "unknown ticket %llu", (unsigned long long)ticket);
    std::unique_lock<std::mutex> lk(E->qmu);
    auto it = E->jobs.find(ticket);
    if (it == E->jobs.end()) return SRAY_ERR_NOTFOUND;
    Job *j = it->second.get();
    if (j->state != 2) {
        if (!wait) return SRAY_ERR_PENDING;
        E->qcv.wait(lk, [&] { return j->state == 2; });
    }
    if (out) {
        out->rgba = j->host_out ? j->host_out : j->dev_out;
        out->rays = j->rays;
        out->latency_ms = j->latency_ms;
        out->queue_ms = j->queue_ms;
    }
    int rc = j->rc;
    std::string err = j->err;
    E->jobs.erase(it);
    lk.unlock();
    if (rc) {
        std::lock_guard<std::recursive_mutex> g(c->mu);
        c->err = err;
    }
    return rc;
}

void engine_destroy(sray_gpu_ctx *c)
{
    Engine *S = c->serve;
    if (S && S->started) {
        {
            std::lock_guard<std::mutex> lk(S->qmu);
            S->stop = true;
            S->qcv.notify_all();
        }
        S->worker.join();
    }
    std::lock_guard<std::recursive_mutex> g(c->mu);
    (void)hipStreamSynchronize(c->stream);
    if (S) {
        std::lock_guard<std::mutex> lk(S->qmu);
        for (auto &kv : S->jobs)
            if (kv.second->state != 2) {
                kv.second->rc = SRAY_ERR_ARG;
                kv.second->state = 2;
            }
        S->qcv.notify_all();
    }
    engine_free(c->plain);
    engine_free(S);
    c->plain = c->serve = nullptr;
}

uint64_t engine_bytes(sray_gpu_ctx *c)
{
    return (c->plain ? c->plain->bytes() : 0) + (c->serve ? c->serve->bytes() : 0);
}

int engine_pmap_busy(sray_gpu_ctx *c, uint32_t pmap_id)
{
    Engine *E = c->serve;
    if (!E || !pmap_id) return SRAY_OK;
    for (uint32_t s = 0; s < E->hi; s++)
        if (E->slot_job[s] && (E->slot_job[s]->pmaps[0] == pmap_id || E->slot_job[s]->pmaps[1] == pmap_id))
            return fail(c, SRAY_ERR_PENDING, "photon map %u is read by requests in flight", pmap_id);
    return SRAY_OK;
}

int engine_scene_busy(sray_gpu_ctx *c, uint32_t scene_id)
{
    Engine *E = c->serve;
    if (!E) return SRAY_OK;
    std::lock_guard<std::mutex> lk(E->qmu);
    for (auto &kv : E->jobs)
        if (kv.second->state != 2 && kv.second->scene_id == scene_id)
            return fail(c, SRAY_ERR_PENDING, "scene %u has requests in flight", scene_id);
    return SRAY_OK;
}

int check_tuning(sray_gpu_ctx *c, const sray_tuning &t)
{
    bool ok = (t.trace_block == 64 || t.trace_block == 128 || t.trace_block == 256) &&
              (t.tile_w == 8 || t.tile_w == 16 || t.tile_w == 32 || t.tile_w == 64) &&
              (t.gather_lanes == 8 || t.gather_lanes == 16 || t.gather_lanes == 32) && t.chunk_log2 >= 16 &&
              t.chunk_log2 <= 21 && (t.gather_cell == 0.25f || t.gather_cell == 0.5f || t.gather_cell == 1.0f || t.gather_cell == 2.0f) &&
              (t.graphs == 0 || t.graphs == 1) && (t.persistent >= 0 && t.persistent <= 2) &&
              (t.lanes == 1 || t.lanes == 2 || t.lanes == 4) && t.sparse_from >= 0 && t.sparse_from <= 64 &&
              (t.chain == 0 || t.chain == 1) && (t.sample_runs == 0 || t.sample_runs == 1);
    return ok ? SRAY_OK : fail(c, SRAY_ERR_ARG, "tuning value outside the supported set");
}

/* median device time of `rounds` renders of params (after one untimed render) */
int time_render(sray_gpu_ctx *c, uint32_t scene_id, const sray_render_params *p, int rounds, float *d_out, double *ms)
{
    sray_stats st;
    int rc = do_render(c, scene_id, p, d_out, &st);
    if (rc) return rc;
    std::vector<double> v;
    for (int r = 0; r < rounds; r++) {
        if ((rc = do_render(c, scene_id, p, d_out, &st))) return rc;
        v.push_back(st.render_ms);
    }
    std::sort(v.begin(), v.end());
    *ms = v[v.size() / 2];
    return SRAY_OK;
}

int do_autotune(sray_gpu_ctx *c, uint32_t scene_id, const sray_render_params *p, int rounds, sray_tuning *best,
                double *best_ms)
{
    int rc = validate_params(c, p);
    if (rc) return rc;
    if (rounds < 1) rounds = 1;
    size_t bytes = (size_t)(p->x1 - p->x0) * (size_t)(p->y1 - p->y0) * 16;
    DevBuf out;
    if ((rc = ensure(c, out, bytes))) return rc;
    sray_tuning cur = c->tune;
    double cur_ms = 0;
    if ((rc = time_render(c, scene_id, p, rounds, dptr<float>(out), &cur_ms))) {
        release(out);
        return rc;
    }
    static const int blocks_[] = {64, 128, 256}, tiles[] = {8, 16, 32, 64}, lanes[] = {8, 16, 32}, chunks[] = {18, 19, 20, 21};
    static const float cells[] = {0.25f, 0.5f, 1.0f, 2.0f};
    bool photons = p->caust_photons || p->gi_photons;
    static const int nlanes[] = {1, 2, 4}, sparse[] = {0, 1, 2, 64};
    for (int knob = 0; knob < 11; knob++) {
        if ((knob == 2 || knob == 4) && !photons) continue;
        int nv = knob == 0 ? 3 : knob == 1 ? 4 : knob == 2 ? 3 : knob == 3 ? 4 : knob == 4 ? 4 : knob == 6 ? 3
               : knob == 7 ? 3 : knob == 8 ? 4 : 2;
        for (int k = 0; k < nv; k++) {
            sray_tuning t = cur;
            if (knob == 0) t.trace_block = blocks_[k];
            else if (knob == 1) t.tile_w = tiles[k];
            else if (knob == 2) t.gather_lanes = lanes[k];
            else if (knob == 3) t.chunk_log2 = chunks[k];
            else if (knob == 4) t.gather_cell = cells[k];
            else if (knob == 5) t.graphs = k;
            else if (knob == 6) t.persistent = k;
            else if (knob == 7) t.lanes = nlanes[k];
            else if (knob == 8) t.sparse_from = sparse[k];
            else if (knob == 9) t.chain = k;
            else t.sample_runs = k;
            if (!memcmp(&t, &cur, sizeof t)) continue;
            c->tune = t;
            double ms = 0;
            if ((rc = time_render(c, scene_id, p, rounds, dptr<float>(out), &ms))) {
                c->tune = cur;
                release(out);
                return rc;
            }
            /* a change must beat the current choice by more than the
             * round-to-round noise of the timings (about 1%), and again
             * when both are timed once more, one right after the other
             * (one lucky median would otherwise move the tuning) */
            if (ms < cur_ms * 0.985) {
                double again_cur = 0, again_t = 0;
                c->tune = cur;
                if (!(rc = time_render(c, scene_id, p, rounds, dptr<float>(out), &again_cur))) {
                    c->tune = t;
                    rc = time_render(c, scene_id, p, rounds, dptr<float>(out), &again_t);
                }
                if (rc) {
                    c->tune = cur;
                    release(out);
                    return rc;
                }
                if (again_t < again_cur * 0.985) {
                    cur = t;
                    cur_ms = again_t;
                } else {
                    cur_ms = again_cur;
                }
            }
        }
        c->tune = cur;
    }
    release(out);
    if (best) *best = cur;
    if (best_ms) *best_ms = cur_ms;
    return SRAY_OK;
}

} // namespace srt

/* ================================================================== */
/* C ABI                                                                */
/* ================================================================== */
extern "C" {

int sray_gpu_primary_rays_device(sray_gpu_ctx *c, uint32_t scene_id, const sray_render_params *p, int sample,
                                 sray_ray *d_rays)
{
    if (!c || !p || !d_rays || sample < 0) return SRAY_ERR_ARG;
    ApiLock lk(c);
    GpuScene *g = find_scene(c, scene_id);
    if (!g) return fail(c, SRAY_ERR_NOTFOUND, "unknown scene %u", scene_id);
    if (p->width <= 0 || p->height <= 0 || p->x0 < 0 || p->y0 < 0 || p->x1 > p->width || p->y1 > p->height ||
        p->x0 >= p->x1 || p->y0 >= p->y1)
        return fail(c, SRAY_ERR_ARG, "bad pixel rectangle");
    RenderDev R;
    fill_render_dev(p, R);
    uint32_t n = (uint32_t)R.W * (uint32_t)(p->y1 - p->y0);
    k_primary<<<blocks(n, 256), 256, 0, c->stream>>>(cam_dev(g->camera, g->dev.cam_tab, g->dev.cam_t0, g->dev.cam_n), R,
                                                     sample, d_rays, n);
    HIPCHK(c, hipGetLastError());
    return SRAY_OK;
}

int sray_gpu_render_device(sray_gpu_ctx *c, uint32_t scene_id, const sray_render_params *p, float *d_rgba,
                           sray_stats *st)
{
    if (!c || !d_rgba) return SRAY_ERR_ARG;
    ApiLock lk(c);
    if (hipSetDevice(c->device) != hipSuccess) return fail(c, SRAY_ERR_HIP, "hipSetDevice failed");
    return do_render(c, scene_id, p, d_rgba, st);
}

int sray_gpu_render(sray_gpu_ctx *c, uint32_t scene_id, const sray_render_params *p, float *rgba, sray_stats *st)
{
    if (!c || !rgba) return SRAY_ERR_ARG;
    ApiLock lk(c);
    if (hipSetDevice(c->device) != hipSuccess) return fail(c, SRAY_ERR_HIP, "hipSetDevice failed");
    int rc = validate_params(c, p);
    if (rc) return rc;
    size_t bytes = (size_t)(p->x1 - p->x0) * (size_t)(p->y1 - p->y0) * 4 * sizeof(float);
    DevBuf out;
    hipError_t e = hipMalloc(&out.p, bytes);
    if (e != hipSuccess) return fail(c, SRAY_ERR_NOMEM, "hipMalloc(%zu) failed", bytes);
    out.bytes = bytes;
    rc = do_render(c, scene_id, p, (float *)out.p, st);
    if (rc == SRAY_OK) {
        e = hipMemcpy(rgba, out.p, bytes, hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = fail(c, SRAY_ERR_HIP, "hipMemcpy failed: %s", hipGetErrorString(e));
    }
    release(out);
    return rc;
}

int sray_gpu_submit(sray_gpu_ctx *c, const sray_request *r, float *rgba, uint64_t *ticket)
{
    if (!c || !r || !rgba || !ticket) return SRAY_ERR_ARG;
    return do_submit(c, r, rgba, ticket);
}

int sray_gpu_submit_many(sray_gpu_ctx *c, const sray_request *r, uint32_t n, float *const *rgba, uint64_t *tickets)
{
    if (!c || (n && (!r || !rgba || !tickets))) return SRAY_ERR_ARG;
    if (n == 0) return SRAY_OK;
    return do_submit_many(c, r, n, rgba, tickets);
}

int sray_gpu_poll(sray_gpu_ctx *c, uint64_t ticket, int wait, sray_result *out)
{
    if (!c) return SRAY_ERR_ARG;
    return do_poll(c, ticket, wait, out);
}

int sray_gpu_set_tuning(sray_gpu_ctx *c, const sray_tuning *t)
{
    if (!c || !t) return SRAY_ERR_ARG;
    ApiLock lk(c);
    int rc = check_tuning(c, *t);
    if (rc) return rc;
    c->tune = *t;
    return SRAY_OK;
}

int sray_gpu_get_tuning(sray_gpu_ctx *c, sray_tuning *t)
{
    if (!c || !t) return SRAY_ERR_ARG;
    ApiLock lk(c);
    *t = c->tune;
    return SRAY_OK;
}

int sray_gpu_autotune(sray_gpu_ctx *c, uint32_t scene_id, const sray_render_params *p, int rounds, sray_tuning *best,
                      double *best_ms)
{
    if (!c || !p) return SRAY_ERR_ARG;
    ApiLock lk(c);
    if (hipSetDevice(c->device) != hipSuccess) return fail(c, SRAY_ERR_HIP, "hipSetDevice failed");
    return do_autotune(c, scene_id, p, rounds, best, best_ms);
}

int sray_gpu_set_admission_cap(sray_gpu_ctx *c, uint64_t bytes)
{
    if (!c) return SRAY_ERR_ARG;
    ApiLock lk(c);
    c->admission_cap = bytes;
    return SRAY_OK;
}

} /* extern "C" */
