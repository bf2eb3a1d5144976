/*
 * render.h — wavefront replacement of the recursive shader
 * (render_block, src/render.cc:194-244; Scene::trace_ray, src/scene.cc:551-569;
 * shade_phong + calc_lighting, src/shader.cc:83-283; shade_undef :36-71).
 *
 * The recursion is linear in the child radiances, so each queued ray carries
 * a throughput weight; a bounce level is: closest-hit trace -> shade (direct
 * terms, one shadow ray per light sample, reflect/refract/diffuse children
 * into fixed slots) -> any-hit trace of the shadow rays -> per-sample
 * accumulation -> stable compaction of the children into the next level.
 * Queues stay ordered by sample slot, so every sample's contributions are
 * summed by one thread in a fixed order (deterministic, no float atomics).
 */
#ifndef SRAY_RENDER_H_
#define SRAY_RENDER_H_

#include "traverse.h"

namespace srd {

struct alignas(16) QRay {
    float4 o;    /* origin, w = time (int bits) */
    float4 d;    /* direction, w = energy */
    float4 w;    /* throughput weight rgb, w = current ior */
    uint4 m;     /* x = sample slot (chunk-local), y = path id, z = iter, w = nstack | flags<<8 */
    float4 stk;  /* ior stack */
};

struct alignas(16) SRay {
    float4 o;    /* origin, w = time (int bits) */
    float4 d;    /* direction (unnormalised: segment to the light), w = 1 valid / 0 unused */
    float4 c;    /* contribution rgb if unoccluded */
};

constexpr uint32_t QF_DIFFUSE = 1u;  /* diffuse GI ray (IS_DIFFUSE_RAY, src/shader.cc:80-81) */

struct RenderDev {
    int32_t width, height, x0, y0, x1, y1, W;
    int32_t min_samples, max_samples;
    float max_var, min_energy;
    int32_t iter, shadow_samples, diffuse_samples, frame_time, mblur, caust, gi;
    uint32_t flags;
    uint64_t seed;
    float aspect;     /* (float)((double)width / height) */
    int32_t L;        /* shadow slots per ray */
    int32_t K;        /* child slots per ray */
};

/* photon-map gather inputs for shading (null when the render has none) */
struct PmapDev {
    const float4 *pos;     /* x,y,z, w = photon index (bits) — kd-tree order */
    const float4 *dir;     /* dir.xyz, w unused */
    const float4 *nrm;     /* norm.xyz */
    const float4 *col;     /* col.rgb */
    const float4 *split;   /* per kd node: split value in .x, axis in .y bits */
    uint32_t n;
    uint32_t depth;
};

} // namespace srd

#endif
