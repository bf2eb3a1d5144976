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

#include "pmap.h"
#include "traverse.h"

namespace srd {

struct alignas(16) QRay {
    float4 o;    /* origin, w = time (int bits) */
    float4 d;    /* direction, w = energy */
    float4 w;    /* throughput weight rgb, w = current ior */
    uint4 m;     /* x = sample slot (chunk-local), y = path id, z = iter,
                  * w = nstack | flags << 8 | request slot << 16 */
    float4 stk;  /* ior stack */
};

/* a shadow ray; its contribution (rgb if unoccluded, w = 1 for a used slot,
 * 0 for an unused one) sits in a parallel float4 array, the only part the
 * accumulation reads */
struct alignas(16) SRay {
    float4 o;    /* origin, w = time (int bits) */
    float4 d;    /* direction (unnormalised: segment to the light), w = request slot (bits) */
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
};

/* photon maps bound for shading (caustic irradiance, GI radiance) */
struct ShadePmaps {
    PmapGrid caust, glob;
    float radius; /* inclusion radius (float, as kd_nearest_range3f) */
    int32_t k;    /* > 0: the k nearest within the radius */
    double scale; /* 1 / (pi r^2) with the double gather distance */
};

/* a photon gather a shaded ray needs, done by its own kernel after shading:
 * mode 1 adds W * caustic irradiance to the ray's direct term
 * (src/shader.cc:231-236), mode 2 sets it to W * (GI radiance * kd) for a
 * diffuse GI ray (src/shader.cc:119-123) */
struct alignas(16) GatherReq {
    float4 pos; /* hit point, w = mode (bits; 0 = none) */
    float4 nrm; /* shading normal (faces the ray) */
    float4 w;   /* throughput weight */
    float4 kd;  /* diffuse colour (mode 2) */
};

/* the camera a request renders through (Camera::get_primary_ray inputs) */
struct CamDev {
    float cam[12];   /* camera -> world */
    float fz;        /* 1/tan(vfov/2), from double on the host */
    float max_dist;
    int32_t shutter;
    int32_t pad;
    /* animated camera: xform (12 floats) per msec from t0, or null */
    const float *tab;
    int32_t t0, n;
};

/* one request of a batch: its parameters, camera, photon maps and where its
 * pixels live in the engine's pixel arena */
struct ReqDev {
    RenderDev P;
    CamDev C;
    ShadePmaps PM;
    uint32_t scene;    /* index into the batch's scene table */
    uint32_t pix_base; /* first arena pixel */
    uint64_t sub_base; /* first per-sample record (max_samples per pixel) */
    uint32_t run;      /* samples the first pass takes of each pixel (a power of
                          two <= 16 that every pixel takes: min_samples, at
                          least 2, at most max_samples); later passes take one */
    uint32_t pad_;
};

/* An active-list entry: the arena pixel in its low 28 bits and, in the top
 * four, the sample's offset from the pixel's samples taken.  A pixel of a
 * request with run r > 1 has r consecutive entries (offsets 0..r-1) in one
 * aligned group of 16 entries, so a primary chunk never splits them. */
constexpr uint32_t ENT_SHIFT = 28, ENT_PIX = (1u << ENT_SHIFT) - 1u;
__host__ __device__ __forceinline__ uint32_t ent_pix(uint32_t e) { return e & ENT_PIX; }
__host__ __device__ __forceinline__ uint32_t ent_off(uint32_t e) { return e >> ENT_SHIFT; }

/* kernel view of a batch */
struct Tabs {
    const SceneDev *scenes;
    const ReqDev *reqs;
    const uint32_t *preq; /* arena pixel -> request slot */
    const int *nsamp;     /* arena pixel -> samples taken */
};

/* the rays of one level step: [off, off + n) of a level queue whose full
 * count *cnt the device wrote (the previous level's compaction, or the ray
 * generator), at most cap of them.  Kernels read the count themselves, so
 * a bounce level's launches need nothing from the host. */
struct StepN {
    const uint32_t *cnt;
    uint32_t off, cap;
};
__device__ __forceinline__ uint32_t step_n(const StepN &s)
{
    const uint32_t c = *s.cnt;
    return c > s.off ? min(c - s.off, s.cap) : 0u;
}

constexpr int SLOT_SHIFT = 16;
constexpr uint32_t MAX_SLOTS = 1024; /* requests in flight in one engine */

} // namespace srd

#endif
