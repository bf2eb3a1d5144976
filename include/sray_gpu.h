/*
 * sray_gpu.h — C ABI of the MI355X (gfx950) ray-tracing hot path.
 *
 * Plain C, POD structs, plain pointers and sizes, int status codes (0 = ok,
 * < 0 = error, text from sray_gpu_last_error()).  No torch, no C++ types.
 *
 * Every entry point names the reference interface it replaces (S-ray 0.1,
 * file:line).  Host buffers are caller-owned and copied in/out; scenes and
 * photon maps uploaded to a context are owned by the context until released.
 * One context = one HIP stream; calls on one context are serialised by an
 * internal mutex, several contexts may share a device.
 */
#ifndef SRAY_GPU_H_
#define SRAY_GPU_H_

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ */
/* status codes                                                        */
/* ------------------------------------------------------------------ */
#define SRAY_OK              0
#define SRAY_ERR_ARG        -1   /* invalid argument / size */
#define SRAY_ERR_HIP        -2   /* a HIP runtime call failed */
#define SRAY_ERR_NOMEM      -3   /* allocation failed or admission cap exceeded */
#define SRAY_ERR_NOTFOUND   -4   /* unknown scene / photon map / ticket */
#define SRAY_ERR_LIMIT      -5   /* a compiled limit (stack depth, ior stack) is exceeded */
#define SRAY_ERR_PARSE      -6   /* scene XML or a photon dump could not be parsed/loaded */
#define SRAY_ERR_PENDING    -7   /* sray_gpu_poll: request not finished yet */
#define SRAY_ERR_NODEVICE   -8   /* no HIP device */

/* ------------------------------------------------------------------ */
/* constants shared with the reference                                 */
/* ------------------------------------------------------------------ */
/* Rays are segments t in (epsilon, 1]; the direction carries the length
 * (RAY_MAGNITUDE, src/opt.h:21).  ERROR_MARGIN / XSMALL_NUMBER live in the
 * un-vendored libvmath; these are the values this build fixes. */
#define SRAY_RAY_MAGNITUDE   10000.0
#define SRAY_ERROR_MARGIN    1e-6
#define SRAY_XSMALL_NUMBER   1e-8

/* object types, src/object.cc:28-29 */
#define SRAY_OBJ_SPHERE      0
#define SRAY_OBJ_CYLINDER    1
#define SRAY_OBJ_MESH        2

/* light types, src/light.cc:23-29 */
#define SRAY_LIGHT_POINT     0
#define SRAY_LIGHT_SPHERE    1
#define SRAY_LIGHT_BOX       2

/* material flags: which attributes the <material> defines
 * (Material::have_attribute, src/material.cc:170-173) and its shader
 * (src/shader.cc:28-34). */
#define SRAY_MAT_HAS_SPECULAR  0x1u
#define SRAY_MAT_HAS_REFLECT   0x2u
#define SRAY_MAT_HAS_REFRACT   0x4u
#define SRAY_MAT_PHONG         0x8u   /* shader="phong"; otherwise shade_undef */
#define SRAY_MAT_HAS_NORMAL    0x10u  /* a "normal" attribute (bump map when mapped) */

/* texture maps of material attributes (MatAttrib::tex, src/material.cc:64-72) */
enum { SRAY_MAP_DIFFUSE, SRAY_MAP_SPECULAR, SRAY_MAP_SHININESS, SRAY_MAP_REFLECT, SRAY_MAP_REFRACT,
       SRAY_MAP_IOR, SRAY_MAP_NORMAL, SRAY_NUM_MAPS };
#define SRAY_TEX_FILTER_LINEAR  0     /* 5-tap cross average (Texture(), src/texture.cc:21-25) */
#define SRAY_TEX_FILTER_NEAREST 1
#define SRAY_TEX_WRAP_REPEAT    0
#define SRAY_TEX_WRAP_CLAMP     1

/* ------------------------------------------------------------------ */
/* flat scene description (host side, what Scene::load produces)       */
/* ------------------------------------------------------------------ */
/* Matrices are row-major 3x4 (rows x,y,z of a 4x4 affine matrix whose last
 * row is 0 0 0 1), applied to column vectors: p' = M * (p, 1). */
typedef struct sray_object_desc {
    int32_t type;          /* SRAY_OBJ_* */
    int32_t material;      /* index into materials, -1 = Scene::default_mat */
    int32_t mesh;          /* index into meshes (SRAY_OBJ_MESH), else -1 */
    float   radius;        /* sphere / cylinder radius */
    float   start, end;    /* cylinder extent along local y */
    float   xform[12];     /* object -> world at the frame time */
    float   inv_xform[12]; /* world -> object */
    /* linear translation over [motion_t0, motion_t1] msec (motion blur):
     * xform(t) = xform with translation + motion * s, s = clamp((t-t0)/(t1-t0)).
     * all zero = static object. */
    float   motion[3];
    int32_t motion_t0, motion_t1;
    /* keyframed motion (XFormNode tracks: rotation, scale, pivot, more
     * than two position keys), evaluated on the host at every msec of the
     * scene's [frame_t0, frame_t1]: 24 floats per msec (xform, then
     * inv_xform), frame_t1 - frame_t0 + 1 rows; NULL = xform / motion
     * above.  A ray outside the interval takes the nearest row. */
    const float *time_xforms;
} sray_object_desc;

typedef struct sray_mesh_desc {
    uint32_t tri_start;    /* first triangle in the scene triangle arrays */
    uint32_t tri_count;
} sray_mesh_desc;

typedef struct sray_material_desc {
    float    diffuse[4];   /* kd */
    float    specular[4];  /* ks */
    float    shininess;    /* default 1 (Material::def_attr, src/material.cc:77) */
    float    reflect;
    float    refract;
    float    ior;
    uint32_t flags;        /* SRAY_MAT_* */
    float    normal;       /* the "normal" attribute's value (bump scale) */
    /* texture of each attribute (SRAY_MAP_*): 1 + index into the scene's
     * textures, 0 = none (a zeroed descriptor has no maps).  The attribute
     * is texture x value (MatAttrib::get_color, src/material.inl:23-28). */
    int32_t  map[SRAY_NUM_MAPS];
    int32_t  pad_;
} sray_material_desc;

typedef struct sray_texture_desc {
    int32_t      width, height;
    const float *texels;   /* width * height RGBA floats, row-major, top row first */
    int32_t      filter;   /* SRAY_TEX_FILTER_* */
    int32_t      wrap;     /* SRAY_TEX_WRAP_* */
    float        xform[12];/* texture-coordinate matrix (a Texture is an XFormNode;
                              identity as loaded from a scene) */
} sray_texture_desc;

typedef struct sray_light_desc {
    int32_t type;          /* SRAY_LIGHT_* */
    float   pos[3];        /* get_position(frame time) */
    float   color[4];
    float   radius;        /* sphere light */
    float   dim[3];        /* box light */
    float   xform[12];     /* box light: get_xform_matrix() at time 0 */
    /* point / sphere light with a position track: the position at every
     * msec of [frame_t0, frame_t1] (3 floats per msec), or NULL */
    const float *time_pos;
} sray_light_desc;

typedef struct sray_camera_desc {
    float   xform[12];     /* camera -> world (Camera::get_matrix) */
    double  vfov;          /* radians (default QUARTER_PI, src/camera.cc:32) */
    float   max_dist;      /* RAY_MAGNITUDE */
    int32_t shutter;       /* msec */
    /* the aspect ratio is not a scene property: Camera() takes it from the
     * render size (src/camera.cc:31), so it comes from sray_render_params */
    /* an animated camera (Camera::get_matrix(ray.time), src/camera.cc:141):
     * xform at every msec of the scene's [frame_t0, frame_t1] (12 floats
     * per msec), or NULL.  Ignored in a request's own camera. */
    const float *time_xforms;
} sray_camera_desc;

typedef struct sray_scene_desc {
    uint32_t                  num_objects;
    const sray_object_desc   *objects;
    uint32_t                  num_meshes;
    const sray_mesh_desc     *meshes;
    uint32_t                  num_triangles;
    const float              *tri_pos;      /* 9 floats/tri, v0 v1 v2, mesh-local */
    const float              *tri_normal;   /* 9 floats/tri or NULL */
    const float              *tri_tangent;  /* 9 floats/tri or NULL */
    const float              *tri_texcoord; /* 6 floats/tri or NULL */
    uint32_t                  num_materials;
    const sray_material_desc *materials;
    uint32_t                  num_lights;
    const sray_light_desc    *lights;
    sray_camera_desc          camera;
    float                     env_ambient[4];
    float                     env_color[4];
    float                     env_ior;
    int32_t                   scene_oct_depth;  /* -soctdepth (default 5) */
    int32_t                   scene_oct_items;  /* -soctitems (default 5) */
    int32_t                   mesh_oct_depth;   /* -moctdepth (default 7) */
    int32_t                   mesh_oct_items;   /* -moctitems (default 14) */
    int32_t                   frame_t0, frame_t1; /* build_tree(t0, t1) interval */
    uint32_t                  num_textures;
    const sray_texture_desc  *textures;
} sray_scene_desc;

/* ------------------------------------------------------------------ */
/* rays and hits (the parity entry)                                    */
/* ------------------------------------------------------------------ */
typedef struct sray_ray {
    float   origin[3];
    float   dir[3];        /* segment end = origin + dir */
    int32_t time;          /* msec */
    int32_t pad_;
} sray_ray;

typedef struct sray_hit {
    float    t;            /* parametric distance, FLT_MAX on miss */
    int32_t  obj;          /* object index, -1 on miss */
    int32_t  prim;         /* triangle index within the mesh, -1 otherwise */
    uint32_t rank;         /* position of the winning occurrence in the scene
                              octree's depth-first item order, among the
                              leaves the walk visited (leaves beyond the hit
                              are skipped); when distinct items tie at t
                              the walk visits every leaf the segment meets,
                              and it is the reference's first occurrence */
} sray_hit;

#define SRAY_TRACE_CLOSEST   0u    /* Scene::cast_ray with a SurfPoint */
#define SRAY_TRACE_ANY       1u    /* Scene::cast_ray(ray) == NULL test (shadow) */
#define SRAY_TRACE_COUNT     2u    /* also fill visit counters (slower) */
#define SRAY_TRACE_WAVE      4u    /* walk with the instance the render path
                                      uses: a wave tests the leaves of its
                                      few lanes still in mesh leaves together
                                      (same results; for tests) */
#define SRAY_TRACE_SPARSE    8u    /* with SRAY_TRACE_WAVE: the instance of the
                                      sparse bounce levels (two triangles in
                                      flight per lane, world ray in
                                      registers) */

typedef struct sray_trace_counts {
    uint64_t nodes;        /* octree node records fetched (children are
                              tested from their parent's box; only those
                              the segment meets are fetched) */
    uint64_t spheres;
    uint64_t cylinders;
    uint64_t triangles;
    uint64_t item_refs;    /* leaf items tested (scene objects, triangles) */
    uint64_t mesh_enters;  /* mesh (second level) traversals started */
    uint64_t item_boxes;   /* mesh leaf item boxes tested (a triangle whose
                              box the segment misses is not tested) */
} sray_trace_counts;

/* ------------------------------------------------------------------ */
/* rendering                                                           */
/* ------------------------------------------------------------------ */
#define SRAY_RENDER_REF_Q1   0x1u  /* reproduce src/render.cc:240 as written */
#define SRAY_RENDER_COUNT    0x2u  /* fill sray_stats visit counters (slower) */

typedef struct sray_render_params {
    int32_t  width, height;         /* full image size (camera aspect) */
    int32_t  x0, y0, x1, y1;        /* pixel rectangle to render (tile); x1/y1 exclusive */
    int32_t  min_samples, max_samples;  /* -r min[-max]; max_samples below 1 renders 1 sample (the
                                       reference would divide by zero samples) */
    float    max_var;               /* -d (default 0.005) */
    float    min_energy;            /* -e (default 1e-4) */
    int32_t  iter;                  /* -i (default 7) */
    int32_t  shadow_samples;        /* -shadowrays (default 1); 0 or less: no area-light samples */
    int32_t  diffuse_samples;       /* -diffuserays (default 1) */
    int32_t  frame_time;            /* msec, render(msec) */
    int32_t  mblur;                 /* -m */
    int32_t  caust_photons;         /* -c : enables caustics gather */
    int32_t  gi_photons;            /* -g : enables GI diffuse rays */
    float    gather_dist;           /* -gatherdisc, relative to the scene octree diagonal */
    uint64_t seed;                  /* counter-based RNG seed (sample jitter, area lights, GI) */
    uint32_t flags;                 /* SRAY_RENDER_* */
    /* photons per gather estimate, the max_photons of
     * PhotonMap::irradiance_est / radiance_est (src/pmap.h:59-60, which
     * the reference leaves unused): 0 (or below) every photon within the
     * radius, as the reference; k > 0 the k nearest within the radius,
     * ties at equal distance by photon index.  The estimate keeps the
     * reference's 1 / (pi r^2) density. */
    int32_t  gather_k;
} sray_render_params;

typedef struct sray_stats {
    uint64_t rays_primary;
    uint64_t rays_shadow;
    uint64_t rays_reflect;
    uint64_t rays_refract;
    uint64_t rays_diffuse;
    uint64_t samples;               /* primary samples over all pixels */
    double   render_ms;             /* device time of the render phase */
    double   accel_ms;              /* scene octree build / upload time */
    sray_trace_counts closest;      /* SRAY_RENDER_COUNT: closest-hit traversal visits */
    sray_trace_counts shadow;       /* SRAY_RENDER_COUNT: shadow (any-hit) traversal visits */
} sray_stats;

/* ------------------------------------------------------------------ */
/* photon maps                                                         */
/* ------------------------------------------------------------------ */
typedef struct sray_photon {
    float pos[3];
    float dir[3];
    float norm[3];
    float col[3];
} sray_photon;

#define SRAY_PMAP_CAUSTIC    0
#define SRAY_PMAP_GLOBAL     1
#define SRAY_GATHER_IRRADIANCE 0   /* PhotonMap::irradiance_est, src/pmap.cc:109-139 */
#define SRAY_GATHER_RADIANCE   1   /* PhotonMap::radiance_est,  src/pmap.cc:143-178 */
/* or-ed into the gather mode: gather over a uniform grid of radius-sized
 * cells (built per radius on first use, the structure the renderer's
 * shading gathers use) instead of the kd-tree; same found set, and with
 * k > 0 the same k nearest (a radix select of dist^2, then photon index) */
#define SRAY_GATHER_GRID       0x10

typedef struct sray_gather_query {
    float pos[3];
    float norm[3];
    float dir[3];          /* radiance mode only */
    float pad_;
} sray_gather_query;

/* photon shooting (shoot_photons, src/render.cc:67-122) */
typedef struct sray_photon_params {
    int32_t  caust_photons;   /* -c: caustic photons shot */
    int32_t  gi_photons;      /* -g: global photons shot */
    float    photon_energy;   /* -photonenergy (default 300) */
    int32_t  iter;            /* -i: bounce limit (ray.iter) */
    int32_t  t0, t1;          /* frame interval (msec) */
    uint64_t seed;            /* counter-based RNG seed */
} sray_photon_params;

typedef struct sray_photon_stats {
    uint64_t shot[2];         /* caustic, global photons shot */
    uint64_t stored[2];       /* caustic, global photons stored */
    uint64_t segments;        /* photon ray segments traced */
    double   trace_ms;        /* device time of photon tracing */
    double   build_ms;        /* device time of the kd-tree builds */
} sray_photon_stats;

/* The 64x32 projection-map cell of a direction, without transcendental
 * functions: the polar index counts cos(k*pi/32) boundaries below y/|d|,
 * the azimuth index counts cos(2*pi*k/64) boundaries in the half plane of
 * z (ProjMap::cell_from_dir, src/projmap.cc:146-163, uses atan2/acos; the
 * result differs only exactly on a cell border). */
#define SRAY_PROJ_U 64
#define SRAY_PROJ_V 32

/* ------------------------------------------------------------------ */
/* batching / serving requests                                         */
/* ------------------------------------------------------------------ */
typedef struct sray_request {
    uint32_t           scene_id;
    sray_render_params params;      /* tile rectangle inside params */
    sray_camera_desc   camera;      /* per-request camera pose */
    int32_t            use_camera;  /* 0 = the scene's own camera */
    int32_t            flags;       /* SRAY_REQUEST_* */
} sray_request;

/* rgba of submit is device memory of the context's device (the image is
 * written there by the device; latency then ends when it is written) */
#define SRAY_REQUEST_DEVICE_OUT 0x1

typedef struct sray_result {
    float   *rgba;                  /* caller-owned, (x1-x0)*(y1-y0)*4 floats */
    uint64_t rays;                  /* ray queries made for this request */
    double   latency_ms;            /* submit -> result on host */
    double   queue_ms;              /* submit -> admitted */
} sray_result;

/* ------------------------------------------------------------------ */
/* opaque context                                                      */
/* ------------------------------------------------------------------ */
typedef struct sray_gpu_ctx sray_gpu_ctx;

/* Replaces rend_init(Image*) + ThreadPool::start (src/render.cc:22-32,
 * src/tpool.cc:48-74): one HIP stream instead of opt.threads workers. */
int  sray_gpu_create(int device, uint32_t flags, sray_gpu_ctx **out);
void sray_gpu_destroy(sray_gpu_ctx *ctx);
const char *sray_gpu_last_error(sray_gpu_ctx *ctx);

/* Replaces Scene::build_tree (src/scene.cc:353-390) + Mesh::build_tree
 * (src/mesh.cc:210-228) + the per-object data the Object::intersect
 * overrides read: builds both octree levels on the host, flattens them and
 * uploads everything to HBM. */
int  sray_gpu_scene_upload(sray_gpu_ctx *ctx, const sray_scene_desc *s, uint32_t *scene_id);
int  sray_gpu_scene_release(sray_gpu_ctx *ctx, uint32_t scene_id);
/* bytes of device memory held by a scene (admission accounting);
 * scene_id 0: everything the context holds (scenes, photon maps, queues) */
int  sray_gpu_scene_bytes(sray_gpu_ctx *ctx, uint32_t scene_id, uint64_t *bytes);

/* Replaces Scene::cast_ray (src/scene.cc:764-794) over the octree path
 * (Octree<T>::intersect, src/octree.inl:194-241).  Host buffers. */
int  sray_gpu_trace(sray_gpu_ctx *ctx, uint32_t scene_id, const sray_ray *rays, uint32_t n,
                    sray_hit *hits, uint32_t flags, sray_trace_counts *counts);
/* Same, on device-resident buffers, enqueued on the context stream. */
int  sray_gpu_trace_device(sray_gpu_ctx *ctx, uint32_t scene_id, const sray_ray *d_rays,
                           uint32_t n, sray_hit *d_hits, uint32_t flags);

/* Replaces render_frame/start_frame/render_block + Scene::trace_ray +
 * shade_phong (src/render.cc:124-244, src/scene.cc:551-569,
 * src/shader.cc:83-283): wavefront render of params' pixel rectangle into
 * caller-owned RGBA float (x1-x0)*(y1-y0)*4, row-major. */
int  sray_gpu_render(sray_gpu_ctx *ctx, uint32_t scene_id, const sray_render_params *p,
                     float *rgba, sray_stats *st);
/* Same, into a device buffer, timed with HIP events on the context stream. */
int  sray_gpu_render_device(sray_gpu_ctx *ctx, uint32_t scene_id, const sray_render_params *p,
                            float *d_rgba, sray_stats *st);

/* Replaces PhotonMap::add_photon/scale_photon_power + the kdtree build
 * (src/pmap.cc:73-100): uploads an identical photon array, builds a
 * balanced kd-tree on the device. */
int  sray_gpu_pmap_upload(sray_gpu_ctx *ctx, uint32_t scene_id, const sray_photon *p,
                          uint32_t n, int kind, uint32_t *pmap_id);
int  sray_gpu_pmap_release(sray_gpu_ctx *ctx, uint32_t pmap_id);
/* Replaces PhotonMap::irradiance_est / radiance_est (src/pmap.cc:107-178).
 * k == 0: fixed-radius range gather (reference semantics);
 * k > 0 : the k nearest within radius, ties broken by photon index. */
int  sray_gpu_pmap_gather(sray_gpu_ctx *ctx, uint32_t pmap_id, const sray_gather_query *q,
                          uint32_t n, float radius, int k, int mode, float *rgb);

/* No counterpart in the reference.  For callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): the fixed-radius gather of
 * sray_gpu_pmap_gather (k == 0, kd-tree; mode SRAY_GATHER_IRRADIANCE or
 * SRAY_GATHER_RADIANCE, no SRAY_GATHER_GRID) on queries d_q and results
 * d_rgb (n * 3 floats) already in device memory.  It allocates nothing:
 * its per-query counters live in a fixed device array of the library,
 * one per GPU, sized for SRAY_NOALLOC_MAX_QUERIES queries; larger n is
 * rejected with SRAY_ERR_LIMIT.  Asynchronous on the context's stream
 * (sray_gpu_synchronize waits).  Calls on one GPU share that array, so
 * calls that overlap on one GPU (other contexts, other threads) race. */
#define SRAY_NOALLOC_MAX_QUERIES (1u << 20)
int  sray_gpu_pmap_gather_device_noalloc(sray_gpu_ctx *ctx, uint32_t pmap_id,
                                         const sray_gather_query *d_q, uint32_t n, float radius,
                                         int mode, float *d_rgb);

/* Replaces shoot_photons + Scene::build_caustics_map/build_global_map +
 * Light::gen_photon/build_projmap (src/render.cc:67-122, src/scene.cc:445-760,
 * src/light.cc:111-205, src/projmap.cc): projection maps on the host,
 * photon paths traced on the device, both maps built as device kd-trees and
 * bound to the scene.  *caustic / *global receive the photon map ids (0 if
 * that map was not requested). */
int  sray_gpu_shoot_photons(sray_gpu_ctx *ctx, uint32_t scene_id, const sray_photon_params *pp,
                            uint32_t *caustic, uint32_t *global, sray_photon_stats *st);
/* photon array of a map in storage order (the order add_photon saw them) */
int  sray_gpu_pmap_download(sray_gpu_ctx *ctx, uint32_t pmap_id, sray_photon *out, uint32_t cap,
                            uint32_t *n);

/* Tuning of the device path (no reference counterpart: the reference's
 * knobs are its thread count and tile size, src/opt.c:338-342).  The values
 * change speed only; every image stays the same, except that the grid cell
 * scale changes the order in which a photon gather's sums are added. */
typedef struct sray_tuning {
    int32_t trace_block;      /* 64, 128 or 256 threads per traversal block */
    int32_t tile_w;           /* 8, 16, 32 or 64: primary rays in tiles of tile_w x (64 / tile_w) pixels */
    int32_t gather_lanes;     /* 8, 16 or 32 lanes sharing one photon-gather query */
    int32_t chunk_log2;       /* 16..21: rays per bounce-level sub-chunk = 2^chunk_log2 */
    float   gather_cell;      /* photon-grid cell size in gather radii: 0.25, 0.5, 1 or 2
                                 (renders with gather_k > 0 use half of it) */
    int32_t graphs;           /* 1: each primary chunk's bounce levels are replayed as a captured
                                 HIP graph when the pass runs on one lane; 0: launched kernel by
                                 kernel (same kernels, same order) */
    int32_t persistent;       /* 1: traversal kernels run as many waves as stay resident, each
                                 looping over runs of rays (16, 32 or 64: the fewest that cover
                                 the level in one run per wave); 2: so below the primary level
                                 only; 0: one run of 64 rays per wave */
    int32_t lanes;            /* 1, 2 or 4 HIP streams a render pass spreads its primary chunks over
                                 (each with its own bounce-level queues), so that one chunk's deep,
                                 sparse levels overlap another chunk's work */
    int32_t sparse_from;      /* 0..64: first bounce level whose traversal kernels keep two triangles
                                 in flight per lane and the world ray in registers (more registers,
                                 fewer waves; for levels with too few rays to hide load latency);
                                 64: none */
    int32_t chain;            /* 1: when no ray of the requests in flight can spawn more than one
                                 child and none uses photon maps (and at most 4 shadow rays per
                                 hit), one lane follows each sample's rays to the end in a single
                                 kernel; 0: always the level queues (same image) */
    int32_t sample_runs;      /* 1: a request's first pass takes the samples every pixel takes
                                 (max(min_samples, 2) up to max_samples, as a power of two <= 16:
                                 all of them when min_samples >= max_samples), each pixel's
                                 samples in consecutive queue entries and summed in sample
                                 order (same image, fewer passes); 0: one sample per pass */
} sray_tuning;
int  sray_gpu_set_tuning(sray_gpu_ctx *ctx, const sray_tuning *t);
int  sray_gpu_get_tuning(sray_gpu_ctx *ctx, sray_tuning *t);
/* Autotuner: renders params of the scene `rounds` times per candidate,
 * varying one knob at a time from the current tuning (coordinate descent
 * over the values above); a change is kept when its median render time
 * beats the current tuning's by 1.5%, twice (the second time with both
 * timed again, one after the other).  Keeps the tuning in the context and
 * returns it with its median render time. */
int  sray_gpu_autotune(sray_gpu_ctx *ctx, uint32_t scene_id, const sray_render_params *p, int rounds,
                       sray_tuning *best, double *best_ms);
/* The octree knobs of a scene description (leaf size and depth of the mesh
 * and scene octrees, src/opt.c:383-387): renders params with each of
 * candidates (mesh items, mesh depth) pairs and returns the fastest pair.
 * Changing them changes only which of two exactly equal hit distances
 * wins (the depth-first rank), never a hit's distance. */
int  sray_gpu_autotune_octree(sray_gpu_ctx *ctx, const sray_scene_desc *s, const sray_render_params *p, int rounds,
                              const int32_t *items, const int32_t *depths, int ncand, int *best, double *best_ms);

/* Batching / serving (no reference counterpart: the reference renders one
 * frame of one scene at a time, src/sray.cc:56-65).  submit queues a tile
 * request (its pixel rectangle of params, through the scene's camera or
 * r->camera) and returns at once; a worker thread admits queued requests
 * first come first served while the context's device memory (scenes,
 * photon maps, queues, requests in flight) stays under the admission cap
 * (default 200 GB), and every pass of the renderer takes one sample of
 * every active pixel of every admitted request, so requests that arrive
 * while others are on the GPU join at the next pass.  rgba (host memory,
 * or device memory with SRAY_REQUEST_DEVICE_OUT) must stay valid until
 * poll reports the ticket done.  poll: SRAY_ERR_PENDING if not done
 * and wait == 0; otherwise the request's status (SRAY_ERR_NOMEM if it can
 * never fit the cap), after which the ticket is forgotten.  A request's
 * image is the same alone and in any mix. */
int  sray_gpu_submit(sray_gpu_ctx *ctx, const sray_request *r, float *rgba, uint64_t *ticket);
/* n requests queued together (one ticket each, in order): all are checked
 * before any is queued, so either all are queued or none (no reference
 * counterpart, as submit) */
int  sray_gpu_submit_many(sray_gpu_ctx *ctx, const sray_request *r, uint32_t n, float *const *rgba,
                          uint64_t *tickets);
int  sray_gpu_poll(sray_gpu_ctx *ctx, uint64_t ticket, int wait, sray_result *out);
int  sray_gpu_set_admission_cap(sray_gpu_ctx *ctx, uint64_t bytes);

/* ------------------------------------------------------------------ */
/* host-only entry points (no device needed)                           */
/* ------------------------------------------------------------------ */
typedef struct sray_host_scene sray_host_scene;

/* Replaces Scene::load(const char*) / load(FILE*) (src/scene.cc:75-183):
 * parses reference-format scene XML (from path, or xml_text when given)
 * into a flat description owned by the returned handle.  A material
 * attribute's "map" file is looked for as find_file does
 * (src/datapath.c:56-77): in the directory of path ("." without one),
 * then in each entry of SRAY_DATA_PATH; an absolute name as it is.  A map
 * that cannot be found or read is reported on stderr and the attribute
 * keeps its plain value (src/material.cc:64-71). */
int  sray_scene_load_xml(const char *path, const char *xml_text, sray_host_scene **out);
const sray_scene_desc *sray_scene_desc_of(const sray_host_scene *hs);
/* Replaces the per-frame part of render(msec) (src/render.cc:34-41) and the
 * XFormNode track evaluation (src/anim.inl:188-275, src/anim.cc:391-427)
 * for a scene loaded from XML: the frame interval msec - shutter/2 ..
 * + shutter, every transform at msec, and per-msec rows of the animated
 * objects, lights and camera over the interval (upload the scene again to
 * render the new frame).  The load itself evaluates msec 0. */
int  sray_scene_set_frame(sray_host_scene *hs, int32_t msec);
void sray_scene_free(sray_host_scene *hs);
const char *sray_host_last_error(void);

/* Flattened octree of one level, as the device sees it (for inspection and
 * host-side tests).  level = -1 for the scene octree, else a mesh index. */
typedef struct sray_octree_info {
    uint32_t num_nodes;
    uint32_t num_leaves;
    uint32_t num_item_refs;
    uint32_t height;
    float    root_min[3], root_max[3];
} sray_octree_info;
int  sray_octree_build_host(const sray_scene_desc *s, int level, sray_octree_info *info,
                            float *node_lo, float *node_hi, uint32_t *items,
                            uint32_t cap_nodes, uint32_t cap_items);

/* Replaces Light::build_projmap (src/light.cc:151-205, 392-432): the
 * projection map of one light over the frame interval [t0, t1] as
 * SRAY_PROJ_U x SRAY_PROJ_V cell bytes (v-major, 1 = full) and its
 * coverage; kind 0 = all objects (global photons), 1 = reflecting or
 * refracting objects (caustic photons).  Host only, the same maps
 * sray_gpu_shoot_photons builds. */
int  sray_projmap_build(const sray_scene_desc *s, int light, int kind, int t0, int t1, uint64_t seed,
                        uint8_t *cells, double *coverage);
/* Replaces ProjMap::cell_from_dir (src/projmap.cc:146-163): the cell of a
 * direction (see SRAY_PROJ_U above). */
int  sray_projmap_cell(float x, float y, float z, int *uc, int *vc);

/* Seeded synthetic workloads (no reference counterpart; the reference ships
 * no scenes): config 0 = 64 spheres, 1 = fBm heightfield of 1M triangles +
 * 40 glass/mirror spheres, 2 = photon room, 3 = batch scene with `variant`
 * triangles.  For configs 0-2, variant > 1 shrinks the scene for tests. */
int  sray_scene_generate(int config, uint64_t seed, int variant, sray_host_scene **out);
/* reference-format XML of a host scene (%.9g floats); *len receives the
 * size including the terminating NUL, buf may be NULL to query it */
int  sray_scene_write_xml(const sray_host_scene *hs, char *buf, size_t cap, size_t *len);

/* Photon-map wire format.  Replaces PhotonMap::dump (src/pmap.cc:180-205):
 * writes "energy %f\n", then one line per photon,
 * "<%f %f %f> <%f %f %f> <%f %f %f> <%f %f %f>\n" (position, direction,
 * normal, colour rgb), as the reference's printf. */
int  sray_photon_dump(const char *path, const sray_photon *p, uint32_t n, float energy);
/* Replaces PhotonMap::restore (src/pmap.cc:207-251): reads a dump (the
 * reference's scanf formats; a short line is an error) and scales every
 * colour by energy / (the dump's energy).  *out (NULL when *n is 0) is
 * freed with sray_photon_free; the photons go to the device with
 * sray_gpu_pmap_upload. */
int  sray_photon_restore(const char *path, float energy, sray_photon **out, uint32_t *n);
void sray_photon_free(sray_photon *p);
/* Replaces Image::save (src/img.cc:80-83, libimago's float formats, which
 * are not available here): the RGBA float framebuffer (row-major, top row
 * first) as a PFM file, "PF\n<w> <h>\n-1.0\n" then RGB float32
 * little-endian, bottom row first. */
int  sray_image_write_pfm(const char *path, const float *rgba, int width, int height);
/* Replaces Image::save (src/img.cc:80-83) for the reference's default
 * output, output.png / frame%04d.png (src/sray.cc:150-155): the RGBA float
 * framebuffer (row-major, top row first) as an 8-bit RGBA PNG, each channel
 * value * 255 truncated toward zero and clamped to [0, 255], the float to
 * byte conversion libimago applies before its PNG writer. */
int  sray_image_write_png(const char *path, const float *rgba, int width, int height);
/* The backend status protocol (emit_status, src/render.cc:279-289): one
 * record is "%c %6d %6d %6d %6d\n" at the start of SRAY_STATUS_RECORD
 * bytes, the rest zero.  'f' frame (x blocks, y blocks, block size,
 * threads), 's' block start (x, y, width, height), 'e' block end (the
 * same), 'd' done (milliseconds, ...). */
#define SRAY_STATUS_RECORD 32
int  sray_status_record(char st, int a1, int a2, int a3, int a4, char *out);
/* Replaces Image::load (src/img.cc:64-78, which reads through libimago
 * as RGBA float): a PNG (every colour type and depth, tRNS transparency;
 * not interlaced), PPM/PGM (P2, P3, P5, P6; maxval up to 65535) or PFM
 * ("PF"/"Pf", either byte order) file, told apart by its first bytes, as
 * RGBA floats, row-major, top row first.  An integer channel reads
 * value / maxval (8-bit: / 255); a file without alpha reads alpha 1.
 * *rgba is freed with sray_image_free. */
int  sray_image_read(const char *path, float **rgba, int *width, int *height);
void sray_image_free(float *rgba);
/* text of the last failure of the I/O calls above (per thread) */
const char *sray_io_last_error(void);
/* Status stream of a context (fd < 0: off, the default).  A render writes
 * 'f' (1, 1, width, lanes), 's' (x0, y0, width, height) when it starts,
 * 'e' (the same) and 'd' (device milliseconds) when it ends.  A submitted
 * request is a block: 's' when it is admitted to the GPU, 'e' when it is
 * done, then 'd' (latency ms, queue ms, 0, 0).  Records are written whole,
 * under a lock, from the thread that renders. */
int  sray_gpu_set_status_fd(sray_gpu_ctx *ctx, int fd);

/* Device memory and synchronisation on the context's device, so callers
 * (tests, benchmarks, the serve loop) need no second GPU runtime in the
 * process.  kind: 0 host->device, 1 device->host, 2 device->device.
 * No reference counterpart (the reference has no device). */
int  sray_gpu_malloc(sray_gpu_ctx *ctx, size_t bytes, void **dptr);
int  sray_gpu_free(sray_gpu_ctx *ctx, void *dptr);
int  sray_gpu_memcpy(sray_gpu_ctx *ctx, void *dst, const void *src, size_t bytes, int kind);
int  sray_gpu_synchronize(sray_gpu_ctx *ctx);

/* Per-kernel device time, measured with HIP events around each launch on
 * the stream it runs on while profiling is on (no reference counterpart;
 * the reference only prints total wall time, src/sray.cc:67-76).  While
 * profiling, renders launch kernel by kernel (no graphs), and a bounce
 * level's kernels run in order on one stream per lane. */
typedef struct sray_kernel_stat {
    char     name[32];
    uint64_t launches;
    double   total_ms;
} sray_kernel_stat;
int  sray_gpu_set_profiling(sray_gpu_ctx *ctx, int on);
/* copies up to cap entries, *n = entries available; reset != 0 clears them */
int  sray_gpu_kernel_stats(sray_gpu_ctx *ctx, sray_kernel_stat *out, int cap, int *n, int reset);

/* sizeof() of the public structs, in declaration order (sray_object_desc,
 * sray_mesh_desc, sray_material_desc, sray_light_desc, sray_camera_desc,
 * sray_scene_desc, sray_ray, sray_hit, sray_trace_counts,
 * sray_render_params, sray_stats, sray_photon, sray_gather_query,
 * sray_request, sray_result, sray_octree_info, sray_kernel_stat,
 * sray_photon_params, sray_photon_stats, sray_tuning) — lets
 * bindings check their
 * layout.  Returns the number of entries written (at most n). */
int  sray_abi_sizes(uint32_t *out, int n);

/* Primary rays of one sample pass (Camera::get_primary_ray,
 * src/camera.cc:119-142) generated by the device raygen kernel into a
 * device buffer of (x1-x0)*(y1-y0) rays, row-major. */
int  sray_gpu_primary_rays_device(sray_gpu_ctx *ctx, uint32_t scene_id,
                                  const sray_render_params *p, int sample, sray_ray *d_rays);

#ifdef __cplusplus
}
#endif
#endif /* SRAY_GPU_H_ */
