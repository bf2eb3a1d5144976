/*
 * oracle.h — CPU restatement of the S-ray 0.1 hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or as the timed CPU baseline).  The product path (sray_amd/, libsray_gpu.so)
 * never links, loads or calls it.
 *
 * Two precisions from one source:
 *   prec = 64 : double, the reference's own arithmetic (SurfPoint and vmath
 *               are double, src/object.h:31-37) = "reference semantics";
 *   prec = 32 : float, strict IEEE (built -ffp-contract=off, no fast-math),
 *               with the same operation order as the device kernels — the
 *               target for bit-exact hit IDs.
 *
 * Pins (tests/test_oracle_kat.py, tests/test_photons_cpu.py,
 * tests/test_host.py): closed forms, brute force and invariants for every
 * function here, except the following, which are parity unpinned:
 *   - the sample points of sphere and box lights (Light::get_point,
 *     src/light.cc:272-275, 346-354) beyond their use in the renders: their
 *     distribution comes from the build's own generator;
 *   - ERROR_MARGIN / XSMALL_NUMBER and the refraction formula, which live in
 *     the un-vendored libvmath: the values and the textbook form are fixed
 *     here (the glass and Snell closed forms pin the form, not libvmath).
 */
#ifndef SRAY_ORACLE_H_
#define SRAY_ORACLE_H_

#include "../include/sray_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_scene oracle_scene;

/* extra per-ray information the oracle records beside sray_hit */
typedef struct oracle_hit_ext {
    int32_t  tie;          /* another distinct item had exactly the winning t */
    int32_t  pad_;
    uint64_t nodes;        /* octree nodes visited (exhaustive traversal) */
    uint64_t prims;        /* primitive tests */
} oracle_hit_ext;

typedef struct oracle_tree_info {
    uint32_t num_nodes;    /* all nodes, empty leaves included */
    uint32_t num_leaves;   /* leaves holding at least one item */
    uint32_t num_item_refs;
    uint32_t height;
    double   root_min[3], root_max[3];
} oracle_tree_info;

/* flags for oracle_scene_create */
#define ORACLE_Q9_RETRAVERSE 0x1u  /* re-traverse a mesh in every scene leaf (src/octree.inl:211-222) */
#define ORACLE_BRUTE_FORCE   0x2u  /* Scene::cast_ray / Mesh::intersect without octrees */
#define ORACLE_BRUTE_GATHER  0x4u  /* photon gathers test every photon (no bucket grid) */

oracle_scene *oracle_scene_create(const sray_scene_desc *desc, int prec, uint32_t flags);
void oracle_scene_destroy(oracle_scene *s);

/* Scene::cast_ray (src/scene.cc:764-794); any != 0: only "is anything hit" */
int oracle_trace(oracle_scene *s, const sray_ray *rays, uint32_t n, sray_hit *hits,
                 oracle_hit_ext *ext, int any, int nthreads);

/* octree of one level (-1 scene, else mesh index): shape and DFS leaf lists.
 * leaf_items receives, for every non-empty leaf in depth-first child order
 * 0..7 (src/octree.inl:226-234), its items in list order; leaf_counts the
 * per-leaf counts. */
int oracle_tree(oracle_scene *s, int level, oracle_tree_info *info, uint32_t *leaf_items,
                uint32_t *leaf_counts, uint32_t cap_items, uint32_t cap_leaves);

/* Camera::get_primary_ray (src/camera.cc:119-142) for every pixel of the
 * params rectangle and one sample index. */
int oracle_primary_rays(oracle_scene *s, const sray_render_params *p, int sample, sray_ray *rays);

/* render_block over the params rectangle (src/render.cc:194-244) with a tile
 * pool of nthreads workers (src/tpool.cc), 48x48 tiles in Gaussian priority
 * order (src/render.cc:144-192).  rgba: (x1-x0)*(y1-y0)*4 floats. */
int oracle_render(oracle_scene *s, const sray_render_params *p, float *rgba, int nthreads,
                  sray_stats *st);

/* photon maps for rendering and gather: kind = SRAY_PMAP_* */
int oracle_set_photons(oracle_scene *s, int kind, const sray_photon *p, uint32_t n);
/* PhotonMap::irradiance_est / radiance_est by brute force over the array
 * (exact range semantics: dist^2 <= r^2 in double on float positions);
 * k > 0: k nearest within radius, ties by photon index. */
int oracle_gather(oracle_scene *s, int kind, const sray_gather_query *q, uint32_t n,
                  float radius, int k, int mode, float *rgb);

/* shoot_photons (src/render.cc:67-122): projection maps, emission and
 * tracing of caustic and global photons; stored photons in storage order.
 * caust/glob: capacity cap_c / cap_g photons; counts returned in n_c / n_g. */
int oracle_shoot_photons(oracle_scene *s, const sray_photon_params *pp, sray_photon *caust, uint32_t cap_c,
                         uint32_t *n_c, sray_photon *glob, uint32_t cap_g, uint32_t *n_g, uint64_t *segments);
/* projection maps of a light (kind 0 = all objects, 1 = specular objects):
 * 64x32 cells as bytes (v-major), and the coverage */
int oracle_projmap(oracle_scene *s, int light, int kind, int t0, int t1, uint64_t seed, uint8_t *cells,
                   double *coverage);
/* ProjMap cell of a direction (the transcendental-free definition) */
int oracle_cell_from_dir(float x, float y, float z, int *uc, int *vc);

/* surface point of each ray's closest hit: 8 floats per ray (pos, normal,
 * u, v; NaN on a miss) */
int oracle_surface(oracle_scene *s, const sray_ray *rays, uint32_t n, float *out);

/* scene octree root diagonal (Scene::build_tree, src/scene.cc:384-387) */
double oracle_scene_diag(oracle_scene *s);

/* n sample points of light `light` (Light::get_point, src/light.cc:106-109,
 * 272-275, 346-354), the draws of samples 0..n-1 of one pixel; 3 doubles each */
int oracle_light_points(oracle_scene *s, int light, uint64_t seed, uint32_t pixel, uint32_t n, double *out);

/* counter-based RNG used by both implementations (exposed for KAT tests) */
uint32_t oracle_rng_u32(uint64_t seed, uint32_t a, uint32_t b, uint32_t c, uint32_t d);
uint64_t oracle_splitmix64(uint64_t x);

#ifdef __cplusplus
}
#endif
#endif
