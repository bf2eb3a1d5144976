/*
 * device.h — device-side data layout and FP32 math of the gfx950 kernels.
 *
 * All arithmetic is IEEE binary32 with contraction off (the file is built
 * with -ffp-contract=off and the pragma below), in the operation order the
 * CPU restatement's float path uses, so hit IDs and spawned rays agree
 * bit for bit:
 *   dot(a,b) = (a.x*b.x + a.y*b.y) + a.z*b.z
 *   point  M*p = ((m0*x + m1*y) + m2*z) + m3,  vector: (m0*x + m1*y) + m2*z
 *   normal (M^-1)^T n = (i0*x + i4*y) + i8*z, ...
 *   normalize = per-component true division by sqrt(dot(v,v))
 */
#ifndef SRAY_DEVICE_H_
#define SRAY_DEVICE_H_

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/sray_gpu.h"
#include "rng.h"

#pragma clang fp contract(off)

namespace srd {

constexpr uint32_t LEAF = 0x80000000u;
constexpr int STACK = 20;      /* per-lane traversal stack entries (LDS) */
constexpr int TRACE_BLOCK = 128;
constexpr int IOR_DEPTH = 4;   /* ior stack carried by queued rays */
constexpr float EPS = (float)SRAY_ERROR_MARGIN;
constexpr float XSMALL = (float)SRAY_XSMALL_NUMBER;

/* Item boxes of mesh leaves.  Every item position of a mesh leaf holds a
 * uint2 whose low six bytes are codes x0 y0 z0 x1 y1 z1 of its triangle's
 * bounding box, relative to the leaf's cell [lo, hi]: per axis, with
 * ext = hi - lo, base = lo - ext * IBOX_GROW and step = ext * IBOX_STEP
 * (float operations), code q in 0..255 is the plane base + q * step (the
 * codes span the cell grown by four cell sizes on each side).  Lower codes
 * are at or below the triangle's minimum, upper codes at or above its
 * maximum.  IBOX_OPEN in the high word marks a triangle whose box leaves
 * that range, or whose test could accept points off its box (non-finite or
 * badly shaped): it is always tested. */
constexpr float IBOX_GROW = 4.0f;
constexpr float IBOX_STEP = 9.0f / 255.0f;
constexpr uint32_t IBOX_OPEN = 0x10000u;

/* object record: traversal reads the first 80 bytes */
struct alignas(16) DevObj {
    float inv[12];   /* world -> object */
    float rad, start, end;
    int32_t type;
    int32_t mat;     /* material index (default material resolved on upload) */
    uint32_t root;   /* mesh: root node of its octree */
    uint32_t tri_base;
    int32_t moving;  /* 0 static, 1 linear translation (motion), 2 keyframed rows (tab) */
    float xf[12];    /* object -> world */
    float motion[3];
    int32_t mt0;
    int32_t mt1;
    int32_t mesh;
    int32_t has_attr; /* mesh: 1 normals, 2 texcoords, 4 tangents */
    int32_t pad;
    /* keyframed motion: xform + inverse (24 floats) per msec from tab_t0 */
    const float *tab;
    int32_t tab_t0, tab_n;
};

/* A scene-level item occurrence: the part of its object's record the
 * traversal reads, stored per occurrence so that one load by occurrence
 * rank brings it (a moving object's matrices still come from its DevObj) */
struct alignas(16) SceneItem {
    float inv[12];
    float rad, start, end;
    int32_t type;
    int32_t obj;     /* object index */
    uint32_t root;   /* mesh: root node of its octree */
    int32_t moving;  /* DevObj::moving */
    int32_t pad;
};

struct alignas(16) DevMat {
    float4 kd, ks;
    float shininess, reflect, refract, ior;
    uint32_t flags, pad0, pad1, pad2;
    /* textures of the attributes (SRAY_MAP_*): 1 + texture index, 0 = none */
    int32_t map[SRAY_NUM_MAPS];
    float nscale; /* the "normal" attribute's value */
};
constexpr uint32_t MAT_MAPPED = 0x100u; /* some attribute has a texture */

/* Texture (src/texture.cc): RGBA texels, filter, wrap, coordinate matrix */
struct alignas(16) DevTex {
    const float4 *texels;
    int32_t w, h, filter, wrap;
    float xf[12];
};

struct alignas(16) DevLight {
    int32_t type;
    float pos[3];
    float color[4];
    float radius;
    float dim[3];
    float xf[12];
    /* a point / sphere light's position (3 floats) per msec from tab_t0 */
    const float *tab;
    int32_t tab_t0, tab_n;
};

/* the row of a per-msec table at a time (the nearest row outside it) */
__device__ __forceinline__ const float *time_row(const float *tab, int t0, int n, int time, int width)
{
    int k = time - t0;
    k = k < 0 ? 0 : (k >= n ? n - 1 : k);
    return tab + (size_t)k * (size_t)width;
}

struct SceneDev {
    const float4 *node;      /* 2 per node: lo (w = first child or item), hi (w = child mask or LEAF | count) */
    const uint32_t *items;
    const SceneItem *sitem;  /* per scene item position (the scene octree's leaves) */
    const DevObj *objs;
    const float4 *tri;       /* 3 per triangle: v0, v1, v2; w = face normal (v1-v0) x (v2-v0) */
    const float4 *ltri;      /* the same, 3 per item position: mesh leaves' triangles in leaf order */
    const uint2 *lbox;       /* item box codes per item position (IBOX_STEP) */
    const float *tri_n;      /* 9 per triangle or null */
    const float *tri_g;      /* 9 per triangle or null */
    const float *tri_uv;     /* 6 per triangle or null */
    const DevMat *mats;
    const DevLight *lights;
    uint32_t scene_root;
    int32_t nlights;
    int32_t nobjs;
    int32_t fast_boxes;      /* every octree box finite, lo <= hi, |bounds| <= 2^120 */
    float4 env_ambient;
    float4 env_color;
    float cam[12];
    float fz;                /* 1/tan(vfov/2), from double on the host */
    float max_dist;
    int32_t shutter;
    float gather_abs;        /* gather radius (absolute) */
    /* animated camera: xform (12 floats) per msec from cam_t0 */
    const float *cam_tab;
    int32_t cam_t0, cam_n;
    const DevTex *texs;
    int32_t ntex;
};

/* ------------------------------------------------------------------ */
/* small vector helpers (explicit order, no contraction)                */
/* ------------------------------------------------------------------ */
struct V3 {
    float x, y, z;
};
__device__ __forceinline__ V3 v3(float x, float y, float z) { return V3{x, y, z}; }
__device__ __forceinline__ V3 sub(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 add(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 mul(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ V3 dvs(V3 a, float s) { return V3{a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ V3 neg(V3 a) { return V3{-a.x, -a.y, -a.z}; }
__device__ __forceinline__ float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V3 cross(V3 a, V3 b)
{
    return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ float length(V3 a) { return sqrtf(dot(a, a)); }
__device__ __forceinline__ V3 normalized(V3 a)
{
    float l = length(a);
    return V3{a.x / l, a.y / l, a.z / l};
}
__device__ __forceinline__ V3 reflect(V3 d, V3 n)
{
    float s = 2.0f * dot(d, n);
    return V3{d.x - n.x * s, d.y - n.y * s, d.z - n.z * s};
}
__device__ __forceinline__ V3 xf_point(const float *m, V3 p)
{
    return V3{m[0] * p.x + m[1] * p.y + m[2] * p.z + m[3], m[4] * p.x + m[5] * p.y + m[6] * p.z + m[7],
              m[8] * p.x + m[9] * p.y + m[10] * p.z + m[11]};
}
__device__ __forceinline__ V3 xf_vec(const float *m, V3 v)
{
    return V3{m[0] * v.x + m[1] * v.y + m[2] * v.z, m[4] * v.x + m[5] * v.y + m[6] * v.z,
              m[8] * v.x + m[9] * v.y + m[10] * v.z};
}
__device__ __forceinline__ V3 xf_normal(const float *i, V3 n)
{
    return V3{i[0] * n.x + i[4] * n.y + i[8] * n.z, i[1] * n.x + i[5] * n.y + i[9] * n.z,
              i[2] * n.x + i[6] * n.y + i[10] * n.z};
}

/* base[key] += v over the lanes of a wave with act set, one atomic per
 * distinct key of the wave instead of one per lane (the lanes of one
 * request are contiguous, so a wave holds one or two keys).  Every lane of
 * the wave must call it. */
template <typename T>
__device__ __forceinline__ void wave_keyed_add(T *base, uint32_t key, T v, bool act)
{
    const uint32_t lane = __lane_id();
    unsigned long long left = __ballot(act && v != (T)0);
    while (left) {
        int lead = __ffsll(left) - 1;
        uint32_t k0 = (uint32_t)__shfl((int)key, lead);
        bool mine = ((left >> lane) & 1ull) && key == k0;
        unsigned long long grp = __ballot(mine);
        T s = mine ? v : (T)0;
        for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
        if (lane == (uint32_t)lead) atomicAdd(&base[k0], s);
        left &= ~grp;
    }
}

} // namespace srd

#endif
