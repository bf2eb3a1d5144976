/*
 * traverse.h — two-level octree traversal on gfx950 (one ray per lane).
 *
 * Replaces Octree<T>::intersect_rec (src/octree.inl:200-241) at both levels:
 * the scene octree over objects and, inside a mesh item, that mesh's octree
 * over triangles (Mesh::intersect, src/mesh.cc:252-308).
 *
 * The reference walks every node whose box meets the segment, children in
 * fixed order 0..7, and keeps the first occurrence of the smallest t (strict
 * '<').  Here children are visited near-to-far (octant bits XOR the ray's
 * sign mask), a node is skipped when its entry distance exceeds the best
 * t found so far (by a small relative slack, below), and candidates are
 * ranked by the
 * pair (t, occurrence rank) — the rank being the item's position in the
 * depth-first item array.
 *
 * Ties.  Skipping far nodes is exact for the distance, and for the winner
 * whenever one item alone holds the smallest t.  When two distinct items
 * tie at it, the reference's winner is the one whose first occurrence
 * comes first in depth-first order, and that occurrence may sit in a leaf
 * the near-to-far walk skipped: a leaf beyond the hit that comes earlier
 * in depth-first order and holds an item whose box reaches into it.  So
 * the walk notes every exact tie between distinct items at the running
 * best t (at either level, a mesh's internal ties included), and a ray that
 * ends with one is walked again without skipping any node: then every
 * occurrence the reference tests is tested, and the (t, rank) order picks
 * the reference's winner.  Ties are rare, so the second walk costs little.
 * To see a tie, the first walk must reach both tied items.  A tied item
 * can sit in a node whose box the ray enters exactly where it hits the
 * item (a flat mesh's box, a triangle whose edge lies on a split plane);
 * the box's entry distance and the item's hit distance are then equal in
 * exact arithmetic but rounded differently.  So a node is skipped only
 * when its entry exceeds the best t by a relative PRUNE_SLACK, far more
 * than the few roundings between the two.
 *
 * Per-lane stack in LDS, 32-bit entries:
 *   ITER   first child block << 8 | pending children whose boxes the
 *          segment meets (order space)                              (bit 31 = 0)
 *   RESUME two words: the next scene item, then E_RES | end of the
 *          leaf's items (a scene leaf interrupted by a mesh)
 *   MARKER E_MARK | scene item rank of the mesh being traversed
 * Children sit in eight-slot blocks (child with octant bits k at first + k),
 * so an ITER entry needs no child mask; four bytes per entry keep the LDS
 * footprint at 80 bytes per lane.
 * A mesh met by the same ray in several scene leaves (Q9) is traversed once:
 * its result is memoised for the ray (for an any-hit ray, its miss).
 * A mesh leaf's triangles are read from a copy in item order (no index
 * load before the vertex loads); the triangle id of a hit is looked up
 * from its item position once the mesh is done.
 */
#ifndef SRAY_TRAVERSE_H_
#define SRAY_TRAVERSE_H_

#include <cfloat>

#include "device.h"

namespace srd {

constexpr uint32_t E_RES = 0x80000000u, E_MARK = 0xC0000000u;

/* a node is skipped when its entry distance exceeds best t * PRUNE_SLACK */
constexpr float PRUNE_SLACK = 1.0f + 0x1p-14f;

struct TRay {
    float ox, oy, oz, dx, dy, dz, ix, iy, iz;
    uint32_t dmask;
    /* the box tests may take their min/max form (box_tests_exact) */
    bool fast;
};

__device__ __forceinline__ TRay make_tray(V3 o, V3 d)
{
    TRay r;
    r.ox = o.x;
    r.oy = o.y;
    r.oz = o.z;
    r.dx = d.x;
    r.dy = d.y;
    r.dz = d.z;
    r.ix = 1.0f / d.x;
    r.iy = 1.0f / d.y;
    r.iz = 1.0f / d.z;
    r.dmask = (d.x < 0.0f ? 1u : 0u) | (d.y < 0.0f ? 2u : 0u) | (d.z < 0.0f ? 4u : 0u);
    r.fast = false;
    return r;
}

/* When the box tests may use their min/max form.  slab() below is the
 * reference's sequence of comparisons and updates.  With no NaN among its
 * values and near <= far on every axis, that sequence accepts exactly when
 * max(entry) <= min(exit), entry < 1 and exit > 0, and its final entry and
 * exit are that max and min: the pairwise checks it makes, together with
 * near <= far per axis, are all nine "entry <= exit" pairs.  Near <= far
 * and no NaN hold when the origin is finite, every 1/d component is finite
 * and non-zero (d finite, non-zero, not subnormal enough to overflow; its
 * sign then matches d's, which picks near and far), and every box of the
 * scene is finite with lo <= hi and |bounds| <= 2^120 (so that no midpoint
 * overflows: lo <= (lo + hi) * 0.5 <= hi).  The scene's condition is
 * checked once at upload (SceneDev::fast_boxes). */
__device__ __forceinline__ bool box_tests_exact(const TRay &r)
{
    auto fin = [](float v) { return __builtin_isfinite(v); };
    auto fnz = [](float v) { return __builtin_isfinite(v) && v != 0.0f; };
    return fin(r.ox) && fin(r.oy) && fin(r.oz) && fnz(r.ix) && fnz(r.iy) && fnz(r.iz);
}

__device__ __forceinline__ TRay make_tray_s(V3 o, V3 d, bool fast_boxes)
{
    TRay r = make_tray(o, d);
    r.fast = fast_boxes && box_tests_exact(r);
    return r;
}

/* AABox::intersect, src/aabb.cc:81-125; tmin_out = entry distance */
__device__ __forceinline__ bool slab(float4 lo, float4 hi, const TRay &r, float &tmin_out)
{
    bool xs = r.dx < 0.0f, ys = r.dy < 0.0f, zs = r.dz < 0.0f;
    float tmin = ((xs ? hi.x : lo.x) - r.ox) * r.ix;
    float tmax = ((xs ? lo.x : hi.x) - r.ox) * r.ix;
    float tymin = ((ys ? hi.y : lo.y) - r.oy) * r.iy;
    float tymax = ((ys ? lo.y : hi.y) - r.oy) * r.iy;
    if (r.fast) {
        const float tzmin = ((zs ? hi.z : lo.z) - r.oz) * r.iz;
        const float tzmax = ((zs ? lo.z : hi.z) - r.oz) * r.iz;
        const float a = fmaxf(fmaxf(tmin, tymin), tzmin), b = fminf(fminf(tmax, tymax), tzmax);
        tmin_out = a;
        return a <= b && a < 1.0f && b > 0.0f;
    }
    if ((tmin > tymax) || (tymin > tmax)) return false;
    if (tymin > tmin) tmin = tymin;
    if (tymax < tmax) tmax = tymax;
    float tzmin = ((zs ? hi.z : lo.z) - r.oz) * r.iz;
    float tzmax = ((zs ? lo.z : hi.z) - r.oz) * r.iz;
    if ((tmin > tzmax) || (tzmin > tmax)) return false;
    if (tzmin > tmin) tmin = tzmin;
    if (tzmax < tmax) tmax = tzmax;
    tmin_out = tmin;
    return (tmin < 1.0f) && (tmax > 0.0f);
}

/* The slab tests of a node's eight children at once.  Children are the
 * octants around the box centre (lo + hi) * 0.5, cut exactly so by the
 * host build, so the values slab() computes on a child's own box are
 * products of the node's bounds and its centre: per axis the near half
 * spans [N, C] and the far half [C, F] in ray parameter (N, F the node's
 * entry / exit products, C the centre's).  Each child then goes through
 * slab()'s own comparison sequence (same floats, same order, so the same
 * result, NaN cases included).  exist and the result are in order space
 * (bit k = child k ^ dmask, near to far); children whose entry distance
 * exceeds bound are dropped. */
__device__ __forceinline__ uint32_t expand_children(float4 lo, float4 hi, const TRay &r, uint32_t exist, float bound)
{
    const float cx = (lo.x + hi.x) * 0.5f, cy = (lo.y + hi.y) * 0.5f, cz = (lo.z + hi.z) * 0.5f;
    const bool xs = r.dx < 0.0f, ys = r.dy < 0.0f, zs = r.dz < 0.0f;
    /* near / far bound picked before the arithmetic: the same values, and
     * results of arithmetic, which the min/max form takes as they are */
    const float nx = ((xs ? hi.x : lo.x) - r.ox) * r.ix, xc = (cx - r.ox) * r.ix, fx = ((xs ? lo.x : hi.x) - r.ox) * r.ix;
    const float ny = ((ys ? hi.y : lo.y) - r.oy) * r.iy, yc = (cy - r.oy) * r.iy, fy = ((ys ? lo.y : hi.y) - r.oy) * r.iy;
    const float nz = ((zs ? hi.z : lo.z) - r.oz) * r.iz, zc = (cz - r.oz) * r.iz, fz = ((zs ? lo.z : hi.z) - r.oz) * r.iz;
    uint32_t out = 0;
    if (r.fast) {
        /* the same tests in min/max form (box_tests_exact), as one compare
         * per child: with no NaN, "exit > 0" is "exit >= 2^-149" (the
         * smallest positive float; denormals are kept), "entry < 1" is
         * "entry <= 1 - 2^-24", so entry <= exit, entry < 1, exit > 0 and
         * entry <= bound together are max(entry, 2^-149) <= min(exit, c)
         * with c = min(bound, 1 - 2^-24), given bound >= 2^-149 (it is
         * +inf, FLT_MAX or a hit distance times PRUNE_SLACK, and hits are
         * at t >= 1e-6); the extra terms go into the per-axis values
         * shared by the children */
        const float lo_t = 0x1p-149f, c = fminf(bound, 0x1.fffffep-1f);
        const float nx2 = fmaxf(nx, lo_t), xc2 = fmaxf(xc, lo_t), fx2 = fminf(fx, c), xc3 = fminf(xc, c);
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const float txmin = (k & 1) ? xc2 : nx2, txmax = (k & 1) ? fx2 : xc3;
            const float tymin = (k & 2) ? yc : ny, tymax = (k & 2) ? fy : yc;
            const float tzmin = (k & 4) ? zc : nz, tzmax = (k & 4) ? fz : zc;
            const float a = fmaxf(fmaxf(txmin, tymin), tzmin), b = fminf(fminf(txmax, tymax), tzmax);
            out |= (a <= b) ? (1u << k) : 0u;
        }
        return out & exist;
    }
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const float txmin = (k & 1) ? xc : nx, txmax = (k & 1) ? fx : xc;
        const float tymin = (k & 2) ? yc : ny, tymax = (k & 2) ? fy : yc;
        const float tzmin = (k & 4) ? zc : nz, tzmax = (k & 4) ? fz : zc;
        float tmin = txmin, tmax = txmax;
        bool hit = !((tmin > tymax) || (tymin > tmax));
        if (tymin > tmin) tmin = tymin;
        if (tymax < tmax) tmax = tymax;
        hit = hit && !((tmin > tzmax) || (tzmin > tmax));
        if (tzmin > tmin) tmin = tzmin;
        if (tzmax < tmax) tmax = tzmax;
        hit = hit && (tmin < 1.0f) && (tmax > 0.0f) && !(tmin > bound);
        out |= hit ? (1u << k) : 0u;
    }
    return out & exist;
}

/* child mask permuted into near-to-far order: bit k <-> child (k ^ dmask) */
__device__ __forceinline__ uint32_t order_mask(uint32_t m, uint32_t dm)
{
    if (dm & 1u) m = ((m & 0x55u) << 1) | ((m >> 1) & 0x55u);
    if (dm & 2u) m = ((m & 0x33u) << 2) | ((m >> 2) & 0x33u);
    if (dm & 4u) m = ((m & 0x0fu) << 4) | ((m >> 4) & 0x0fu);
    return m;
}

/* object matrices at a ray time: linear translation keys, or the
 * keyframed row of that msec */
__device__ __forceinline__ void obj_inv_at(const DevObj &o, int time, float I[12])
{
    /* the matrix's source chosen first, then loaded once (loading the
     * static one and overwriting it held both in registers) */
    const float *m = o.moving == 2 ? time_row(o.tab, o.tab_t0, o.tab_n, time, 24) + 12 : o.inv;
#pragma unroll
    for (int k = 0; k < 12; k++) I[k] = m[k];
    if (o.moving != 1) return;
    float s;
    if (time <= o.mt0) s = 0.0f;
    else if (time >= o.mt1) s = 1.0f;
    else s = (float)(time - o.mt0) / (float)(o.mt1 - o.mt0);
    V3 dl = v3(o.motion[0] * s, o.motion[1] * s, o.motion[2] * s);
    V3 rd = xf_vec(o.inv, dl);
    I[3] = I[3] - rd.x;
    I[7] = I[7] - rd.y;
    I[11] = I[11] - rd.z;
}
__device__ __forceinline__ void obj_xf_at(const DevObj &o, int time, float M[12])
{
    const float *m = o.moving == 2 ? time_row(o.tab, o.tab_t0, o.tab_n, time, 24) : o.xf;
#pragma unroll
    for (int k = 0; k < 12; k++) M[k] = m[k];
    if (o.moving != 1) return;
    float s;
    if (time <= o.mt0) s = 0.0f;
    else if (time >= o.mt1) s = 1.0f;
    else s = (float)(time - o.mt0) / (float)(o.mt1 - o.mt0);
    M[3] = M[3] + o.motion[0] * s;
    M[7] = M[7] + o.motion[1] * s;
    M[11] = M[11] + o.motion[2] * s;
}

/* Sphere::intersect, src/sphere.cc:85-132 (t only).  The quadratic is
 * evaluated in double: with |dir| = 1e4 the discriminant cancels about
 * three digits, which in float would put hit points 1e-5 off the surface
 * and self-shadow where the reference (double) does not. */
__device__ __forceinline__ bool sphere_td(double rad, V3 o, V3 d, double &t)
{
    double ox = o.x, oy = o.y, oz = o.z, dx = d.x, dy = d.y, dz = d.z;
    double a = dx * dx + dy * dy + dz * dz;
    double b = 2.0 * dx * ox + 2.0 * dy * oy + 2.0 * dz * oz;
    double c = ox * ox + oy * oy + oz * oz - rad * rad;
    double disc = b * b - 4.0 * a * c;
    if (disc < 0.0) return false;
    double sd = sqrt(disc);
    double t1 = (-b + sd) / (2.0 * a);
    double t2 = (-b - sd) / (2.0 * a);
    const double e = SRAY_ERROR_MARGIN;
    if ((t1 < e && t2 < e) || (t1 > 1.0 && t2 > 1.0)) return false;
    if (t1 < e) t1 = t2;
    if (t2 < e) t2 = t1;
    t = t1 < t2 ? t1 : t2;
    return true;
}
__device__ __forceinline__ bool sphere_t(float rad, V3 o, V3 d, float &t)
{
    double td;
    if (!sphere_td((double)rad, o, d, td)) return false;
    t = (float)td;
    return true;
}

/* Cylinder::intersect, src/cylinder.cc:125-185 (t only; double as above) */
__device__ __forceinline__ bool cylinder_td(double rad, double ys, double ye, V3 o, V3 d, double &t)
{
    double ox = o.x, oy = o.y, oz = o.z, dx = d.x, dy = d.y, dz = d.z;
    double a = dx * dx + dz * dz;
    double b = 2.0 * dx * ox + 2.0 * dz * oz;
    double c = ox * ox + oz * oz - rad * rad;
    double disc = b * b - 4.0 * a * c;
    if (disc < 0.0) return false;
    double sd = sqrt(disc);
    double t1 = (-b + sd) / (2.0 * a);
    double t2 = (-b - sd) / (2.0 * a);
    const double e = SRAY_ERROR_MARGIN;
    if ((t1 < e && t2 < e) || (t1 > 1.0 && t2 > 1.0)) return false;
    double y1 = oy + dy * t1;
    double y2 = oy + dy * t2;
    bool v0 = true, v1 = true;
    if (t1 < e || y1 < ys || y1 >= ye) {
        v0 = false;
        t1 = t2;
    }
    if (t2 < e || y2 < ys || y2 >= ye) {
        v1 = false;
        t2 = t1;
    }
    if (!v0 && !v1) return false;
    t = t1 < t2 ? t1 : t2;
    return true;
}
__device__ __forceinline__ bool cylinder_t(float rad, float ys, float ye, V3 o, V3 d, float &t)
{
    double td;
    if (!cylinder_td((double)rad, (double)ys, (double)ye, o, d, td)) return false;
    t = (float)td;
    return true;
}

/* Triangle::intersect + check_clock, src/mesh.cc:117-165, 310-315.  A
 * plane distance strictly beyond bound cannot win (the caller's best is
 * closer), so the edge tests are skipped for it; equal distances go on. */
__device__ __forceinline__ bool tri_t(float4 A, float4 B, float4 C, const TRay &r, float &t, float bound = FLT_MAX)
{
    V3 v0 = v3(A.x, A.y, A.z), v1 = v3(B.x, B.y, B.z), v2 = v3(C.x, C.y, C.z);
    V3 n = v3(A.w, B.w, C.w); /* (v1-v0) x (v2-v0), stored at upload */
    V3 d = v3(r.dx, r.dy, r.dz), o = v3(r.ox, r.oy, r.oz);
    float ndd = dot(n, d);
    if (fabsf(ndd) < EPS) return false;
    V3 vo = sub(o, v0);
    float tt = -dot(n, vo) / ndd;
    if (tt < EPS || tt > 1.0f || tt > bound) return false;
    V3 p = add(o, mul(d, tt));
    if (!(dot(cross(sub(v1, v0), sub(p, v0)), n) >= 0.0f)) return false;
    if (!(dot(cross(sub(v2, v1), sub(p, v1)), n) >= 0.0f)) return false;
    if (!(dot(cross(sub(v0, v2), sub(p, v2)), n) >= 0.0f)) return false;
    t = tt;
    return true;
}

struct TraceCounts {
    uint32_t nodes, spheres, cylinders, triangles, item_refs, mesh_enters, item_boxes;
};
constexpr int NCOUNT = 7; /* device counters per kind of traversal */

/* adds a lane's visit counts to the device counters */
template <bool COUNT>
__device__ __forceinline__ void flush_counts(const TraceCounts &cnt, unsigned long long *c)
{
    if (!COUNT) return;
    atomicAdd(&c[0], (unsigned long long)cnt.nodes);
    atomicAdd(&c[1], (unsigned long long)cnt.spheres);
    atomicAdd(&c[2], (unsigned long long)cnt.cylinders);
    atomicAdd(&c[3], (unsigned long long)cnt.triangles);
    atomicAdd(&c[4], (unsigned long long)cnt.item_refs);
    atomicAdd(&c[5], (unsigned long long)cnt.mesh_enters);
    atomicAdd(&c[6], (unsigned long long)cnt.item_boxes);
}

struct HitRec {
    float t;
    int32_t obj;
    uint32_t tri;  /* global triangle index, 0xffffffff if none */
    uint32_t rank;
};

/* where traverse() takes the world ray from: registers, or a queue record
 * re-read when needed (at the start, entering and leaving a mesh), which
 * keeps the world ray out of registers during the walk */
struct RayRegs {
    V3 o, d;
    int time;
    __device__ __forceinline__ void get(V3 &wo, V3 &wd, int &t) const
    {
        wo = o;
        wd = d;
        t = time;
    }
};
template <class R> struct RayRecord {
    const R *base;
    uint32_t i;
    __device__ __forceinline__ void get(V3 &wo, V3 &wd, int &t) const
    {
        uint32_t k = i;
        asm volatile("" : "+v"(k)); /* re-read from memory each time, not kept */
        float4 a = base[k].o, b = base[k].d;
        wo = v3(a.x, a.y, a.z);
        wd = v3(b.x, b.y, b.z);
        t = __float_as_int(a.w);
    }
};

/* wave-wide minimum of a 64-bit key (every lane of the wave takes part) */
__device__ __forceinline__ unsigned long long wave_min_u64(unsigned long long v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        unsigned long long o = __shfl_xor(v, off);
        v = o < v ? o : v;
    }
    return v;
}
__device__ __forceinline__ float read_lane_f(float v, int j)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), j));
}
__device__ __forceinline__ uint32_t read_lane_u(uint32_t v, int j)
{
    return (uint32_t)__builtin_amdgcn_readlane((int)v, j);
}
template <class T> __device__ __forceinline__ const T *read_lane_p(const T *p, int j)
{
    const uint64_t v = (uint64_t)(uintptr_t)p;
    const uint32_t lo = read_lane_u((uint32_t)v, j), hi = read_lane_u((uint32_t)(v >> 32), j);
    return (const T *)(uintptr_t)(((uint64_t)hi << 32) | lo);
}

/* Up to this many lanes of a wave in mesh leaves at once, the whole wave
 * tests their leaves, one triangle per lane, several leaves per round when
 * their items fit the wave; with more, each lane walks its own leaf (item
 * boxes first).  Long rays (grazing a heightfield, say) keep a wave alive
 * long after its other lanes are done, and set how long a bounce level
 * takes; in such a wave one step of 64 loads replaces a leaf's serial chain
 * of loads.  Measured on the heightfield with one leaf per round
 * (closest-hit / any-hit thresholds 1/2, 2/4, 4/8, 8/8, 8/16, 12/8, 16/8
 * gave 6.06, 5.84, 5.36, 5.23, 5.29, 5.37, 5.65 ms per frame): 8 for both. */
template <bool ANY> constexpr int coop_lanes() { return 8; }

/* Node records a lane fetches per round of the walk before the wave moves
 * on to the item tests.  Walking on until every lane has items (no cap)
 * keeps the lanes that reached a leaf early waiting through the other
 * lanes' node visits; with a cap they test their items in the same round.
 * Measured on the heightfield (frame ms, caps 1 / 2 / 3 / 4 / none):
 * 5.90 / 5.29 / 5.33 / 5.27-5.38 / 5.85. */
constexpr int NODE_ROUND = 2;

/* Triangles a lane keeps in flight when it tests its own leaf.  The deep
 * bounce levels hold few rays (two or three waves per SIMD), too few to
 * hide one load latency per triangle with other waves' work; the SPARSE
 * instances, launched for those levels, load SPARSE_TRIS triangles at once.
 * With item boxes tested first, few triangles of a leaf remain: 2 was
 * measured faster than 4 (4.84 -> 4.79 ms per heightfield frame). */
constexpr int SPARSE_TRIS = 2;

/* The fields of a scene descriptor the walk reads, for a wave whose lanes
 * all walk the same scene: read once into scalar registers
 * (trav_scene_uniform).  Read through the descriptor in memory instead,
 * the node arrays' addresses are loaded again at every node visit, a
 * second memory round trip in front of the node's own load. */
struct TravScene {
    const float4 *node;
    const uint32_t *items;
    const SceneItem *sitem;
    const DevObj *objs;
    const float4 *ltri;
    const uint2 *lbox;
    uint32_t scene_root;
    int32_t fast_boxes;
};

template <class T> __device__ __forceinline__ T *uniform_ptr(T *p)
{
    const uint64_t v = (uint64_t)(uintptr_t)p;
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return (T *)(uintptr_t)(((uint64_t)hi << 32) | lo);
}

/* S must be the same descriptor in every active lane */
__device__ __forceinline__ TravScene trav_scene_uniform(const SceneDev &S)
{
    TravScene t;
    t.node = uniform_ptr(S.node);
    t.items = uniform_ptr(S.items);
    t.sitem = uniform_ptr(S.sitem);
    t.objs = uniform_ptr(S.objs);
    t.ltri = uniform_ptr(S.ltri);
    t.lbox = uniform_ptr(S.lbox);
    t.scene_root = __builtin_amdgcn_readfirstlane(S.scene_root);
    t.fast_boxes = __builtin_amdgcn_readfirstlane(S.fast_boxes);
    return t;
}

/* Scene::cast_ray over the octrees.  ANY: stop at the first hit (the
 * reference's shadow test cast_ray(ray) != NULL, src/shader.cc:266).
 * stk points at this lane's entry 0; consecutive entries are `stride`
 * apart (LDS, lane-interleaved).
 *
 * A node record is fetched only for a child whose box the segment meets:
 * when a node is entered its children are tested from its own box
 * (expand_children) and only the hit ones are pushed.  A popped child is
 * tested again from its fetched box (the same floats, so it passes) to
 * prune it against the best distance found since the push, and the pseudo
 * entries of the roots get their real test there.  In min/max form that
 * test is left out for inner nodes: their children's tests, against the
 * same bound, then all fail.
 *
 * COOP: every lane of the wave calls this (live = false for a lane without
 * a ray) and stays in the walk until all lanes are done, so that sparse
 * waves can test leaves together.  A leaf tested by the wave yields the
 * same winner as tested by its own lane: the smallest (t, item position)
 * key, merged with the rule the serial loop applies. */
template <bool ANY, bool COUNT, bool COOP, bool SPARSE, class SRC, class SC>
__device__ __forceinline__ HitRec traverse_src(const SC &S, const SRC &src, bool live, uint32_t *stk,
                                               int stride, TraceCounts &cnt)
{
    HitRec best;
    best.t = FLT_MAX;
    best.obj = -1;
    best.tri = 0xffffffffu;
    best.rank = 0xffffffffu;
    TRay r;
    int level = 0;
    float m_t = FLT_MAX;
    uint32_t m_rank = 0xffffffffu; /* item position of the mesh's best triangle */
    int cache_obj = -1;
    float cache_t = FLT_MAX;
    uint32_t cache_tri = 0xffffffffu;
    uint32_t item_cur = 0, item_end = 0;
    int cur_mesh = -1;
    int sp = 0;
    bool done = !live;
    /* the cell of the mesh leaf being tested (its item boxes are coded
     * relative to it) */
    float4 c_lo = make_float4(0.0f, 0.0f, 0.0f, 0.0f), c_hi = c_lo;
    /* TF_TIE: distinct items tie at best.t; TF_MTIE: the same within the
     * mesh being walked; TF_EXH: the second, exhaustive walk of a tie ray */
    uint32_t tf = 0;
    constexpr uint32_t TF_TIE = 1u, TF_MTIE = 2u, TF_EXH = 4u;
    /* a triangle of the mesh being walked at item position k (tie_in: it
     * ties with another triangle of its leaf) */
    auto mesh_offer = [&](float t, uint32_t k, bool tie_in) {
        const bool eq = t == m_t;
        if (eq && !(tf & TF_MTIE) && (tie_in || S.items[k] != S.items[m_rank])) tf |= TF_MTIE;
        if (t < m_t || (eq && k < m_rank)) {
            if (!eq) tf = (tf & ~TF_MTIE) | (tie_in ? TF_MTIE : 0u);
            m_t = t;
            m_rank = k;
        }
    };
    /* a scene item at occurrence rank: object oi (triangle tri of a mesh) */
    auto scene_offer = [&](float t, int oi, uint32_t tri, uint32_t rank, bool tie_in) {
        const bool eq = t == best.t;
        if (eq && (tie_in || oi != best.obj || tri != best.tri)) tf |= TF_TIE;
        if (t < best.t || (eq && rank < best.rank)) {
            if (!eq) tf = (tf & ~TF_TIE) | (tie_in ? TF_TIE : 0u);
            best.t = t;
            best.obj = oi;
            best.tri = tri;
            best.rank = rank;
        }
    };
    if (live) {
        V3 wo, wd;
        int time;
        src.get(wo, wd, time);
        r = make_tray_s(wo, wd, S.fast_boxes); /* world ray; the mesh's object ray inside a mesh */
        /* pseudo entry whose only child (slot 0) is the scene root */
        stk[0] = (S.scene_root << 8) | (1u << r.dmask);
        sp = 1;
    } else {
        r = make_tray(v3(0.0f, 0.0f, 0.0f), v3(1.0f, 1.0f, 1.0f));
    }
    for (;;) {
        int nv = 0; /* node records fetched this round */
        if (!done) for (;;) {
            /* while-while: walk the stack until this lane has items to test
             * (or has fetched NODE_ROUND node records this round), so lanes
             * meet at the item tests instead of alternating node visits and
             * leaf loops */
            while (item_cur >= item_end && sp > 0 && (!COOP || nv < NODE_ROUND)) {
                uint32_t e = stk[(sp - 1) * stride];
                if ((e & E_MARK) == E_MARK) {
                    sp--;
                    uint32_t rank = e & 0x3fffffffu;
                    const int oi = cur_mesh; /* the object at rank */
                    cache_obj = oi;
                    if (!ANY) {
                        uint32_t m_tri = m_rank != 0xffffffffu ? S.items[m_rank] : 0xffffffffu;
                        cache_t = m_t;
                        cache_tri = m_tri;
                        if (m_tri != 0xffffffffu) scene_offer(m_t, oi, m_tri, rank, (tf & TF_MTIE) != 0u);
                    }
                    {
                        V3 o2, d2;
                        int t2;
                        src.get(o2, d2, t2);
                        r = make_tray_s(o2, d2, S.fast_boxes); /* back to the world ray (same floats) */
                    }
                    level = 0;
                    continue;
                }
                if (e & E_RES) {
                    item_end = e & 0x3fffffffu;
                    item_cur = stk[(sp - 2) * stride];
                    sp -= 2;
                    continue;
                }
                uint32_t pend = e & 0xffu;
                uint32_t k = (uint32_t)__builtin_ctz(pend);
                pend &= pend - 1u;
                if (pend == 0u) sp--;
                else stk[(sp - 1) * stride] = (e & ~0xffu) | pend;
                uint32_t child = (e >> 8) + (k ^ r.dmask);
                const float4 *nd = S.node + 2 * (size_t)child;
                float4 lo = nd[0], hi = nd[1];
                nv++;
                if (COUNT) cnt.nodes++;
                uint32_t meta = __float_as_uint(hi.w), first = __float_as_uint(lo.w);
                float bound = FLT_MAX;
                if (!ANY && !(tf & TF_EXH)) bound = (level ? fminf(best.t, m_t) : best.t) * PRUNE_SLACK;
                if (r.fast && !(meta & LEAF)) {
                    /* an inner node in min/max form: every child's entry is
                     * at least the node's and its exit at most the node's,
                     * so when the node's own test fails (a miss, or an
                     * entry beyond bound) every child test fails too, and
                     * the node's test is left out */
                    uint32_t hitm = expand_children(lo, hi, r, order_mask(meta, r.dmask), bound);
                    if (hitm) stk[(sp++) * stride] = (first << 8) | hitm;
                    continue;
                }
                float tmin;
                if (!slab(lo, hi, r, tmin)) continue;
                if (!ANY && tmin > bound) continue;
                if (meta & LEAF) {
                    item_cur = first;
                    item_end = first + (meta & ~LEAF);
                    c_lo = lo;
                    c_hi = hi;
                } else {
                    uint32_t hitm = expand_children(lo, hi, r, order_mask(meta, r.dmask), bound);
                    if (hitm) stk[(sp++) * stride] = (first << 8) | hitm;
                }
            }
            if (item_cur < item_end) break;
            if (sp > 0) break; /* this round's node records fetched */
            /* stack empty: a tie ray is walked again, skipping nothing */
            if (!ANY && (tf & (TF_TIE | TF_EXH)) == TF_TIE) {
                tf = TF_EXH;
                best.t = FLT_MAX;
                best.obj = -1;
                best.tri = 0xffffffffu;
                best.rank = 0xffffffffu;
                m_t = FLT_MAX;
                m_rank = 0xffffffffu;
                cache_obj = -1;
                cache_t = FLT_MAX;
                cache_tri = 0xffffffffu;
                cur_mesh = -1;
                level = 0;
                V3 wo, wd;
                int time;
                src.get(wo, wd, time);
                r = make_tray_s(wo, wd, S.fast_boxes);
                stk[0] = (S.scene_root << 8) | (1u << r.dmask);
                sp = 1;
                continue;
            }
            done = true;
            break;
        }
        if (COOP) {
            if (__ballot(!done) == 0ull) break;
        } else if (done) {
            break;
        }
        const bool in_leaf = !done && level != 0 && item_cur < item_end;
        if (COOP) {
            unsigned long long lm = __ballot(in_leaf);
            if (lm != 0ull && __popcll(lm) <= coop_lanes<ANY>()) {
                /* The wave tests the leaf lanes' leaves, one triangle per
                 * lane: in rounds of consecutive leaf lanes whose items fit
                 * the 64 lanes together (a leaf of more items alone, in
                 * passes of 64).  Each leaf's winner is the smallest (t, item
                 * position) key over its own lanes, merged with the rule
                 * the serial loop applies. */
                const uint32_t lane = __lane_id();
                const uint32_t mycnt = in_leaf ? item_end - item_cur : 0u;
                while (lm != 0ull) {
                    constexpr int NSEG = coop_lanes<ANY>();
                    int segl[NSEG];
                    uint32_t segp[NSEG];
                    bool segv[NSEG];
                    uint32_t total = 0;
                    bool open = true;
#pragma unroll
                    for (int q = 0; q < NSEG; q++) {
                        segv[q] = false;
                        segl[q] = 0;
                        segp[q] = total;
                        if (open && lm != 0ull) {
                            const int j = __ffsll(lm) - 1;
                            const uint32_t cj = read_lane_u(mycnt, j);
                            if (q > 0 && total + cj > 64u) {
                                open = false;
                            } else {
                                segv[q] = true;
                                segl[q] = j;
                                segp[q] = total;
                                total += cj;
                                lm &= lm - 1ull;
                                if (total > 64u) open = false;
                            }
                        }
                    }
                    if (total > 64u) {
                        /* one leaf of more than 64 items: passes of 64 */
                        const int j = segl[0];
                        TRay rj;
                        rj.ox = read_lane_f(r.ox, j);
                        rj.oy = read_lane_f(r.oy, j);
                        rj.oz = read_lane_f(r.oz, j);
                        rj.dx = read_lane_f(r.dx, j);
                        rj.dy = read_lane_f(r.dy, j);
                        rj.dz = read_lane_f(r.dz, j);
                        const uint32_t cur = read_lane_u(item_cur, j), end = read_lane_u(item_end, j);
                        const float bnd = ANY ? FLT_MAX : read_lane_f(fminf(m_t, best.t), j);
                        unsigned long long key = ~0ull;
                        bool ltie = false; /* two of this lane's triangles tie at its best */
                        for (uint32_t base = cur; base < end; base += 64u) {
                            const uint32_t k = base + lane;
                            if (k < end) {
                                const float4 *L = S.ltri + (size_t)k * 3;
                                float4 A = L[0], B = L[1], C = L[2];
                                if (COUNT) {
                                    cnt.item_refs++;
                                    cnt.triangles++;
                                }
                                float t;
                                if (tri_t(A, B, C, rj, t, bnd)) {
                                    const unsigned long long kk = ((unsigned long long)__float_as_uint(t) << 32) | k;
                                    if (!ANY && (uint32_t)(key >> 32) == __float_as_uint(t)) ltie = true;
                                    key = kk < key ? kk : key;
                                }
                            }
                        }
                        const unsigned long long mine = key;
                        key = wave_min_u64(key);
                        /* positions within one leaf are distinct triangles:
                         * any other lane at the leaf's best t is a tie */
                        bool leaf_tie = false;
                        if (!ANY)
                            leaf_tie = __ballot(mine != ~0ull && (uint32_t)(mine >> 32) == (uint32_t)(key >> 32) &&
                                                (mine != key || ltie)) != 0ull;
                        if (lane == (uint32_t)j) {
                            if (key != ~0ull) {
                                const float t = __uint_as_float((uint32_t)(key >> 32));
                                const uint32_t kpos = (uint32_t)key;
                                if (ANY) {
                                    best.t = t;
                                    best.tri = S.items[kpos];
                                    best.obj = cur_mesh;
                                    done = true;
                                } else {
                                    mesh_offer(t, kpos, leaf_tie);
                                }
                            }
                            item_cur = item_end;
                        }
                        continue;
                    }
                    /* a packed round: lane l tests item l - segp[q] of the
                     * leaf of segment q (segp[q] <= l < segp[q + 1]) */
                    int mseg = 0;
#pragma unroll
                    for (int q = 1; q < NSEG; q++)
                        if (segv[q] && lane >= segp[q]) mseg = q;
                    int src = segl[0];
                    uint32_t off = lane - segp[0];
#pragma unroll
                    for (int q = 1; q < NSEG; q++)
                        if (mseg == q) {
                            src = segl[q];
                            off = lane - segp[q];
                        }
                    TRay rj;
                    rj.ox = __shfl(r.ox, src);
                    rj.oy = __shfl(r.oy, src);
                    rj.oz = __shfl(r.oz, src);
                    rj.dx = __shfl(r.dx, src);
                    rj.dy = __shfl(r.dy, src);
                    rj.dz = __shfl(r.dz, src);
                    const uint32_t k = (uint32_t)__shfl((int)item_cur, src) + off;
                    const float bnd = ANY ? FLT_MAX : __shfl(fminf(m_t, best.t), src);
                    unsigned long long key = ~0ull;
                    if (lane < total) {
                        const float4 *L = S.ltri + (size_t)k * 3;
                        float4 A = L[0], B = L[1], C = L[2];
                        if (COUNT) {
                            cnt.item_refs++;
                            cnt.triangles++;
                        }
                        float t;
                        if (tri_t(A, B, C, rj, t, bnd)) key = ((unsigned long long)__float_as_uint(t) << 32) | k;
                    }
#pragma unroll
                    for (int q = 0; q < NSEG; q++) {
                        if (!segv[q]) continue;
                        const unsigned long long mine = (lane < total && mseg == q) ? key : ~0ull;
                        const unsigned long long kmin = wave_min_u64(mine);
                        bool leaf_tie = false;
                        if (!ANY)
                            leaf_tie = __ballot(mine != ~0ull && (uint32_t)(mine >> 32) == (uint32_t)(kmin >> 32) &&
                                                mine != kmin) != 0ull;
                        if (lane == (uint32_t)segl[q]) {
                            if (kmin != ~0ull) {
                                const float t = __uint_as_float((uint32_t)(kmin >> 32));
                                const uint32_t kpos = (uint32_t)kmin;
                                if (ANY) {
                                    best.t = t;
                                    best.tri = S.items[kpos];
                                    best.obj = cur_mesh;
                                    done = true;
                                } else {
                                    mesh_offer(t, kpos, leaf_tie);
                                }
                            }
                            item_cur = item_end;
                        }
                    }
                }
                continue;
            }
        }
        if (COOP && in_leaf && r.fast) {
            /* (only in the wave-wide walks of the level kernels: in the
             * one-lane walks of the chain and photon kernels the registers
             * it holds cost more than it saves; spheres 1601 -> 1742 Mrays/s
             * without it there) */
            /* mesh leaf, item boxes first: a triangle is tested only when
             * the segment meets its box, grown by a margin far above the
             * roundings of the triangle test (which accepts points a few
             * ulps of the coordinates off the triangle) and of these box
             * tests, so no hit the triangle test would report is skipped
             * (the codes are defined with IBOX_STEP in device.h). */
            const float lo3[3] = {c_lo.x, c_lo.y, c_lo.z}, hi3[3] = {c_hi.x, c_hi.y, c_hi.z};
            const float o3[3] = {r.ox, r.oy, r.oz}, i3[3] = {r.ix, r.iy, r.iz};
            float mag = 0.0f;
#pragma unroll
            for (int a = 0; a < 3; a++) mag = fmaxf(mag, fmaxf(fabsf(o3[a]), fmaxf(fabsf(lo3[a]), fabsf(hi3[a]))));
            const float mg = mag * 0x1p-14f;
            /* per axis, the padded planes of code q lie at t = fma(q, dt, t0)
             * (the padding folded into t0); the roundings of this form are
             * far below the padding */
            float tl0[3], th0[3], dt[3];
#pragma unroll
            for (int a = 0; a < 3; a++) {
                const float ext = hi3[a] - lo3[a];
                const float base = lo3[a] - ext * IBOX_GROW;
                dt[a] = (ext * IBOX_STEP) * i3[a];
                tl0[a] = ((base - mg) - o3[a]) * i3[a];
                th0[a] = ((base + mg) - o3[a]) * i3[a];
            }
            /* a hit the triangle test reports lies at t <= tb */
            const float tb = ANY ? 1.0f : fminf(fminf(m_t, best.t), 1.0f);
            auto box_met = [&](uint2 q) {
                float near = -INFINITY, far = INFINITY;
#pragma unroll
                for (int a = 0; a < 3; a++) {
                    const float ql = (float)((q.x >> (8 * a)) & 0xffu);
                    const float qh = (float)(a == 0 ? q.x >> 24 : (q.y >> (8 * (a - 1))) & 0xffu);
                    const float tl = fmaf(ql, dt[a], tl0[a]), th = fmaf(qh, dt[a], th0[a]);
                    near = fmaxf(near, fminf(tl, th));
                    far = fminf(far, fmaxf(tl, th));
                }
                return (q.y & IBOX_OPEN) != 0u || (near <= far && far >= 0.0f && near <= tb);
            };
            /* boxes per load round: 8 for closest-hit rays, 16 for any-hit
             * rays, which often end at their first hit (measured: shadow
             * trace 2.09 -> 1.96 ms per heightfield frame) */
            constexpr uint32_t IB = ANY ? 16 : 8;
            constexpr int TG = SPARSE ? SPARSE_TRIS : 1;    /* triangles per load round */
            while (item_cur < item_end && !done) {
                const uint32_t nb = min(IB, item_end - item_cur);
                uint2 qb[IB];
#pragma unroll
                for (uint32_t j = 0; j < IB; j++) qb[j] = j < nb ? S.lbox[item_cur + j] : make_uint2(0u, 0u);
                if (COUNT) cnt.item_boxes += nb;
                uint32_t mask = 0;
#pragma unroll
                for (uint32_t j = 0; j < IB; j++) mask |= (j < nb && box_met(qb[j])) ? 1u << j : 0u;
                while (mask != 0u && !done) {
                    uint32_t kk[TG];
                    bool ok[TG];
                    float4 V[3 * TG];
#pragma unroll
                    for (int u = 0; u < TG; u++) {
                        ok[u] = mask != 0u;
                        kk[u] = item_cur + (ok[u] ? (uint32_t)__builtin_ctz(mask) : 0u);
                        mask &= ok[u] ? mask - 1u : ~0u;
                    }
#pragma unroll
                    for (int u = 0; u < TG; u++)
                        if (ok[u]) {
                            const float4 *L = S.ltri + (size_t)kk[u] * 3;
                            V[3 * u] = L[0];
                            V[3 * u + 1] = L[1];
                            V[3 * u + 2] = L[2];
                        }
#pragma unroll
                    for (int u = 0; u < TG; u++) {
                        if (!ok[u] || done) continue;
                        if (COUNT) {
                            cnt.item_refs++;
                            cnt.triangles++;
                        }
                        float t;
                        if (tri_t(V[3 * u], V[3 * u + 1], V[3 * u + 2], r, t, ANY ? FLT_MAX : fminf(m_t, best.t))) {
                            if (ANY) {
                                best.t = t;
                                best.tri = S.items[kk[u]];
                                best.obj = cur_mesh;
                                done = true;
                            } else {
                                mesh_offer(t, kk[u], false);
                            }
                        }
                    }
                }
                item_cur += nb;
                /* one batch per round: the wave's other lanes walk on
                 * meanwhile (measured: 5.02 -> 4.93 ms per heightfield
                 * frame) */
                if (COOP) break;
            }
            continue;
        }
        if (in_leaf) {
            /* mesh leaf: all its triangles, from the leaf-ordered copies
             * (dense levels: one per step, as a second triangle's registers
             * in flight cost a wave of occupancy and more than its latency
             * hiding gains; sparse levels: SPARSE_TRIS per step, tested in
             * order) */
            if (SPARSE) {
                constexpr int U = SPARSE_TRIS;
                for (; item_cur + (uint32_t)U <= item_end && !done; item_cur += U) {
                    const float4 *L = S.ltri + (size_t)item_cur * 3;
                    float4 V[3 * U];
#pragma unroll
                    for (int u = 0; u < 3 * U; u++) V[u] = L[u];
#pragma unroll
                    for (int u = 0; u < U; u++) {
                        float t;
                        if (!done && tri_t(V[3 * u], V[3 * u + 1], V[3 * u + 2], r, t,
                                           ANY ? FLT_MAX : fminf(m_t, best.t))) {
                            const uint32_t k = item_cur + (uint32_t)u;
                            if (ANY) {
                                best.t = t;
                                best.tri = S.items[k];
                                best.obj = cur_mesh;
                                done = true;
                            } else {
                                mesh_offer(t, k, false);
                            }
                        }
                    }
                }
                if (done) continue;
            }
            for (; item_cur < item_end;) {
                const float4 *L = S.ltri + (size_t)item_cur * 3;
                float4 A = L[0], B = L[1], C = L[2];
                if (COUNT) {
                    cnt.item_refs++;
                    cnt.triangles++;
                }
                float t;
                if (tri_t(A, B, C, r, t, ANY ? FLT_MAX : fminf(m_t, best.t))) {
                    if (ANY) {
                        best.t = t;
                        best.tri = S.items[item_cur];
                        best.obj = cur_mesh;
                        done = true;
                        break;
                    }
                    mesh_offer(t, item_cur, false);
                }
                item_cur++;
            }
            continue;
        }
        if (done || item_cur >= item_end) continue; /* no items this round */
        uint32_t rank = item_cur++;
        /* the occurrence's record (one load); a moving object's matrices
         * from its DevObj */
        const SceneItem &ob = S.sitem[rank];
        const int oi = ob.obj;
        if (COUNT) cnt.item_refs++;
        const int type = ob.type;
        auto inv_at = [&](int time, float I[12]) {
            if (ob.moving == 0) {
#pragma unroll
                for (int k = 0; k < 12; k++) I[k] = ob.inv[k];
            } else {
                obj_inv_at(S.objs[oi], time, I);
            }
        };
        if (type == SRAY_OBJ_MESH) {
            if (oi == cache_obj) {
                /* this ray has traversed this mesh already (in another scene
                 * leaf): same result; an any-hit ray only gets here after
                 * finding nothing in it */
                if (!ANY && cache_tri != 0xffffffffu) scene_offer(cache_t, oi, cache_tri, rank, false);
                continue;
            }
            if (COUNT) cnt.mesh_enters++;
            if (item_cur < item_end) {
                stk[(sp++) * stride] = item_cur;
                stk[(sp++) * stride] = E_RES | item_end;
            }
            stk[(sp++) * stride] = E_MARK | rank;
            float I[12];
            {
                V3 o2, d2;
                int t2;
                src.get(o2, d2, t2);
                inv_at(t2, I);
                r = make_tray_s(xf_point(I, o2), xf_vec(I, d2), S.fast_boxes);
            }
            stk[(sp++) * stride] = (ob.root << 8) | (1u << r.dmask);
            item_cur = item_end = 0;
            level = 1;
            cur_mesh = oi;
            m_t = FLT_MAX;
            m_rank = 0xffffffffu;
            tf &= ~TF_MTIE;
            continue;
        }
        float I[12];
        V3 lo, ld;
        {
            V3 o2, d2;
            int t2;
            src.get(o2, d2, t2);
            inv_at(t2, I);
            lo = xf_point(I, o2);
            ld = xf_vec(I, d2);
        }
        float t;
        bool h;
        if (type == SRAY_OBJ_SPHERE) {
            if (COUNT) cnt.spheres++;
            h = sphere_t(ob.rad, lo, ld, t);
        } else {
            if (COUNT) cnt.cylinders++;
            h = cylinder_t(ob.rad, ob.start, ob.end, lo, ld, t);
        }
        if (h) {
            if (ANY) {
                best.t = t;
                best.obj = oi;
                done = true;
                continue;
            }
            scene_offer(t, oi, 0xffffffffu, rank, false);
        }
    }
    return best;
}

template <bool ANY, bool COUNT>
__device__ __forceinline__ HitRec traverse(const SceneDev &S, V3 wo, V3 wd, int time, uint32_t *stk, int stride,
                                           TraceCounts &cnt)
{
    RayRegs src{wo, wd, time};
    return traverse_src<ANY, COUNT, false, false>(S, src, true, stk, stride, cnt);
}

} // namespace srd

#endif
