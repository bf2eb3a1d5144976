/*
 * oracle.cc — CPU restatement of the S-ray 0.1 ray-tracing hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h): the checker for the device path
 * and the timed CPU baseline.  Nothing in the product loads it.
 *
 * Each function names the reference file:line it restates.  The reference
 * sources are GPLv3 C++ (S-ray 0.1, src/); this file is an independent
 * restatement of their behaviour, templated on the scalar type so that one
 * source gives both the reference's double arithmetic and a strict float
 * variant whose operation order equals the device kernels'.
 *
 * Arithmetic conventions (both precisions; the device kernels use the same):
 *   dot(a,b)        = (a.x*b.x + a.y*b.y) + a.z*b.z
 *   cross(a,b)      = (a.y*b.z - a.z*b.y, a.z*b.x - a.x*b.z, a.x*b.y - a.y*b.x)
 *   point  M*p      = ((m0*x + m1*y) + m2*z) + m3          (row-major 3x4)
 *   vector M*v      =  (m0*x + m1*y) + m2*z
 *   normal M^-T*n   =  (i0*x + i4*y) + i8*z, (i1 ..) , (i2 ..)
 *   normalize(v)    = v / sqrt(dot(v,v)) per component (true division)
 *   reflect(d,n)    = d - n*(2*dot(d,n))
 * libvmath (where these live in the reference) is not available, so their
 * exact form is fixed here; see the parity notes in the test suite.
 */
#include "oracle.h"

#include <algorithm>
#include <array>
#include <atomic>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

namespace orc {

/* ------------------------------------------------------------------ */
/* counter-based RNG (replaces libc rand()/vmath frand/sphrand, Q20)   */
/* ------------------------------------------------------------------ */
static inline uint64_t splitmix64(uint64_t x)
{
    /* SplitMix64 (Steele, Lea, Flood 2014): one step from state x */
    uint64_t z = x + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static inline uint32_t rng_u32(uint64_t seed, uint32_t a, uint32_t b, uint32_t c, uint32_t d)
{
    uint64_t h = splitmix64(seed);
    h = splitmix64(h ^ (uint64_t)a);
    h = splitmix64(h ^ (((uint64_t)b << 32) | (uint64_t)c));
    h = splitmix64(h ^ (uint64_t)d);
    return (uint32_t)(h >> 32);
}

/* uniform in [0,1) with 24 bits: exactly representable in float */
static inline float rng_unit(uint64_t seed, uint32_t a, uint32_t b, uint32_t c, uint32_t d)
{
    return (float)(rng_u32(seed, a, b, c, d) >> 8) * (1.0f / 16777216.0f);
}

/* RNG "slot" word: which random site a draw belongs to
 *   kind<<30 | i1<<20 | i2<<10 | attempt<<3 | dim
 * kind 0 camera (dim 0,1 jitter, 2 motion-blur time), 1 light sample
 * (i1 light, i2 shadow sample), 2 diffuse GI direction (i1 sample index),
 * 3 photon emission/bounce.  i1, i2 < 1024, attempt < 128, dim < 8. */
static inline uint32_t slot(uint32_t kind, uint32_t i1, uint32_t i2, uint32_t att, uint32_t dim)
{
    return (kind << 30) | (i1 << 20) | (i2 << 10) | (att << 3) | dim;
}
enum : uint32_t { K_CAMERA = 0u, K_LIGHT = 1u, K_DIFFUSE = 2u, K_PHOTON = 3u };

/* ------------------------------------------------------------------ */
/* vector math                                                          */
/* ------------------------------------------------------------------ */
template <class T> struct V3 {
    T x, y, z;
};
template <class T> static inline V3<T> v3(T x, T y, T z) { return V3<T>{x, y, z}; }
template <class T> static inline V3<T> operator+(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> static inline V3<T> operator-(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> static inline V3<T> operator-(V3<T> a) { return {-a.x, -a.y, -a.z}; }
template <class T> static inline V3<T> operator*(V3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <class T> static inline V3<T> operator/(V3<T> a, T s) { return {a.x / s, a.y / s, a.z / s}; }
template <class T> static inline T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> static inline V3<T> cross(V3<T> a, V3<T> b)
{
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class T> static inline T length(V3<T> a) { return std::sqrt(dot(a, a)); }
template <class T> static inline V3<T> normalized(V3<T> a)
{
    T len = length(a);
    return {a.x / len, a.y / len, a.z / len};
}
template <class T> static inline V3<T> reflect(V3<T> d, V3<T> n)
{
    T s = (T)2 * dot(d, n);
    return {d.x - n.x * s, d.y - n.y * s, d.z - n.z * s};
}

template <class T> struct Mat {
    T m[12];
};
template <class T> static inline Mat<T> mat_from(const float *f)
{
    Mat<T> r;
    for (int i = 0; i < 12; i++) r.m[i] = (T)f[i];
    return r;
}
template <class T> static inline V3<T> xf_point(const Mat<T> &M, V3<T> p)
{
    const T *m = M.m;
    return {m[0] * p.x + m[1] * p.y + m[2] * p.z + m[3], m[4] * p.x + m[5] * p.y + m[6] * p.z + m[7],
            m[8] * p.x + m[9] * p.y + m[10] * p.z + m[11]};
}
template <class T> static inline V3<T> xf_vec(const Mat<T> &M, V3<T> v)
{
    const T *m = M.m;
    return {m[0] * v.x + m[1] * v.y + m[2] * v.z, m[4] * v.x + m[5] * v.y + m[6] * v.z,
            m[8] * v.x + m[9] * v.y + m[10] * v.z};
}
/* n' = (M^-1)^T n, given M^-1 */
template <class T> static inline V3<T> xf_normal(const Mat<T> &I, V3<T> n)
{
    const T *m = I.m;
    return {m[0] * n.x + m[4] * n.y + m[8] * n.z, m[1] * n.x + m[5] * n.y + m[9] * n.z,
            m[2] * n.x + m[6] * n.y + m[10] * n.z};
}

template <class T> struct C4 {
    T x, y, z, w;
};
template <class T> static inline C4<T> c4(T x, T y, T z, T w) { return C4<T>{x, y, z, w}; }
template <class T> static inline C4<T> operator+(C4<T> a, C4<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w}; }
template <class T> static inline C4<T> operator-(C4<T> a, C4<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w}; }
template <class T> static inline C4<T> operator*(C4<T> a, C4<T> b) { return {a.x * b.x, a.y * b.y, a.z * b.z, a.w * b.w}; }
template <class T> static inline C4<T> operator*(C4<T> a, T s) { return {a.x * s, a.y * s, a.z * s, a.w * s}; }
template <class T> static inline C4<T> cfrom(const float *f) { return {(T)f[0], (T)f[1], (T)f[2], (T)f[3]}; }

/* ------------------------------------------------------------------ */
/* ray                                                                  */
/* ------------------------------------------------------------------ */
constexpr int IOR_STACK = 16;

template <class T> struct Ray {
    V3<T> o, d;
    int time = 0;
    int iter = 0;
    T energy = 1;
    /* vmath's ior stack: current ior + the iors of the media entered */
    T ior = 1;
    int nstack = 0;
    T stack[IOR_STACK];

    void enter(T new_ior)
    {
        if (nstack < IOR_STACK) stack[nstack++] = ior;
        ior = new_ior;
    }
    void leave() { ior = nstack > 0 ? stack[--nstack] : (T)1; }
    /* ratio (from / to) used by refraction, vmath Ray::calc_ior */
    T calc_ior(bool entering, T mat_ior) const
    {
        if (entering) return ior / mat_ior;
        T to = nstack > 0 ? stack[nstack - 1] : (T)1;
        return ior / to;
    }
};

/* refraction of a unit incident d about n (n faces the incident side);
 * total internal reflection returns the mirror direction. */
template <class T> static inline V3<T> refract(V3<T> d, V3<T> n, T ior)
{
    T cos_inc = -dot(d, n);
    T radical = (T)1 + ior * ior * (cos_inc * cos_inc - (T)1);
    if (radical < (T)0) return reflect(d, n);
    T beta = ior * cos_inc - std::sqrt(radical);
    return {d.x * ior + n.x * beta, d.y * ior + n.y * beta, d.z * ior + n.z * beta};
}

/* ------------------------------------------------------------------ */
/* AABox, src/aabb.cc                                                   */
/* ------------------------------------------------------------------ */
template <class T> struct Box {
    V3<T> mn, mx;
};

/* AABox::in_box, src/aabb.cc:53-65 (closed intervals) */
template <class T> static inline bool in_box(const Box<T> &a, const Box<T> &b)
{
    if (a.mn.x > b.mx.x || b.mn.x > a.mx.x) return false;
    if (a.mn.y > b.mx.y || b.mn.y > a.mx.y) return false;
    if (a.mn.z > b.mx.z || b.mn.z > a.mx.z) return false;
    return true;
}

/* AABox::intersect, src/aabb.cc:81-125 (Williams et al. slab test on the
 * segment t in [0,1]); relies on IEEE 1/0 = inf */
template <class T> static inline bool box_hit(const Box<T> &b, const Ray<T> &r)
{
    const V3<T> bb[2] = {b.mn, b.mx};
    int xs = r.d.x < (T)0;
    T ix = (T)1 / r.d.x;
    T tmin = (bb[xs].x - r.o.x) * ix;
    T tmax = (bb[1 - xs].x - r.o.x) * ix;
    int ys = r.d.y < (T)0;
    T iy = (T)1 / r.d.y;
    T tymin = (bb[ys].y - r.o.y) * iy;
    T tymax = (bb[1 - ys].y - r.o.y) * iy;
    if ((tmin > tymax) || (tymin > tmax)) return false;
    if (tymin > tmin) tmin = tymin;
    if (tymax < tmax) tmax = tymax;
    int zs = r.d.z < (T)0;
    T iz = (T)1 / r.d.z;
    T tzmin = (bb[zs].z - r.o.z) * iz;
    T tzmax = (bb[1 - zs].z - r.o.z) * iz;
    if ((tmin > tzmax) || (tzmin > tmax)) return false;
    if (tzmin > tmin) tmin = tzmin;
    if (tzmax < tmax) tmax = tzmax;
    return (tmin < (T)1) && (tmax > (T)0);
}

/* ------------------------------------------------------------------ */
/* Octree<T>, src/octree.h + src/octree.inl                             */
/* ------------------------------------------------------------------ */
template <class T> struct OctNode {
    Box<T> box;
    std::vector<uint32_t> items;
    int child[8];
    uint32_t rank0 = 0; /* DFS occurrence rank of items[0] (leaves) */
};

template <class T> struct Octree {
    int items_per_node = 5, max_depth = 5;
    std::vector<Box<T>> item_box;
    Box<T> bounds;
    std::vector<OctNode<T>> nodes;
    uint32_t height = 0;

    Octree() { clear(); }
    /* Octree::clear, src/octree.inl:90-98 */
    void clear()
    {
        nodes.clear();
        item_box.clear();
        bounds.mn = v3<T>(FLT_MAX, FLT_MAX, FLT_MAX);
        bounds.mx = v3<T>(-FLT_MAX, -FLT_MAX, -FLT_MAX);
    }
    /* Octree::add, src/octree.inl:100-115 */
    void add(const Box<T> &b)
    {
        item_box.push_back(b);
        if (b.mn.x < bounds.mn.x) bounds.mn.x = b.mn.x;
        if (b.mn.y < bounds.mn.y) bounds.mn.y = b.mn.y;
        if (b.mn.z < bounds.mn.z) bounds.mn.z = b.mn.z;
        if (b.mx.x > bounds.mx.x) bounds.mx.x = b.mx.x;
        if (b.mx.y > bounds.mx.y) bounds.mx.y = b.mx.y;
        if (b.mx.z > bounds.mx.z) bounds.mx.z = b.mx.z;
    }
    int new_node()
    {
        nodes.emplace_back();
        for (int i = 0; i < 8; i++) nodes.back().child[i] = -1;
        return (int)nodes.size() - 1;
    }
    /* Octree::build, src/octree.inl:118-147 */
    void build()
    {
        nodes.clear();
        int root = new_node();
        nodes[root].box = bounds;
        nodes[root].items.resize(item_box.size());
        for (size_t i = 0; i < item_box.size(); i++) nodes[root].items[i] = (uint32_t)i;
        height = 0;
        if ((int)nodes[root].items.size() > items_per_node) subdivide(root, 0);
        /* occurrence ranks in depth-first leaf order */
        uint32_t r = 0;
        assign_ranks(0, r);
    }
    void assign_ranks(int n, uint32_t &r)
    {
        nodes[n].rank0 = r;
        r += (uint32_t)nodes[n].items.size();
        for (int i = 0; i < 8; i++)
            if (nodes[n].child[i] >= 0) assign_ranks(nodes[n].child[i], r);
    }
    /* Octree::subdivide, src/octree.inl:150-191: always 8 children, octant
     * numbering of :161-169, closed-interval overlap assignment */
    void subdivide(int node, int lvl)
    {
        if ((int)nodes[node].items.size() <= items_per_node || lvl >= max_depth) return;
        Box<T> box = nodes[node].box;
        V3<T> c = v3<T>((box.mn.x + box.mx.x) * (T)0.5, (box.mn.y + box.mx.y) * (T)0.5,
                        (box.mn.z + box.mx.z) * (T)0.5);
        Box<T> cb[8];
        cb[0] = {box.mn, c};
        cb[1] = {v3<T>(c.x, box.mn.y, box.mn.z), v3<T>(box.mx.x, c.y, c.z)};
        cb[2] = {v3<T>(c.x, c.y, box.mn.z), v3<T>(box.mx.x, box.mx.y, c.z)};
        cb[3] = {v3<T>(box.mn.x, c.y, box.mn.z), v3<T>(c.x, box.mx.y, c.z)};
        cb[4] = {v3<T>(box.mn.x, box.mn.y, c.z), v3<T>(c.x, c.y, box.mx.z)};
        cb[5] = {v3<T>(c.x, box.mn.y, c.z), v3<T>(box.mx.x, c.y, box.mx.z)};
        cb[6] = {c, box.mx};
        cb[7] = {v3<T>(box.mn.x, c.y, c.z), v3<T>(c.x, box.mx.y, box.mx.z)};
        if ((uint32_t)(lvl + 1) > height) height = (uint32_t)(lvl + 1);
        for (int i = 0; i < 8; i++) {
            int ch = new_node();
            nodes[node].child[i] = ch;
            nodes[ch].box = cb[i];
            std::vector<uint32_t> its;
            for (uint32_t it : nodes[node].items)
                if (in_box(item_box[it], cb[i])) its.push_back(it);
            nodes[ch].items.swap(its);
            subdivide(ch, lvl + 1);
        }
        nodes[node].items.clear();
        nodes[node].items.shrink_to_fit();
    }
};

/* running closest-hit record of the exhaustive DFS, src/octree.inl:200-241:
 * strict '<' keeps the first occurrence in depth-first order */
template <class T> struct Best {
    T t = (T)FLT_MAX;
    int64_t item = -1;
    int64_t sub = -1; /* triangle of a mesh item */
    uint32_t rank = 0;
    int tie = 0;
    bool found() const { return item >= 0; }
    void offer(T tt, int64_t it, int64_t s, uint32_t rk)
    {
        if (tt < t) {
            t = tt;
            item = it;
            sub = s;
            rank = rk;
            tie = 0;
        } else if (tt == t && item >= 0 && (it != item || s != sub)) {
            tie = 1;
        }
    }
};

/* Octree<T>::intersect_rec, src/octree.inl:200-241, as a depth-first walk:
 * every node whose box meets the segment, children 0..7 in order, leaf
 * items in list order.  fn(item, rank, best) tests one item occurrence. */
template <class T, class F>
static bool oct_walk(const Octree<T> &tr, int n, const Ray<T> &r, F &fn, uint64_t &nodes_visited)
{
    const OctNode<T> &nd = tr.nodes[n];
    nodes_visited++;
    if (!box_hit(nd.box, r)) return false;
    if (!nd.items.empty()) {
        for (size_t i = 0; i < nd.items.size(); i++)
            if (fn(nd.items[i], nd.rank0 + (uint32_t)i)) return true;
    } else {
        for (int i = 0; i < 8; i++)
            if (nd.child[i] >= 0 && oct_walk(tr, nd.child[i], r, fn, nodes_visited)) return true;
    }
    return false;
}

/* ------------------------------------------------------------------ */
/* scene                                                                */
/* ------------------------------------------------------------------ */
template <class T> struct SurfPoint {
    T dist;
    V3<T> pos, normal, tangent;
    T u = 0, v = 0;
};

template <class T> struct Object {
    int type, mat, mesh;
    T rad, start, end;
    Mat<T> xf, inv;
    V3<T> motion;
    int mt0, mt1;
    bool moving;
    std::vector<float> rows; /* xform + inverse per msec of the frame interval (keyframed) */
};

template <class T> struct Tri {
    V3<T> v[3];
    V3<T> n[3], g[3];
    T tu[3], tv[3];
};

template <class T> struct MeshT {
    uint32_t start, count;
    Octree<T> tree;
};

template <class T> struct Light {
    int type;
    V3<T> pos;
    C4<T> color;
    T radius;
    V3<T> dim;
    Mat<T> xf;
    std::vector<float> rows; /* position per msec of the frame interval (a position track) */
};

template <class T> struct Material {
    C4<T> kd, ks;
    T shininess, reflect, refract, ior;
    uint32_t flags;
    int map[SRAY_NUM_MAPS] = {0, 0, 0, 0, 0, 0, 0}; /* 1 + texture index, 0 = none */
    T nscale = 0;                                  /* the "normal" attribute's value */
};

/* Texture (src/texture.h:37-60): RGBA float texels, top row first */
struct Texture {
    std::vector<float> px;
    int w = 0, h = 0, filter = SRAY_TEX_FILTER_LINEAR, wrap = SRAY_TEX_WRAP_REPEAT;
    float xf[12] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0};
};

struct Photon {
    float pos[3], dir[3], norm[3], col[3];
};

struct Counters {
    uint64_t nodes = 0, prims = 0;
};

/* Photons bucketed in a uniform grid of cells of side h >= the gather
 * radius, in index order within a cell.  A photon that passes the range
 * test d^2 <= r^2 is within r (to a rounding) of the query on every axis,
 * so with h = 1.01 r its cell is the query's or a neighbour: the 27 cells
 * around the query's hold every photon the brute-force loop would accept.
 * The candidates then take the brute-force test and order, so results are
 * identical to testing every photon (kdtree's kd_nearest_range3f is a
 * range search, src/pmap.cc:114,149: only the found set matters). */
struct PhotonGrid {
    double rf = -1.0; /* float-rounded radius it was built for (-1: none) */
    double o[3] = {0.0, 0.0, 0.0}, h = 1.0;
    int64_t n[3] = {0, 0, 0};
    std::vector<uint32_t> start, idx;
};

template <class T> struct Scene {
    std::vector<Object<T>> objs;
    std::vector<MeshT<T>> meshes;
    std::vector<Tri<T>> tris;
    std::vector<Material<T>> mats;
    std::vector<Texture> texs;
    std::vector<Light<T>> lights;
    Mat<T> cam_xf;
    double vfov;
    T max_dist;
    int shutter;
    C4<T> env_ambient, env_color;
    Octree<T> tree;
    bool use_tree = true;
    bool q9_retraverse = false;
    double gather_dist_abs = 0;
    double diag = 0;
    int frame_t0 = 0, nrow = 1;
    std::vector<float> cam_rows; /* camera matrix per msec (an animated camera) */
    std::vector<Photon> pm[2];
    PhotonGrid grid[2];
    bool brute_gather = false;
    static constexpr T eps = (T)SRAY_ERROR_MARGIN;

    /* (re)builds the bucket grid of map kind for a gather radius (not
     * thread-safe: called before a render's or a gather's workers start) */
    void build_grid(int kind, double radius)
    {
        const double rf = (double)(float)radius;
        PhotonGrid &g = grid[kind];
        if (brute_gather || g.rf == rf || !(rf > 0.0)) return;
        g = PhotonGrid();
        const std::vector<Photon> &ph = pm[kind];
        double mn[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, mx[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
        for (const Photon &p : ph)
            for (int d = 0; d < 3; d++) {
                mn[d] = std::min(mn[d], (double)p.pos[d]);
                mx[d] = std::max(mx[d], (double)p.pos[d]);
            }
        if (ph.empty())
            for (int d = 0; d < 3; d++) mn[d] = mx[d] = 0.0;
        for (int d = 0; d < 3; d++)
            if (!std::isfinite(mn[d]) || !std::isfinite(mx[d])) return; /* brute force for this map */
        double h = rf * 1.01;
        for (;;) {
            double cells = 1.0;
            for (int d = 0; d < 3; d++) cells *= std::floor((mx[d] - mn[d]) / h) + 1.0;
            if (cells <= (double)(1u << 24)) break;
            h *= 2.0;
        }
        for (int d = 0; d < 3; d++) {
            g.o[d] = mn[d];
            g.n[d] = (int64_t)std::floor((mx[d] - mn[d]) / h) + 1;
        }
        g.h = h;
        const size_t nc = (size_t)(g.n[0] * g.n[1] * g.n[2]);
        std::vector<uint32_t> cell(ph.size());
        g.start.assign(nc + 1, 0);
        for (size_t i = 0; i < ph.size(); i++) {
            int64_t c[3];
            for (int d = 0; d < 3; d++)
                c[d] = std::min<int64_t>(std::max<int64_t>((int64_t)std::floor(((double)ph[i].pos[d] - g.o[d]) / h), 0),
                                         g.n[d] - 1);
            cell[i] = (uint32_t)((c[2] * g.n[1] + c[1]) * g.n[0] + c[0]);
            g.start[cell[i] + 1]++;
        }
        for (size_t c = 0; c < nc; c++) g.start[c + 1] += g.start[c];
        g.idx.resize(ph.size());
        std::vector<uint32_t> fill(g.start.begin(), g.start.end() - 1);
        for (size_t i = 0; i < ph.size(); i++) g.idx[fill[cell[i]]++] = (uint32_t)i;
        g.rf = rf;
    }
    void photons_changed(int kind) { grid[kind] = PhotonGrid(); }

    /* the row of a per-msec table at a time: the scene's frame interval,
     * the nearest row outside it */
    size_t row_of(int time) const
    {
        int k = time - frame_t0;
        return (size_t)(k < 0 ? 0 : k >= nrow ? nrow - 1 : k);
    }
    /* Light::get_position(msec) (src/anim.cc XFormNode tracks, evaluated
     * per msec by the scene loader) */
    V3<T> light_pos(const Light<T> &lt, int time) const
    {
        if (lt.rows.empty()) return lt.pos;
        const float *r = &lt.rows[row_of(time) * 3];
        return v3<T>(r[0], r[1], r[2]);
    }

    /* object matrices at a time: keyframed rows (XFormNode::get_xform_matrix
     * and its inverse at each msec, src/anim.cc:391-427), or linear
     * translation keys (src/anim.inl:188-261) */
    void xform_at(const Object<T> &o, int time, Mat<T> &M, Mat<T> &I) const
    {
        if (!o.rows.empty()) {
            const float *r = &o.rows[row_of(time) * 24];
            M = mat_from<T>(r);
            I = mat_from<T>(r + 12);
            return;
        }
        M = o.xf;
        I = o.inv;
        if (!o.moving) return;
        T s;
        if (time <= o.mt0) s = 0;
        else if (time >= o.mt1) s = 1;
        else s = (T)(time - o.mt0) / (T)(o.mt1 - o.mt0);
        V3<T> dlt = o.motion * s;
        M.m[3] = M.m[3] + dlt.x;
        M.m[7] = M.m[7] + dlt.y;
        M.m[11] = M.m[11] + dlt.z;
        V3<T> rd = xf_vec(o.inv, dlt);
        I.m[3] = I.m[3] - rd.x;
        I.m[7] = I.m[7] - rd.y;
        I.m[11] = I.m[11] - rd.z;
    }

    /* Object::calc_bounds overrides: sphere src/sphere.cc:27-34, cylinder
     * src/cylinder.cc:63-86, mesh src/mesh.cc:169-188 */
    Box<T> bounds(const Object<T> &o, int time) const
    {
        Mat<T> M, I;
        xform_at(o, time, M, I);
        Box<T> b;
        if (o.type == SRAY_OBJ_SPHERE) {
            V3<T> p = xf_point(M, v3<T>(0, 0, 0));
            V3<T> rv = v3<T>(o.rad, o.rad, o.rad);
            b.mn = p - rv;
            b.mx = p + rv;
        } else if (o.type == SRAY_OBJ_CYLINDER) {
            V3<T> s1 = xf_point(M, v3<T>(0, o.start, 0));
            V3<T> s2 = xf_point(M, v3<T>(0, o.end, 0));
            V3<T> rv = v3<T>(o.rad, o.rad, o.rad);
            V3<T> a0 = s1 - rv, a1 = s1 + rv, b0 = s2 - rv, b1 = s2 + rv;
            b.mn = v3<T>(std::min(a0.x, b0.x), std::min(a0.y, b0.y), std::min(a0.z, b0.z));
            b.mx = v3<T>(std::max(a1.x, b1.x), std::max(a1.y, b1.y), std::max(a1.z, b1.z));
            /* MIN/MAX macros: (a)<(b)?(a):(b) — std::min/max agree for non-NaN */
        } else {
            b.mn = v3<T>(FLT_MAX, FLT_MAX, FLT_MAX);
            b.mx = v3<T>(-FLT_MAX, -FLT_MAX, -FLT_MAX);
            const MeshT<T> &m = meshes[o.mesh];
            for (uint32_t f = 0; f < m.count; f++) {
                const Tri<T> &tr = tris[m.start + f];
                for (int k = 0; k < 3; k++) {
                    V3<T> p = xf_point(M, tr.v[k]);
                    if (p.x < b.mn.x) b.mn.x = p.x;
                    if (p.y < b.mn.y) b.mn.y = p.y;
                    if (p.z < b.mn.z) b.mn.z = p.z;
                    if (p.x > b.mx.x) b.mx.x = p.x;
                    if (p.y > b.mx.y) b.mx.y = p.y;
                    if (p.z > b.mx.z) b.mx.z = p.z;
                }
            }
        }
        return b;
    }

    /* Triangle::calc_bounds with the identity matrix, src/mesh.cc:38-54 */
    static Box<T> tri_box(const Tri<T> &tr)
    {
        Box<T> b;
        b.mn = v3<T>(FLT_MAX, FLT_MAX, FLT_MAX);
        b.mx = v3<T>(-FLT_MAX, -FLT_MAX, -FLT_MAX);
        for (int k = 0; k < 3; k++) {
            V3<T> p = tr.v[k];
            if (p.x < b.mn.x) b.mn.x = p.x;
            if (p.y < b.mn.y) b.mn.y = p.y;
            if (p.z < b.mn.z) b.mn.z = p.z;
            if (p.x > b.mx.x) b.mx.x = p.x;
            if (p.y > b.mx.y) b.mx.y = p.y;
            if (p.z > b.mx.z) b.mx.z = p.z;
        }
        return b;
    }

    /* ---------------- primitives ---------------- */

    /* Sphere::intersect, src/sphere.cc:85-132 (t only; attributes later).
     * The quadratic is evaluated in double for both precisions: with
     * |dir| = 1e4, b*b - 4ac cancels about three digits, which in float
     * puts hit points 1e-5 off the surface (self-shadowing the reference,
     * computing in double, does not have). */
    static bool sphere_td(double rad, V3<T> lo, V3<T> ld, double &t)
    {
        double ox = lo.x, oy = lo.y, oz = lo.z, dx = ld.x, dy = ld.y, dz = ld.z;
        double a = dx * dx + dy * dy + dz * dz;
        double b = 2.0 * dx * ox + 2.0 * dy * oy + 2.0 * dz * oz;
        double c = ox * ox + oy * oy + oz * oz - rad * rad;
        double d = b * b - 4.0 * a * c;
        if (d < 0.0) return false;
        double sd = std::sqrt(d);
        double t1 = (-b + sd) / (2.0 * a);
        double t2 = (-b - sd) / (2.0 * a);
        const double e = SRAY_ERROR_MARGIN;
        if ((t1 < e && t2 < e) || (t1 > 1.0 && t2 > 1.0)) return false;
        if (t1 < e) t1 = t2;
        if (t2 < e) t2 = t1;
        t = t1 < t2 ? t1 : t2; /* may exceed 1 (Q4) */
        return true;
    }
    bool sphere_t(const Object<T> &o, const Ray<T> &lr, T &t) const
    {
        double td;
        if (!sphere_td((double)o.rad, lr.o, lr.d, td)) return false;
        t = (T)td;
        return true;
    }

    /* Cylinder::intersect, src/cylinder.cc:125-185 (quadratic in double, as
     * for the sphere) */
    static bool cylinder_td(double rad, double ys, double ye, V3<T> lo, V3<T> ld, double &t)
    {
        double ox = lo.x, oy = lo.y, oz = lo.z, dx = ld.x, dy = ld.y, dz = ld.z;
        double a = dx * dx + dz * dz;
        double b = 2.0 * dx * ox + 2.0 * dz * oz;
        double c = ox * ox + oz * oz - rad * rad;
        double d = b * b - 4.0 * a * c;
        if (d < 0.0) return false;
        double sd = std::sqrt(d);
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
    bool cylinder_t(const Object<T> &o, const Ray<T> &lr, T &t) const
    {
        double td;
        if (!cylinder_td((double)o.rad, (double)o.start, (double)o.end, lr.o, lr.d, td)) return false;
        t = (T)td;
        return true;
    }

    /* Triangle::intersect + check_clock, src/mesh.cc:117-165, 310-315 */
    bool tri_t(const Tri<T> &tr, const Ray<T> &lr, T &t) const
    {
        V3<T> e1 = tr.v[1] - tr.v[0];
        V3<T> e2 = tr.v[2] - tr.v[0];
        V3<T> n = cross(e1, e2);
        T ndd = dot(n, lr.d);
        if (std::fabs(ndd) < eps) return false;
        V3<T> vo = lr.o - tr.v[0];
        T tt = -dot(n, vo) / ndd;
        if (tt < eps || tt > (T)1) return false;
        V3<T> p = lr.o + lr.d * tt;
        for (int i = 0; i < 3; i++) {
            V3<T> a = tr.v[i], b = tr.v[(i + 1) % 3];
            V3<T> tst = cross(b - a, p - a);
            if (!(dot(tst, n) >= (T)0)) return false;
        }
        t = tt;
        return true;
    }

    /* ---------------- scene-level traversal ---------------- */

    /* per-ray memo of mesh results: a mesh in k scene leaves gives the same
     * closest triangle each time (Q9), so it is traversed once per ray
     * unless q9_retraverse asks for the reference's repeated work */
    struct MeshCache {
        struct E {
            int obj = -1;
            bool found = false;
            int tie = 0;
            T t = 0;
            int64_t tri = -1;
        } e[8];
        int next = 0;
        E *find(int oi)
        {
            for (auto &x : e)
                if (x.obj == oi) return &x;
            return nullptr;
        }
        E *slot(int oi)
        {
            E *x = find(oi);
            if (x) return x;
            x = &e[next];
            next = (next + 1) & 7;
            return x;
        }
    };

    /* Mesh::intersect, src/mesh.cc:252-308: closest triangle in local space */
    bool mesh_closest(const Object<T> &o, const Ray<T> &wr, T &t, int64_t &tri, Counters &cnt, int &tie) const
    {
        Mat<T> M, I;
        xform_at(o, wr.time, M, I);
        Ray<T> lr = wr;
        lr.o = xf_point(I, wr.o);
        lr.d = xf_vec(I, wr.d);
        const MeshT<T> &m = meshes[o.mesh];
        Best<T> best;
        if (use_tree) {
            auto fn = [&](uint32_t it, uint32_t rk) {
                T tt;
                cnt.prims++;
                if (tri_t(tris[m.start + it], lr, tt)) best.offer(tt, it, -1, rk);
                return false;
            };
            oct_walk(m.tree, 0, lr, fn, cnt.nodes);
        } else {
            /* brute force, src/mesh.cc:271-287 (sp0.dist starts at DBL_MAX) */
            best.t = (T)DBL_MAX;
            for (uint32_t f = 0; f < m.count; f++) {
                T tt;
                cnt.prims++;
                if (tri_t(tris[m.start + f], lr, tt)) best.offer(tt, f, -1, f);
            }
        }
        tie = best.tie;
        if (!best.found()) return false;
        t = best.t;
        tri = best.item;
        return true;
    }

    bool item_t(int oi, const Ray<T> &wr, T &t, int64_t &sub, Counters &cnt, MeshCache *mc, int &mtie) const
    {
        const Object<T> &o = objs[oi];
        sub = -1;
        mtie = 0;
        if (o.type == SRAY_OBJ_MESH) {
            if (mc && !q9_retraverse) {
                typename MeshCache::E *e = mc->find(oi);
                if (e) {
                    t = e->t;
                    sub = e->tri;
                    mtie = e->tie;
                    return e->found;
                }
            }
            bool f = mesh_closest(o, wr, t, sub, cnt, mtie);
            if (mc) {
                typename MeshCache::E *e = mc->slot(oi);
                e->obj = oi;
                e->found = f;
                e->tie = mtie;
                e->t = t;
                e->tri = sub;
            }
            return f;
        }
        Mat<T> M, I;
        xform_at(o, wr.time, M, I);
        Ray<T> lr = wr;
        lr.o = xf_point(I, wr.o);
        lr.d = xf_vec(I, wr.d);
        cnt.prims++;
        if (o.type == SRAY_OBJ_SPHERE) return sphere_t(o, lr, t);
        return cylinder_t(o, lr, t);
    }

    /* Scene::cast_ray, src/scene.cc:764-794 */
    Best<T> cast(const Ray<T> &r, Counters &cnt, bool any) const
    {
        Best<T> best;
        MeshCache mc;
        if (use_tree) {
            if (tree.nodes.empty()) return best;
            auto fn = [&](uint32_t it, uint32_t rk) {
                T t;
                int64_t sub;
                int mtie = 0;
                if (item_t((int)it, r, t, sub, cnt, &mc, mtie)) {
                    bool won = t < best.t;
                    best.offer(t, it, sub, rk);
                    if (won) best.tie = mtie;
                }
                return any && best.found();
            };
            oct_walk(tree, 0, r, fn, cnt.nodes);
        } else {
            /* brute force with the extra sp.dist > ERROR_MARGIN filter (Q5) */
            best.t = (T)DBL_MAX;
            for (size_t i = 0; i < objs.size(); i++) {
                T t;
                int64_t sub;
                int mtie = 0;
                if (item_t((int)i, r, t, sub, cnt, &mc, mtie) && t > eps) {
                    bool won = t < best.t;
                    best.offer(t, (int64_t)i, sub, (uint32_t)i);
                    if (won) best.tie = mtie;
                    if (any) break;
                }
            }
        }
        return best;
    }

    /* ---------------- hit attributes ---------------- */

    /* interp_face, src/mesh.cc:321-445 (Glassner ch.2 s.3.3, E. Haines):
     * inverse bilinear (u,v) of the degenerate patch p00 p10 p01=p11 */
    void interp_face(const Tri<T> &tr, V3<T> ispos_t, V3<T> normal_t, V3<T> &on, V3<T> &og, T &ou, T &ov) const
    {
        /* evaluated in double for both precisions: for the degenerate quad
         * (v0 v1 v2 v2) du2 is zero up to rounding, and in float that
         * rounding noise passes the XSMALL test and the quadratic branch
         * returns garbage; the reference computes in double */
        typedef double D;
        auto d3 = [](V3<T> a) { return v3<D>((D)a.x, (D)a.y, (D)a.z); };
        V3<D> ispos = d3(ispos_t), normal = d3(normal_t);
        V3<D> p00 = d3(tr.v[0]), p10 = d3(tr.v[1]), p01 = d3(tr.v[2]), p11 = d3(tr.v[2]);
        V3<D> pa = (p00 - p10) + (p11 - p01);
        V3<D> pb = p10 - p00;
        V3<D> pc = p01 - p00;
        V3<D> pd = p00;
        V3<D> na = cross(pa, normal), nb = cross(pb, normal), nc = cross(pc, normal);
        D du0 = dot(nc, pd);
        D du1 = dot(na, pd) + dot(nc, pb);
        D du2 = dot(na, pb);
        D dv0 = dot(nb, pd);
        D dv1 = dot(na, pd) + dot(nb, pc);
        D dv2 = dot(na, pc);
        const D xs = (D)SRAY_XSMALL_NUMBER;
        D ru, rv;
        if (std::fabs(du2) <= xs) {
            D divisor = du1 - dot(na, ispos);
            ru = std::fabs(divisor) > xs ? ((dot(nc, ispos) - du0) / divisor) : (D)0;
        } else {
            V3<D> qux = na / ((D)2 * du2);
            V3<D> quy = (-nc) / du2;
            D dux = -du1 / ((D)2 * du2);
            D duy = du0 / du2;
            D ka = dux + dot(qux, ispos);
            D kb = duy + dot(quy, ispos);
            D u0 = ka - std::sqrt(ka * ka - kb);
            D u1 = ka + std::sqrt(ka * ka - kb);
            ru = (u0 < (D)0 || u0 > (D)1) ? u1 : u0;
        }
        if (std::fabs(dv2) <= xs) {
            D divisor = dv1 - dot(na, ispos);
            rv = std::fabs(divisor) > xs ? ((dot(nb, ispos) - dv0) / divisor) : (D)0;
        } else {
            V3<D> qvx = na / ((D)2 * dv2);
            V3<D> qvy = (-nb) / dv2;
            D dvx = -dv1 / ((D)2 * dv2);
            D dvy = dv0 / dv2;
            D ka = dvx + dot(qvx, ispos);
            D kb = dvy + dot(qvy, ispos);
            D v0 = ka - std::sqrt(ka * ka - kb);
            D v1 = ka + std::sqrt(ka * ka - kb);
            rv = (v0 < (D)0 || v0 > (D)1) ? v1 : v0;
        }
        auto lerp = [](D a, D b, D t) { return a + (b - a) * t; };
        auto l2 = [&](T a00, T a10, T a01) {
            D s0 = lerp((D)a00, (D)a01, rv);
            D s1 = lerp((D)a10, (D)a01, rv);
            return (T)lerp(s0, s1, ru);
        };
        ou = l2(tr.tu[0], tr.tu[1], tr.tu[2]);
        ov = l2(tr.tv[0], tr.tv[1], tr.tv[2]);
        on = v3<T>(l2(tr.n[0].x, tr.n[1].x, tr.n[2].x), l2(tr.n[0].y, tr.n[1].y, tr.n[2].y),
                   l2(tr.n[0].z, tr.n[1].z, tr.n[2].z));
        og = v3<T>(l2(tr.g[0].x, tr.g[1].x, tr.g[2].x), l2(tr.g[0].y, tr.g[1].y, tr.g[2].y),
                   l2(tr.g[0].z, tr.g[1].z, tr.g[2].z));
    }

    /* SurfPoint of the winning item: src/sphere.cc:107-129,
     * src/cylinder.cc:166-182, src/mesh.cc:147-163 + 293-306 */
    SurfPoint<T> surf(const Ray<T> &wr, int oi, int64_t sub, T t) const
    {
        const Object<T> &o = objs[oi];
        Mat<T> M, I;
        xform_at(o, wr.time, M, I);
        V3<T> lo = xf_point(I, wr.o);
        V3<T> ld = xf_vec(I, wr.d);
        SurfPoint<T> sp;
        sp.dist = t;
        V3<T> p = lo + ld * t;
        if (o.type != SRAY_OBJ_MESH) {
            /* the surface point of a quadric from the double-precision root */
            double td;
            bool ok = o.type == SRAY_OBJ_SPHERE
                          ? sphere_td((double)o.rad, lo, ld, td)
                          : cylinder_td((double)o.rad, (double)o.start, (double)o.end, lo, ld, td);
            if (ok)
                p = v3<T>((T)((double)lo.x + (double)ld.x * td), (T)((double)lo.y + (double)ld.y * td),
                          (T)((double)lo.z + (double)ld.z * td));
        }
        if (o.type == SRAY_OBJ_SPHERE) {
            V3<T> nrm = p / o.rad;
            V3<T> tg = cross(v3<T>(0, 1, 0), nrm);
            /* texcoord from SphVector (theta = atan2(z,x) + pi, phi = acos(y/|p|)) */
            T r = length(p);
            T theta = std::atan2(p.z, p.x) + (T)M_PI;
            T phi = r > (T)0 ? std::acos(std::max((T)-1, std::min((T)1, p.y / r))) : (T)0;
            sp.u = theta / (T)(2 * M_PI);
            sp.v = phi / (T)M_PI;
            sp.pos = xf_point(M, p);
            sp.normal = xf_normal(I, nrm);
            sp.tangent = xf_normal(I, tg);
        } else if (o.type == SRAY_OBJ_CYLINDER) {
            V3<T> nrm = v3<T>(p.x, (T)0 * p.y, p.z) / o.rad;
            V3<T> tg = cross(v3<T>(0, 1, 0), nrm);
            T theta = std::atan2(p.z, p.x) + (T)M_PI;
            sp.u = theta / (T)(2 * M_PI);
            sp.v = p.y;
            sp.pos = xf_point(M, p);
            sp.normal = xf_normal(I, nrm);
            sp.tangent = xf_normal(I, tg);
        } else {
            const Tri<T> &tr = tris[meshes[o.mesh].start + (uint32_t)sub];
            V3<T> e1 = tr.v[1] - tr.v[0];
            V3<T> e2 = tr.v[2] - tr.v[0];
            V3<T> gn = cross(e1, e2);
            V3<T> in, ig;
            T iu, iv;
            interp_face(tr, p, gn, in, ig, iu, iv);
            sp.u = iu;
            sp.v = iv;
            V3<T> nrm = dot(in, in) < (T)SRAY_XSMALL_NUMBER ? gn : in;
            sp.pos = xf_point(M, p);
            sp.normal = normalized(xf_normal(I, nrm));
            V3<T> tg = xf_normal(I, ig);
            T tl = length(tg);
            sp.tangent = tl > (T)0 ? tg / tl : tg;
        }
        return sp;
    }

    /* ---------------- photon gather, src/pmap.cc:107-178 ---------------- */
    /* kd_nearest_range3f takes float coordinates and a float radius */
    void gather_list(int kind, V3<T> pos, double radius, int k, std::vector<uint32_t> &out) const
    {
        out.clear();
        const std::vector<Photon> &ph = pm[kind];
        double rf = (double)(float)radius;
        double r2 = rf * rf;
        std::vector<std::pair<double, uint32_t>> cand;
        auto test = [&](uint32_t i) {
            double dx = (double)ph[i].pos[0] - (double)(float)pos.x;
            double dy = (double)ph[i].pos[1] - (double)(float)pos.y;
            double dz = (double)ph[i].pos[2] - (double)(float)pos.z;
            double d2 = dx * dx + dy * dy + dz * dz;
            if (d2 <= r2) cand.push_back({d2, i});
        };
        const PhotonGrid &g = grid[kind];
        if (!brute_gather && g.rf == rf) {
            const double q[3] = {(double)(float)pos.x, (double)(float)pos.y, (double)(float)pos.z};
            double lo[3], hi[3];
            for (int d = 0; d < 3; d++) {
                const double f = std::floor((q[d] - g.o[d]) / g.h);
                lo[d] = std::max(f - 1.0, 0.0);
                hi[d] = std::min(f + 1.0, (double)(g.n[d] - 1));
                if (!(lo[d] <= hi[d])) return; /* no photon within reach */
            }
            for (int64_t z = (int64_t)lo[2]; z <= (int64_t)hi[2]; z++)
                for (int64_t y = (int64_t)lo[1]; y <= (int64_t)hi[1]; y++)
                    for (int64_t x = (int64_t)lo[0]; x <= (int64_t)hi[0]; x++) {
                        const size_t c = (size_t)((z * g.n[1] + y) * g.n[0] + x);
                        for (uint32_t j = g.start[c]; j < g.start[c + 1]; j++) test(g.idx[j]);
                    }
            /* (the candidates are put in ascending index order below, the
             * brute-force loop's order) */
        } else {
            for (uint32_t i = 0; i < ph.size(); i++) test(i);
        }
        if (k > 0 && (int)cand.size() > k) {
            std::sort(cand.begin(), cand.end());
            cand.resize(k);
        }
        std::sort(cand.begin(), cand.end(), [](auto &a, auto &b) { return a.second < b.second; });
        for (auto &c : cand) out.push_back(c.second);
    }
    C4<T> irradiance(V3<T> pos, V3<T> norm, double radius, int k) const
    {
        std::vector<uint32_t> l;
        gather_list(SRAY_PMAP_CAUSTIC, pos, radius, k, l);
        C4<T> irr = c4<T>(0, 0, 0, 1);
        if (l.size() < 8) return c4<T>(0, 0, 0, 0);
        T sx = 0, sy = 0, sz = 0;
        for (uint32_t i : l) {
            const Photon &p = pm[SRAY_PMAP_CAUSTIC][i];
            V3<T> d = v3<T>(p.dir[0], p.dir[1], p.dir[2]);
            if (dot(d, norm) < (T)0) {
                sx += (T)p.col[0];
                sy += (T)p.col[1];
                sz += (T)p.col[2];
            }
        }
        T sc = (T)(1.0 / (M_PI * (radius * radius)));
        irr.x = sx * sc;
        irr.y = sy * sc;
        irr.z = sz * sc;
        irr.w = 0;
        return irr;
    }
    C4<T> radiance(V3<T> pos, V3<T> norm, double radius, int k) const
    {
        static const T max_cos = (T)std::cos(M_PI / 2.5);
        std::vector<uint32_t> l;
        gather_list(SRAY_PMAP_GLOBAL, pos, radius, k, l);
        T sx = 0, sy = 0, sz = 0;
        for (uint32_t i : l) {
            const Photon &p = pm[SRAY_PMAP_GLOBAL][i];
            V3<T> d = v3<T>(p.dir[0], p.dir[1], p.dir[2]);
            V3<T> pn = v3<T>(p.norm[0], p.norm[1], p.norm[2]);
            T ndotl = dot(-d, norm);
            if (ndotl > (T)0 && dot(pn, norm) > max_cos) {
                sx += (T)p.col[0] * ndotl;
                sy += (T)p.col[1] * ndotl;
                sz += (T)p.col[2] * ndotl;
            }
        }
        T sc = (T)(1.0 / (M_PI * (radius * radius)));
        return c4<T>(sx * sc, sy * sc, sz * sc, 0);
    }

    /* ---------------- shading ---------------- */

    struct ShadeCtx {
        const sray_render_params *p;
        uint64_t seed;
        uint32_t pixel, sample;
        uint64_t *ray_counts; /* primary, shadow, reflect, refract, diffuse */
    };

    /* light sample point: Light::get_point src/light.cc:106-109,
     * SphLight src/light.cc:272-275, BoxLight src/light.cc:346-354 (time 0) */
    V3<T> light_point(const Light<T> &lt, const ShadeCtx &sc, uint32_t path, uint32_t li, uint32_t si, int tm) const
    {
        if (lt.type == SRAY_LIGHT_BOX) {
            T fx = (T)rng_unit(sc.seed, sc.pixel, sc.sample, path, slot(K_LIGHT, li, si, 0, 0));
            T fy = (T)rng_unit(sc.seed, sc.pixel, sc.sample, path, slot(K_LIGHT, li, si, 0, 1));
            T fz = (T)rng_unit(sc.seed, sc.pixel, sc.sample, path, slot(K_LIGHT, li, si, 0, 2));
            V3<T> v = v3<T>(fx * lt.dim.x - (T)0.5 * lt.dim.x, fy * lt.dim.y - (T)0.5 * lt.dim.y,
                            fz * lt.dim.z - (T)0.5 * lt.dim.z);
            return xf_point(lt.xf, v);
        }
        if (lt.type == SRAY_LIGHT_SPHERE) {
            V3<T> s = sph_rand(sc, path, K_LIGHT, li, si, lt.radius);
            return light_pos(lt, tm) + s;
        }
        return light_pos(lt, tm);
    }

    /* uniform point on a sphere of radius r (Marsaglia 1972), a bounded
     * rejection loop over counter-based draws */
    V3<T> sph_rand(const ShadeCtx &sc, uint32_t path, uint32_t kind, uint32_t i1, uint32_t i2, T r) const
    {
        for (uint32_t a = 0; a < 64; a++) {
            float u1 = rng_unit(sc.seed, sc.pixel, sc.sample, path, slot(kind, i1, i2, a, 0));
            float u2 = rng_unit(sc.seed, sc.pixel, sc.sample, path, slot(kind, i1, i2, a, 1));
            float x1 = 2.0f * u1 - 1.0f, x2 = 2.0f * u2 - 1.0f;
            float s = x1 * x1 + x2 * x2;
            if (s < 1.0f && s > 0.0f) {
                float q = 2.0f * std::sqrt(1.0f - s);
                V3<T> d = v3<T>((T)(x1 * q), (T)(x2 * q), (T)(1.0f - 2.0f * s));
                return d * r;
            }
        }
        return v3<T>(0, 0, r);
    }

    /* calc_lighting, src/shader.cc:250-283 */
    void calc_lighting(C4<T> &diff, C4<T> &spec, const Light<T> &lt, uint32_t li, T shininess, V3<T> pt,
                       V3<T> norm, V3<T> vdir, int tm, const ShadeCtx &sc, uint32_t path) const
    {
        int ns = lt.type != SRAY_LIGHT_POINT ? sc.p->shadow_samples : 1;
        diff = c4<T>(0, 0, 0, 0);
        spec = c4<T>(0, 0, 0, 0);
        for (int i = 0; i < ns; i++) {
            V3<T> lpos = light_point(lt, sc, path, li, (uint32_t)i, tm);
            V3<T> ldir = lpos - pt;
            T ldist = length(ldir);
            Ray<T> sr;
            sr.o = pt;
            sr.d = ldir;
            sr.time = tm;
            ldir = ldir / ldist;
            Counters cnt;
            sc.ray_counts[1]++;
            if (!cast(sr, cnt, true).found()) {
                V3<T> vref = reflect(vdir, norm);
                T ndotl = std::max(dot(norm, ldir), (T)0);
                T rdotl = std::max(dot(vref, ldir), (T)0);
                C4<T> lcol = lt.color; /* attenuation is always 1 (Q12) */
                diff = diff + lcol * ndotl;
                spec = spec + lcol * (T)std::pow(rdotl, shininess);
            }
        }
        if (ns > 1) {
            T s = (T)1 / (T)ns;
            diff = diff * s;
            spec = spec * s;
        }
    }

    /* Texture::get_texel, src/texture.cc:85-101: repeat wraps with C's
     * remainder (negatives moved up by the size); clamp holds the index to
     * [0, size] — the size itself included, one texel past the row — and the
     * row-major index is kept, where one past the last texel reads the last */
    static C4<T> texel(const Texture &tx, int x, int y)
    {
        if (tx.wrap == SRAY_TEX_WRAP_CLAMP) {
            x = x < 0 ? 0 : (x > tx.w ? tx.w : x);
            y = y < 0 ? 0 : (y > tx.h ? tx.h : y);
        } else {
            x %= tx.w;
            y %= tx.h;
            if (x < 0) x += tx.w;
            if (y < 0) y += tx.h;
        }
        long long k = (long long)y * tx.w + x;
        const long long last = (long long)tx.w * tx.h - 1;
        if (k > last) k = last;
        const float *p = &tx.px[(size_t)k * 4];
        return c4<T>(p[0], p[1], p[2], p[3]);
    }
    /* (int)(c * size), C truncation, held to +-2^30 texels */
    static int tex_coord(T c, int size)
    {
        T f = c * (T)size;
        f = std::max(std::min(f, (T)1073741824.0), (T)-1073741824.0);
        return (int)f;
    }
    /* Texture::lookup, src/texture.cc:60-83: (u, v, 0, 1) through the
     * texture matrix, then one texel or the 5-texel cross averaged */
    static C4<T> tex_lookup(const Texture &tx, T u, T v)
    {
        const T x = (T)tx.xf[0] * u + (T)tx.xf[1] * v + (T)tx.xf[2] * (T)0 + (T)tx.xf[3];
        const T y = (T)tx.xf[4] * u + (T)tx.xf[5] * v + (T)tx.xf[6] * (T)0 + (T)tx.xf[7];
        const int ix = tex_coord(x, tx.w), iy = tex_coord(y, tx.h);
        if (tx.filter == SRAY_TEX_FILTER_NEAREST) return texel(tx, ix, iy);
        C4<T> a = texel(tx, ix, iy), b = texel(tx, ix + 1, iy), c = texel(tx, ix - 1, iy), d = texel(tx, ix, iy + 1),
              e = texel(tx, ix, iy - 1);
        C4<T> r;
        r.x = ((((a.x + b.x) + c.x) + d.x) + e.x) / (T)5;
        r.y = ((((a.y + b.y) + c.y) + d.y) + e.y) / (T)5;
        r.z = ((((a.z + b.z) + c.z) + d.z) + e.z) / (T)5;
        r.w = ((((a.w + b.w) + c.w) + d.w) + e.w) / (T)5;
        return r;
    }
    /* MatAttrib::get_color / get_value, src/material.inl:18-28 */
    C4<T> attr_color(int map, const C4<T> &col, const SurfPoint<T> &sp) const
    {
        if (map <= 0) return col;
        C4<T> t = tex_lookup(texs[map - 1], sp.u, sp.v);
        return c4<T>(t.x * col.x, t.y * col.y, t.z * col.z, t.w * col.w);
    }
    T attr_value(int map, T val, const SurfPoint<T> &sp) const
    {
        if (map <= 0) return val;
        return tex_lookup(texs[map - 1], sp.u, sp.v).x * val;
    }
    /* get_bump_normal, src/shader.cc:293-316: without a map the unflipped
     * surface normal; with one, the map's colour * 2 - 1 (x, y scaled by
     * the attribute's value), normalised, in the frame (tangent,
     * normal x tangent, normal) */
    V3<T> bump_normal(const Material<T> &m, const SurfPoint<T> &sp) const
    {
        if (m.map[SRAY_MAP_NORMAL] <= 0) return sp.normal;
        C4<T> c = tex_lookup(texs[m.map[SRAY_MAP_NORMAL] - 1], sp.u, sp.v);
        V3<T> n = v3<T>(c.x * (T)2 - (T)1, c.y * (T)2 - (T)1, c.z * (T)2 - (T)1);
        n.x = n.x * m.nscale;
        n.y = n.y * m.nscale;
        n = normalized(n);
        V3<T> b = cross(sp.normal, sp.tangent);
        return v3<T>(sp.tangent.x * n.x + b.x * n.y + sp.normal.x * n.z, sp.tangent.y * n.x + b.y * n.y + sp.normal.y * n.z,
                     sp.tangent.z * n.x + b.z * n.y + sp.normal.z * n.z);
    }

    static T fresnel(T r, T cosa)
    {
        /* src/shader.cc:285-291 */
        T inv = (T)1 - cosa;
        T inv_sq = inv * inv;
        return r + ((T)1 - r) * inv_sq * inv_sq * inv;
    }

    /* shade_undef, src/shader.cc:36-71 (checker from texcoords) */
    C4<T> shade_undef(const Ray<T> &ray, const SurfPoint<T> &sp, const ShadeCtx &sc, uint32_t path) const
    {
        int x = (int)(sp.u * (T)20);
        int y = (int)(sp.v * (T)10);
        C4<T> tex = ((x % 2) == (y % 2)) ? c4<T>((T)1, (T)0.3, (T)0.2, 1) : c4<T>((T)0.2, (T)0.3, (T)1, 1);
        C4<T> diff = c4<T>((T)0.1, (T)0.1, (T)0.1, 1), spec = c4<T>(0, 0, 0, 1);
        for (size_t i = 0; i < lights.size(); i++) {
            const Light<T> &lt = lights[i];
            V3<T> lpos = light_point(lt, sc, path, (uint32_t)i, 0, ray.time);
            V3<T> ldir = lpos - sp.pos;
            T ldist = length(ldir);
            Ray<T> sr;
            sr.o = sp.pos;
            sr.d = ldir;
            sr.time = ray.time;
            ldir = ldir / ldist;
            V3<T> vdir = -normalized(ray.d);
            V3<T> hdir = normalized(ldir + vdir);
            Counters cnt;
            sc.ray_counts[1]++;
            if (!cast(sr, cnt, true).found()) {
                T ndotl = std::max(dot(sp.normal, ldir), (T)0);
                T ndoth = std::max(dot(sp.normal, hdir), (T)0);
                diff = diff + lt.color * ndotl;
                spec = spec + lt.color * (T)std::pow(ndoth, (T)50);
            }
        }
        return diff * tex + spec;
    }

    /* shade_phong, src/shader.cc:83-247 */
    C4<T> shade_phong(const Ray<T> &ray, const SurfPoint<T> &sp, const Material<T> &mat, const ShadeCtx &sc,
                      uint32_t path) const
    {
        bool entering;
        V3<T> normal;
        if (dot(ray.d, sp.normal) > (T)0) {
            normal = -sp.normal;
            entering = false;
        } else {
            normal = sp.normal;
            entering = true;
        }
        T ray_mag = length(ray.d);
        V3<T> incident = ray.d / ray_mag;

        C4<T> ks = c4<T>(0, 0, 0, 1);
        T refl_fact = 0, trans_fact = 0;
        C4<T> kd = attr_color(mat.map[SRAY_MAP_DIFFUSE], mat.kd, sp);
        if (mat.flags & SRAY_MAT_HAS_SPECULAR) ks = attr_color(mat.map[SRAY_MAP_SPECULAR], mat.ks, sp);
        T shininess = attr_value(mat.map[SRAY_MAP_SHININESS], mat.shininess, sp);

        if (ray.iter == INT_MAX) {
            /* diffuse (GI) ray: radiance estimate from the global map */
            C4<T> rad = radiance(sp.pos, normal, gather_dist_abs, std::max(sc.p->gather_k, 0));
            return rad * kd;
        }
        if (mat.flags & SRAY_MAT_HAS_REFLECT) refl_fact = attr_value(mat.map[SRAY_MAP_REFLECT], mat.reflect, sp);
        if (mat.flags & SRAY_MAT_HAS_REFRACT) trans_fact = attr_value(mat.map[SRAY_MAP_REFRACT], mat.refract, sp);
        T mat_ior = attr_value(mat.map[SRAY_MAP_IOR], mat.ior, sp);
        T ior = ray.calc_ior(entering, mat_ior);
        if (mat.flags & SRAY_MAT_HAS_NORMAL) normal = bump_normal(mat, sp);

        C4<T> diff = env_ambient * kd;
        C4<T> spec = c4<T>(0, 0, 0, 1);
        for (size_t i = 0; i < lights.size(); i++) {
            C4<T> d, s;
            calc_lighting(d, s, lights[i], (uint32_t)i, shininess, sp.pos, normal, incident, ray.time, sc, path);
            diff = diff + d;
            spec = spec + s;
        }

        C4<T> gi = c4<T>(0, 0, 0, 0);
        if (sc.p->gi_photons) {
            for (int i = 0; i < sc.p->diffuse_samples; i++) {
                V3<T> dir = sph_rand(sc, path, K_DIFFUSE, (uint32_t)i, 0, (T)1);
                T ndotl = dot(dir, normal);
                if (ndotl < (T)0) dir = -dir;
                T diff_energy = ndotl * ray.energy;
                if (diff_energy > (T)sc.p->min_energy) {
                    Ray<T> dr = ray;
                    dr.iter = INT_MAX;
                    dr.energy = diff_energy;
                    dr.o = sp.pos;
                    dr.d = dir * ray_mag;
                    sc.ray_counts[4]++;
                    gi = gi + trace(dr, sc, path | 0x40000000u | ((uint32_t)i << 20)) * ndotl;
                }
            }
            if (sc.p->diffuse_samples > 1) gi = gi * ((T)1 / (T)sc.p->diffuse_samples);
        }

        C4<T> refl = c4<T>(0, 0, 0, 0), refr = c4<T>(0, 0, 0, 0);
        T refl_energy = refl_fact * ray.energy;
        if (refl_energy > (T)sc.p->min_energy && ray.iter > 0) {
            Ray<T> rr = ray;
            rr.energy = refl_energy;
            rr.iter--;
            rr.o = sp.pos;
            rr.d = reflect(rr.d, normal);
            sc.ray_counts[2]++;
            refl = trace(rr, sc, path * 2) * refl_fact;
        }
        T trans_energy = trans_fact * ray.energy;
        if (trans_energy > (T)sc.p->min_energy && ray.iter > 0) {
            Ray<T> tr = ray;
            if (entering) tr.enter(mat_ior);
            else tr.leave();
            tr.energy = trans_energy;
            tr.iter--;
            tr.o = sp.pos;
            tr.d = refract(ray.d / ray_mag, normal, ior) * ray_mag;
            if (dot(tr.d, normal) > (T)0) {
                /* total internal reflection: undo the medium change */
                if (entering) tr.leave();
                else tr.enter(mat_ior);
            }
            sc.ray_counts[3]++;
            refr = trace(tr, sc, path * 2 + 1) * trans_fact;
        }

        C4<T> irrad = c4<T>(0, 0, 0, 0);
        if (sc.p->caust_photons) irrad = irradiance(sp.pos, normal, gather_dist_abs, std::max(sc.p->gather_k, 0));

        T ray_dot_n = dot(-ray.d / ray_mag, normal);
        T sqrt_fres_0 = trans_fact > (T)0 ? (ior - (T)1) / (ior + (T)1) : (T)1;
        T fres = fresnel(sqrt_fres_0 * sqrt_fres_0, ray_dot_n);
        C4<T> reflrefr = refr + (refl - refr) * fres;
        return (diff + gi) * kd + (spec + reflrefr) * ks + irrad;
    }

    /* Scene::trace_ray, src/scene.cc:551-569 */
    C4<T> trace(const Ray<T> &ray, const ShadeCtx &sc, uint32_t path) const
    {
        Counters cnt;
        Best<T> b = cast(ray, cnt, false);
        if (b.found()) {
            SurfPoint<T> sp = surf(ray, (int)b.item, b.sub, b.t);
            const Object<T> &o = objs[b.item];
            C4<T> col;
            if (o.mat >= 0 && (mats[o.mat].flags & SRAY_MAT_PHONG)) col = shade_phong(ray, sp, mats[o.mat], sc, path);
            else col = shade_undef(ray, sp, sc, path);
            col.w = 1;
            return col;
        }
        return env_color;
    }

    /* Camera::get_primary_ray, src/camera.cc:119-142 */
    Ray<T> primary(int x, int y, int sub, int time, const sray_render_params *p, uint64_t seed) const
    {
        T ysz = 2;
        /* Camera(): aspect = opt.width / opt.height in double (src/camera.cc:31) */
        T aspect = (T)((double)p->width / (double)p->height);
        T xsz = aspect * ysz;
        T px = ((T)x / (T)p->width) * xsz - xsz / (T)2;
        T py = (T)1 - ((T)y / (T)p->height) * ysz;
        uint32_t pix = (uint32_t)y * (uint32_t)p->width + (uint32_t)x;
        if (sub > 0) {
            T f0 = (T)rng_unit(seed, pix, (uint32_t)sub, 0, slot(K_CAMERA, 0, 0, 0, 0));
            T f1 = (T)rng_unit(seed, pix, (uint32_t)sub, 0, slot(K_CAMERA, 0, 0, 0, 1));
            px = px + xsz * (f0 - (T)0.5) / (T)p->width;
            py = py + ysz * (f1 - (T)0.5) / (T)p->height;
        }
        Ray<T> r;
        r.o = v3<T>(0, 0, 0);
        /* 1/tan(vfov/2) is evaluated in double for both precisions, so the
         * float path does not depend on a float tan() implementation */
        T fz = (T)(1.0 / std::tan(0.5 * vfov));
        r.d = v3<T>(px * max_dist, py * max_dist, fz * max_dist);
        r.time = time;
        if (p->mblur && sub > 0 && shutter > 0) {
            uint32_t u = rng_u32(seed, pix, (uint32_t)sub, 0, slot(K_CAMERA, 0, 0, 0, 2));
            r.time += (int)(u % (uint32_t)shutter) - shutter / 2;
        }
        /* ray.transformed(get_matrix(ray.time)) */
        const Mat<T> cm = cam_rows.empty() ? cam_xf : mat_from<T>(&cam_rows[row_of(r.time) * 12]);
        Ray<T> w = r;
        w.o = xf_point(cm, r.o);
        w.d = xf_vec(cm, r.d);
        return w;
    }
};

/* ------------------------------------------------------------------ */
/* photon shooting, src/render.cc:67-122, src/scene.cc:445-760,          */
/* src/light.cc:111-205, 392-432, src/projmap.cc                         */
/* ------------------------------------------------------------------ */
enum : uint32_t { P_TIME = 0, P_LPOINT = 1, P_EMIT = 2, P_RR = 3, P_DIFF = 4, P_PROJ = 5 };
constexpr uint32_t PHOTON_ATTEMPTS = 1u << 17;
constexpr int PHOTON_MAX_BOUNCES = 1024;

/* photon draws: a = emission index, b = kind<<24 | light, c = bounce */
static inline float prand(uint64_t seed, uint32_t a, uint32_t b, uint32_t c, uint32_t purpose, uint32_t att,
                          uint32_t dim)
{
    return rng_unit(seed, a, b, c, slot(K_PHOTON, purpose, att >> 7, att & 127u, dim));
}

/* one Marsaglia attempt; false if rejected */
static inline bool marsaglia(float u1, float u2, float out[3])
{
    float x1 = 2.0f * u1 - 1.0f, x2 = 2.0f * u2 - 1.0f;
    float s = x1 * x1 + x2 * x2;
    if (!(s < 1.0f && s > 0.0f)) return false;
    float q = 2.0f * std::sqrt(1.0f - s);
    out[0] = x1 * q;
    out[1] = x2 * q;
    out[2] = 1.0f - 2.0f * s;
    return true;
}

/* ProjMap::cell_from_dir without atan2/acos: boundary cosines in float */
struct CellTables {
    float cv[SRAY_PROJ_V + 1];  /* cos(k*pi/32) */
    float cu[SRAY_PROJ_U + 1];  /* cos(2*pi*k/64) */
    CellTables()
    {
        for (int k = 0; k <= SRAY_PROJ_V; k++) cv[k] = (float)std::cos((double)k * M_PI / SRAY_PROJ_V);
        for (int k = 0; k <= SRAY_PROJ_U; k++) cu[k] = (float)std::cos(2.0 * M_PI * (double)k / SRAY_PROJ_U);
    }
};
static const CellTables kCell;

static inline void cell_from_dir(float x, float y, float z, int &uc, int &vc)
{
    float r = std::sqrt(x * x + y * y + z * z);
    float cy = y / r;
    int v = 0;
    for (int k = 1; k < SRAY_PROJ_V; k++) v += (cy <= kCell.cv[k]) ? 1 : 0;
    float rho = std::sqrt(x * x + z * z);
    int u;
    if (!(rho > 0.0f)) {
        u = x < 0.0f ? SRAY_PROJ_U / 2 : 0;
    } else {
        float cx = x / rho;
        if (!(z < 0.0f)) {
            u = 0;
            for (int k = 1; k <= SRAY_PROJ_U / 2; k++) u += (cx <= kCell.cu[k]) ? 1 : 0;
        } else {
            u = SRAY_PROJ_U / 2;
            for (int k = SRAY_PROJ_U / 2 + 1; k < SRAY_PROJ_U; k++) u += (cx >= kCell.cu[k]) ? 1 : 0;
        }
    }
    uc = std::min(u, SRAY_PROJ_U - 1);
    vc = std::min(v, SRAY_PROJ_V - 1);
}

/* 64x32 map of directions from a light towards objects */
struct ProjMapT {
    std::vector<uint8_t> cells;
    std::vector<double> area;
    double total = 0, coverage = 0;
    ProjMapT() : cells(SRAY_PROJ_U * SRAY_PROJ_V, 0), area(SRAY_PROJ_U * SRAY_PROJ_V)
    {
        /* ProjMap::set_subdiv, src/projmap.cc:44-74 */
        for (int i = 0; i < SRAY_PROJ_U * SRAY_PROJ_V; i++) {
            int uc = i % SRAY_PROJ_U, vc = i / SRAY_PROJ_U;
            double v[4][3];
            for (int k = 0; k < 4; k++) corner(uc, vc, k, v[k]);
            auto tri = [&](const double *a, const double *b, const double *c) {
                double e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]}, e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
                double cx = e1[1] * e2[2] - e1[2] * e2[1], cy = e1[2] * e2[0] - e1[0] * e2[2],
                       cz = e1[0] * e2[1] - e1[1] * e2[0];
                return std::sqrt(cx * cx + cy * cy + cz * cz) * 0.5;
            };
            area[i] = tri(v[0], v[1], v[2]) + tri(v[0], v[2], v[3]);
            total += area[i];
        }
    }
    /* ProjMap::cell_corner_dir, src/projmap.cc:129-144 */
    static void corner(int uc, int vc, int k, double out[3])
    {
        static const int offs[4][2] = {{0, 0}, {1, 0}, {1, 1}, {0, 1}};
        double u = (double)(uc + offs[k][0]) * (1.0 / SRAY_PROJ_U);
        double v = (double)(vc + offs[k][1]) * (1.0 / SRAY_PROJ_V);
        double th = u * M_PI * 2.0, ph = v * M_PI;
        out[0] = std::sin(ph) * std::cos(th);
        out[1] = std::cos(ph);
        out[2] = std::sin(ph) * std::sin(th);
    }
    void set(int i, bool val)
    {
        if ((cells[i] != 0) != val) coverage += val ? area[i] / total : -area[i] / total;
        cells[i] = val ? 1 : 0;
    }
    void set_all() { std::fill(cells.begin(), cells.end(), 1), coverage = 1.0; }
    double get_coverage() const { return std::min(std::max(coverage, 0.0), 1.0); }
};

/* ::build_projmap, src/light.cc:392-432 */
static void build_projmap(ProjMapT &map, const double lpos[3], const std::vector<std::array<double, 6>> &boxes)
{
    for (const auto &b : boxes) {
        if (lpos[0] >= b[0] && lpos[1] >= b[1] && lpos[2] >= b[2] && lpos[0] < b[3] && lpos[1] < b[4] && lpos[2] < b[5]) {
            map.set_all();
            return;
        }
        double c[3] = {(b[0] + b[3]) / 2.0, (b[1] + b[4]) / 2.0, (b[2] + b[5]) / 2.0};
        double r2 = (b[0] - c[0]) * (b[0] - c[0]) + (b[1] - c[1]) * (b[1] - c[1]) + (b[2] - c[2]) * (b[2] - c[2]);
        double v[3] = {c[0] - lpos[0], c[1] - lpos[1], c[2] - lpos[2]};
        double dist = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        int uc, vc;
        cell_from_dir((float)v[0], (float)v[1], (float)v[2], uc, vc);
        map.set(vc * SRAY_PROJ_U + uc, true);
        for (int j = 0; j < SRAY_PROJ_U * SRAY_PROJ_V; j++) {
            int u = j % SRAY_PROJ_U, w = j / SRAY_PROJ_U;
            for (int k = 0; k < 4; k++) {
                double d[3];
                ProjMapT::corner(u, w, k, d);
                double p[3] = {lpos[0] + d[0] * dist - c[0], lpos[1] + d[1] * dist - c[1], lpos[2] + d[2] * dist - c[2]};
                if (p[0] * p[0] + p[1] * p[1] + p[2] * p[2] <= r2) {
                    map.set(j, true);
                    break;
                }
            }
        }
    }
}

template <class T> struct PhotonTracer {
    Scene<T> &S;
    const sray_photon_params *pp;
    uint64_t segments = 0;
    std::vector<ProjMapT> proj, spec;

    PhotonTracer(Scene<T> &s, const sray_photon_params *p) : S(s), pp(p) {}

    /* light point: point = position, sphere = pos + sphrand(r), box = box
     * point at time 0 (src/light.cc:106-109, 272-275, 346-354) */
    bool light_point(const Light<T> &lt, uint64_t seed, uint32_t a, uint32_t b, uint32_t c, uint32_t purpose, int msec,
                     double out[3]) const
    {
        const V3<T> lpos = S.light_pos(lt, msec);
        if (lt.type == SRAY_LIGHT_BOX) {
            T f[3];
            for (int k = 0; k < 3; k++) f[k] = (T)prand(seed, a, b, c, purpose, 0, (uint32_t)k);
            V3<T> v = v3<T>(f[0] * lt.dim.x - (T)0.5 * lt.dim.x, f[1] * lt.dim.y - (T)0.5 * lt.dim.y,
                            f[2] * lt.dim.z - (T)0.5 * lt.dim.z);
            V3<T> w = xf_point(lt.xf, v);
            out[0] = (double)w.x, out[1] = (double)w.y, out[2] = (double)w.z;
            return true;
        }
        if (lt.type == SRAY_LIGHT_SPHERE) {
            for (uint32_t at = 0; at < 128; at++) {
                float d[3];
                if (marsaglia(prand(seed, a, b, c, purpose, at, 0), prand(seed, a, b, c, purpose, at, 1), d)) {
                    V3<T> w = lpos + v3<T>((T)d[0], (T)d[1], (T)d[2]) * lt.radius;
                    out[0] = (double)w.x, out[1] = (double)w.y, out[2] = (double)w.z;
                    return true;
                }
            }
        }
        out[0] = (double)lpos.x, out[1] = (double)lpos.y, out[2] = (double)lpos.z;
        return true;
    }

    /* Light::build_proj_intern, src/light.cc:151-205 */
    void build_maps(int t0, int t1)
    {
        proj.assign(S.lights.size(), ProjMapT());
        spec.assign(S.lights.size(), ProjMapT());
        int tm[2] = {t0, t1};
        int ntm = t0 == t1 ? 1 : 2;
        for (size_t li = 0; li < S.lights.size(); li++) {
            const Light<T> &lt = S.lights[li];
            int samples = lt.type == SRAY_LIGHT_SPHERE ? 8 : 1;
            for (int ti = 0; ti < ntm; ti++) {
                std::vector<std::array<double, 6>> all, caust;
                for (size_t i = 0; i < S.objs.size(); i++) {
                    Box<T> b = S.bounds(S.objs[i], tm[ti]);
                    std::array<double, 6> bb = {(double)b.mn.x, (double)b.mn.y, (double)b.mn.z,
                                                (double)b.mx.x, (double)b.mx.y, (double)b.mx.z};
                    all.push_back(bb);
                    int m = S.objs[i].mat;
                    /* the attributes' colours, without their maps (src/light.cc:173-176) */
                    bool specobj = false;
                    if (m >= 0) {
                        const Material<T> &mm = S.mats[m];
                        T rl = (mm.flags & SRAY_MAT_HAS_REFLECT) ? mm.reflect : (T)1;
                        T rr = (mm.flags & SRAY_MAT_HAS_REFRACT) ? mm.refract : (T)1;
                        specobj = rl > (T)0 || rr > (T)0;
                    }
                    if (specobj) caust.push_back(bb);
                }
                for (int sm = 0; sm < samples; sm++) {
                    double lp[3];
                    light_point(lt, pp->seed, 0xffffffffu, (uint32_t)li, (uint32_t)(ti << 8 | sm), P_PROJ, tm[ti], lp);
                    build_projmap(proj[li], lp, all);
                    build_projmap(spec[li], lp, caust);
                }
            }
        }
    }

    struct Ph {
        int type; /* 0 direct, 1 caustic, 2 global */
        V3<T> pos, dir, norm;
        T col[3];
    };

    /* material-driven interaction shared by both tracers */
    struct Hit {
        SurfPoint<T> sp;
        const Material<T> *mat;
        V3<T> normal;
        bool entering;
        T ray_mag;
    };
    bool cast_hit(const Ray<T> &r, Hit &h)
    {
        Counters cnt;
        Best<T> b = S.cast(r, cnt, false);
        segments++;
        if (!b.found()) return false;
        h.sp = S.surf(r, (int)b.item, b.sub, b.t);
        int m = S.objs[b.item].mat;
        static Material<T> defm = [] {
            Material<T> d;
            d.kd = c4<T>(1, 1, 1, 1);
            d.ks = c4<T>(0, 0, 0, 1);
            d.shininess = 1;
            d.reflect = 0;
            d.refract = 0;
            d.ior = (T)1.5;
            d.flags = SRAY_MAT_HAS_SPECULAR | SRAY_MAT_HAS_REFLECT | SRAY_MAT_HAS_REFRACT;
            return d;
        }();
        h.mat = m >= 0 ? &S.mats[m] : &defm;
        h.ray_mag = length(r.d);
        if (dot(r.d, h.sp.normal) > (T)0) {
            h.normal = -h.sp.normal;
            h.entering = false;
        } else {
            h.normal = h.sp.normal;
            h.entering = true;
        }
        return true;
    }

    /* Material::get_color / get_value return the default attribute
     * (1,1,1,1) for a missing attribute (src/material.cc:77,164-168) */
    C4<T> attr_specular(const Material<T> &m, const SurfPoint<T> &sp) const
    {
        return (m.flags & SRAY_MAT_HAS_SPECULAR) ? S.attr_color(m.map[SRAY_MAP_SPECULAR], m.ks, sp) : c4<T>(1, 1, 1, 1);
    }
    T attr_reflect(const Material<T> &m, const SurfPoint<T> &sp) const
    {
        return (m.flags & SRAY_MAT_HAS_REFLECT) ? S.attr_value(m.map[SRAY_MAP_REFLECT], m.reflect, sp) : (T)1;
    }
    T attr_refract(const Material<T> &m, const SurfPoint<T> &sp) const
    {
        return (m.flags & SRAY_MAT_HAS_REFRACT) ? S.attr_value(m.map[SRAY_MAP_REFRACT], m.refract, sp) : (T)1;
    }

    /* reflect_ray / refract_ray (libvmath, fixed here as in shade_phong) */
    static Ray<T> reflect_ray(const Ray<T> &r, V3<T> n)
    {
        Ray<T> o = r;
        o.d = reflect(r.d, n);
        return o;
    }
    static Ray<T> refract_ray(const Ray<T> &r, V3<T> n, T mat_ior, bool entering, T ray_mag)
    {
        T ior = r.calc_ior(entering, mat_ior);
        Ray<T> o = r;
        if (entering) o.enter(mat_ior);
        else o.leave();
        o.d = refract(r.d / ray_mag, n, ior) * ray_mag;
        if (dot(o.d, n) > (T)0) {
            if (entering) o.leave();
            else o.enter(mat_ior);
        }
        return o;
    }

    /* Scene::trace_caustics_photon, src/scene.cc:571-647 (iterative) */
    bool trace_caustics(Ray<T> r, Ph &ph, uint32_t a, uint32_t b, uint32_t &bounce)
    {
        for (int guard = 0; guard < PHOTON_MAX_BOUNCES; guard++) {
            if (r.iter <= 0) return false;
            Hit h;
            if (!cast_hit(r, h)) return false;
            const Material<T> &m = *h.mat;
            T refl = attr_reflect(m, h.sp), refr = attr_refract(m, h.sp), mat_ior = S.attr_value(m.map[SRAY_MAP_IOR], m.ior, h.sp);
            C4<T> spc = attr_specular(m, h.sp);
            T ior = r.calc_ior(h.entering, mat_ior);
            if (m.flags & SRAY_MAT_HAS_NORMAL) h.normal = S.bump_normal(m, h.sp); /* src/scene.cc:603-606 */
            T ray_dot_n = dot(-r.d / h.ray_mag, h.normal);
            T sf0 = refr > (T)0 ? (ior - (T)1) / (ior + (T)1) : (T)1;
            T fres = Scene<T>::fresnel(sf0 * sf0, ray_dot_n);
            refl = fres * refl;
            refr = ((T)1 - fres) * refr;
            T rnum = (T)prand(pp->seed, a, b, bounce, P_RR, 0, 0);
            bounce++;
            if (rnum < refl) {
                Ray<T> nr = reflect_ray(r, h.normal);
                nr.iter--;
                nr.o = h.sp.pos;
                ph.type = 1;
                for (int k = 0; k < 3; k++) ph.col[k] *= (&spc.x)[k];
                r = nr;
            } else if (rnum < refr) {
                Ray<T> nr = refract_ray(r, h.normal, mat_ior, h.entering, h.ray_mag);
                nr.o = h.sp.pos;
                ph.type = 1;
                for (int k = 0; k < 3; k++) ph.col[k] *= (&spc.x)[k];
                r = nr;
            } else {
                ph.dir = normalized(r.d);
                ph.pos = h.sp.pos;
                ph.norm = h.normal;
                return true;
            }
        }
        return false;
    }

    /* Scene::trace_global_photon, src/scene.cc:650-760 (iterative) */
    bool trace_global(Ray<T> r, Ph &ph, uint32_t a, uint32_t b, uint32_t &bounce)
    {
        for (int guard = 0; guard < PHOTON_MAX_BOUNCES; guard++) {
            if (r.iter <= 0) return false;
            V3<T> incident = normalized(r.d);
            Hit h;
            if (!cast_hit(r, h)) return false;
            const Material<T> &m = *h.mat;
            C4<T> diff = S.attr_color(m.map[SRAY_MAP_DIFFUSE], m.kd, h.sp), spc = attr_specular(m, h.sp);
            T diff_avg = (diff.x + diff.y + diff.z) / (T)3;
            T spec_avg = (spc.x + spc.y + spc.z) / (T)3;
            T rnum = (T)prand(pp->seed, a, b, bounce, P_RR, 0, 0);
            uint32_t here = bounce++;
            if (rnum < diff_avg) {
                /* diffuse reflection: uniform direction flipped into the hemisphere */
                float d[3] = {0, 0, 1};
                for (uint32_t at = 0; at < 128; at++)
                    if (marsaglia(prand(pp->seed, a, b, here, P_DIFF, at, 0), prand(pp->seed, a, b, here, P_DIFF, at, 1), d))
                        break;
                V3<T> dir = v3<T>((T)d[0], (T)d[1], (T)d[2]);
                if (dot(dir, h.normal) < (T)0) dir = -dir;
                T ndotl = dot(-incident, h.normal);
                Ray<T> nr = r;
                nr.d = dir * (T)SRAY_RAY_MAGNITUDE;
                nr.iter--;
                nr.o = h.sp.pos;
                ph.type = 2;
                ph.col[0] *= ndotl * (diff.x / diff_avg);
                ph.col[1] *= ndotl * (diff.y / diff_avg);
                ph.col[2] *= ndotl * (diff.z / diff_avg);
                r = nr;
            } else if (rnum < diff_avg + spec_avg) {
                T refl = attr_reflect(m, h.sp), refr = attr_refract(m, h.sp),
                  mat_ior = S.attr_value(m.map[SRAY_MAP_IOR], m.ior, h.sp);
                T ior = r.calc_ior(h.entering, mat_ior);
                T ray_dot_n = dot(-r.d / h.ray_mag, h.normal);
                T sf0 = refr > (T)0 ? (ior - (T)1) / (ior + (T)1) : (T)1;
                T fres = Scene<T>::fresnel(sf0 * sf0, ray_dot_n);
                refl = fres * refl;
                refr = ((T)1 - fres) * refr;
                T rn2 = (T)prand(pp->seed, a, b, here, P_RR, 0, 1);
                if (rn2 < refl) {
                    Ray<T> nr = reflect_ray(r, h.normal);
                    nr.iter--;
                    nr.o = h.sp.pos;
                    ph.type = 1;
                    for (int k = 0; k < 3; k++) ph.col[k] *= (&spc.x)[k];
                    return trace_caustics(nr, ph, a, b, bounce);
                } else if (rn2 < refr) {
                    Ray<T> nr = refract_ray(r, h.normal, mat_ior, h.entering, h.ray_mag);
                    nr.o = h.sp.pos;
                    ph.type = 1;
                    for (int k = 0; k < 3; k++) ph.col[k] *= (&spc.x)[k];
                    return trace_caustics(nr, ph, a, b, bounce);
                }
                return false; /* specular BRDF lobe: not implemented in the reference either */
            } else {
                ph.dir = incident;
                ph.pos = h.sp.pos;
                ph.norm = h.normal;
                return true;
            }
        }
        return false;
    }

    /* Light::gen_photon, src/light.cc:111-148, with the direction drawn by
     * rejection towards full projection-map cells (src/projmap.cc:194-208) */
    Ph gen_photon(size_t li, int kind, uint32_t a, uint32_t b, int msec, Ray<T> &r)
    {
        const Light<T> &lt = S.lights[li];
        Ph ph;
        ph.type = 0;
        double lp[3];
        light_point(lt, pp->seed, a, b, 0, P_LPOINT, msec, lp);
        const ProjMapT &pm = kind == 0 ? spec[li] : proj[li];
        double cov = pm.get_coverage();
        bool use = cov != 0.0;
        double e = (double)pp->photon_energy * (use ? cov : 1.0);
        for (int k = 0; k < 3; k++) ph.col[k] = (T)((double)(&lt.color.x)[k] * e);
        float d[3] = {0, 0, 1};
        for (uint32_t at = 0; at < PHOTON_ATTEMPTS; at++) {
            if (!marsaglia(prand(pp->seed, a, b, 0, P_EMIT, at, 0), prand(pp->seed, a, b, 0, P_EMIT, at, 1), d)) continue;
            if (!use) break;
            int uc, vc;
            cell_from_dir(d[0], d[1], d[2], uc, vc);
            if (pm.cells[vc * SRAY_PROJ_U + uc]) break;
        }
        r.o = v3<T>((T)lp[0], (T)lp[1], (T)lp[2]);
        r.d = v3<T>((T)d[0], (T)d[1], (T)d[2]) * (T)SRAY_RAY_MAGNITUDE;
        return ph;
    }

    /* Scene::build_caustics_map (kind 0) / build_global_map (kind 1) */
    void shoot(int kind, int num, std::vector<sray_photon> &out)
    {
        out.clear();
        if (num <= 0 || S.lights.empty()) return;
        double acc = 0;
        std::vector<double> inten(S.lights.size());
        for (size_t i = 0; i < S.lights.size(); i++) {
            const Light<T> &l = S.lights[i];
            inten[i] = ((double)l.color.x + (double)l.color.y + (double)l.color.z) / 3.0;
            acc += inten[i];
        }
        int t0 = pp->t0, t1 = pp->t1;
        for (size_t li = 0; li < S.lights.size(); li++) {
            int nphot = (int)std::ceil((double)num * (inten[li] / acc));
            size_t first = out.size();
            uint32_t b = ((uint32_t)kind << 24) | (uint32_t)li;
            for (int j = 0; j < nphot; j++) {
                uint32_t a = (uint32_t)j;
                int msec = t0 == t1 ? t0 : (int)(rng_u32(pp->seed, a, b, 0, slot(K_PHOTON, P_TIME, 0, 0, 0)) %
                                                 (uint32_t)(t1 - t0)) + t0;
                Ray<T> r;
                Ph ph = gen_photon(li, kind, a, b, msec, r);
                r.time = msec;
                r.iter = pp->iter;
                uint32_t bounce = 1;
                bool ok = kind == 0 ? (trace_caustics(r, ph, a, b, bounce) && ph.type == 1)
                                    : trace_global(r, ph, a, b, bounce);
                if (!ok) continue;
                sray_photon sp;
                sp.pos[0] = (float)ph.pos.x, sp.pos[1] = (float)ph.pos.y, sp.pos[2] = (float)ph.pos.z;
                sp.dir[0] = (float)ph.dir.x, sp.dir[1] = (float)ph.dir.y, sp.dir[2] = (float)ph.dir.z;
                sp.norm[0] = (float)ph.norm.x, sp.norm[1] = (float)ph.norm.y, sp.norm[2] = (float)ph.norm.z;
                for (int k = 0; k < 3; k++) sp.col[k] = (float)ph.col[k];
                out.push_back(sp);
            }
            /* PhotonMap::scale_photon_power(1/nphot), src/pmap.cc:90-100 */
            if (nphot) {
                double sc = 1.0 / (double)nphot;
                for (size_t k = first; k < out.size(); k++)
                    for (int c = 0; c < 3; c++) out[k].col[c] = (float)((double)out[k].col[c] * sc);
            }
        }
    }
};

/* ------------------------------------------------------------------ */
/* scene construction from the flat description                         */
/* ------------------------------------------------------------------ */
template <class T> static void load(Scene<T> &S, const sray_scene_desc *d, uint32_t flags)
{
    S.use_tree = !(flags & ORACLE_BRUTE_FORCE);
    S.q9_retraverse = (flags & ORACLE_Q9_RETRAVERSE) != 0;
    S.brute_gather = (flags & ORACLE_BRUTE_GATHER) != 0;
    S.frame_t0 = d->frame_t0;
    S.nrow = d->frame_t1 >= d->frame_t0 ? d->frame_t1 - d->frame_t0 + 1 : 1;
    const size_t nrow = (size_t)S.nrow;
    S.tris.resize(d->num_triangles);
    for (uint32_t i = 0; i < d->num_triangles; i++) {
        Tri<T> &t = S.tris[i];
        for (int k = 0; k < 3; k++) {
            const float *p = d->tri_pos + (size_t)i * 9 + k * 3;
            t.v[k] = v3<T>(p[0], p[1], p[2]);
            if (d->tri_normal) {
                const float *n = d->tri_normal + (size_t)i * 9 + k * 3;
                t.n[k] = v3<T>(n[0], n[1], n[2]);
            } else t.n[k] = v3<T>(0, 0, 0);
            if (d->tri_tangent) {
                const float *g = d->tri_tangent + (size_t)i * 9 + k * 3;
                t.g[k] = v3<T>(g[0], g[1], g[2]);
            } else t.g[k] = v3<T>(0, 0, 0);
            if (d->tri_texcoord) {
                t.tu[k] = d->tri_texcoord[(size_t)i * 6 + k * 2];
                t.tv[k] = d->tri_texcoord[(size_t)i * 6 + k * 2 + 1];
            } else t.tu[k] = t.tv[k] = 0;
        }
    }
    S.meshes.resize(d->num_meshes);
    for (uint32_t m = 0; m < d->num_meshes; m++) {
        /* Mesh::build_tree, src/mesh.cc:210-228 (local space, at load) */
        MeshT<T> &mt = S.meshes[m];
        mt.start = d->meshes[m].tri_start;
        mt.count = d->meshes[m].tri_count;
        mt.tree.max_depth = d->mesh_oct_depth;
        mt.tree.items_per_node = d->mesh_oct_items;
        for (uint32_t f = 0; f < mt.count; f++) mt.tree.add(Scene<T>::tri_box(S.tris[mt.start + f]));
        mt.tree.build();
    }
    S.objs.resize(d->num_objects);
    for (uint32_t i = 0; i < d->num_objects; i++) {
        const sray_object_desc &od = d->objects[i];
        Object<T> &o = S.objs[i];
        o.type = od.type;
        o.mat = od.material;
        o.mesh = od.mesh;
        o.rad = od.radius;
        o.start = od.start;
        o.end = od.end;
        o.xf = mat_from<T>(od.xform);
        o.inv = mat_from<T>(od.inv_xform);
        o.motion = v3<T>(od.motion[0], od.motion[1], od.motion[2]);
        o.mt0 = od.motion_t0;
        o.mt1 = od.motion_t1;
        o.moving = (od.motion[0] != 0 || od.motion[1] != 0 || od.motion[2] != 0) && od.motion_t1 > od.motion_t0;
        if (od.time_xforms) o.rows.assign(od.time_xforms, od.time_xforms + (size_t)nrow * 24);
    }
    S.mats.resize(d->num_materials);
    for (uint32_t i = 0; i < d->num_materials; i++) {
        const sray_material_desc &md = d->materials[i];
        Material<T> &m = S.mats[i];
        m.kd = cfrom<T>(md.diffuse);
        m.ks = cfrom<T>(md.specular);
        m.shininess = md.shininess;
        m.reflect = md.reflect;
        m.refract = md.refract;
        m.ior = md.ior;
        m.flags = md.flags;
        for (int k = 0; k < SRAY_NUM_MAPS; k++) m.map[k] = md.map[k];
        m.nscale = md.normal;
    }
    S.texs.resize(d->num_textures);
    for (uint32_t i = 0; i < d->num_textures; i++) {
        const sray_texture_desc &td = d->textures[i];
        Texture &t = S.texs[i];
        t.w = td.width;
        t.h = td.height;
        t.filter = td.filter;
        t.wrap = td.wrap;
        std::memcpy(t.xf, td.xform, sizeof t.xf);
        t.px.assign(td.texels, td.texels + (size_t)td.width * td.height * 4);
    }
    S.lights.resize(d->num_lights);
    for (uint32_t i = 0; i < d->num_lights; i++) {
        const sray_light_desc &ld = d->lights[i];
        Light<T> &l = S.lights[i];
        l.type = ld.type;
        l.pos = v3<T>(ld.pos[0], ld.pos[1], ld.pos[2]);
        l.color = cfrom<T>(ld.color);
        l.radius = ld.radius;
        l.dim = v3<T>(ld.dim[0], ld.dim[1], ld.dim[2]);
        l.xf = mat_from<T>(ld.xform);
        if (ld.time_pos && ld.type != SRAY_LIGHT_BOX) l.rows.assign(ld.time_pos, ld.time_pos + (size_t)nrow * 3);
    }
    S.cam_xf = mat_from<T>(d->camera.xform);
    if (d->camera.time_xforms) S.cam_rows.assign(d->camera.time_xforms, d->camera.time_xforms + (size_t)nrow * 12);
    S.vfov = d->camera.vfov;
    S.max_dist = d->camera.max_dist;
    S.shutter = d->camera.shutter;
    S.env_ambient = cfrom<T>(d->env_ambient);
    S.env_color = cfrom<T>(d->env_color);

    /* Scene::build_tree, src/scene.cc:353-390 */
    S.tree.clear();
    S.tree.max_depth = d->scene_oct_depth;
    S.tree.items_per_node = d->scene_oct_items;
    int t0 = d->frame_t0, t1 = d->frame_t1;
    for (uint32_t i = 0; i < d->num_objects; i++) {
        Box<T> a = S.bounds(S.objs[i], t0), b = S.bounds(S.objs[i], t1), u;
        u.mn = v3<T>(std::min(a.mn.x, b.mn.x), std::min(a.mn.y, b.mn.y), std::min(a.mn.z, b.mn.z));
        u.mx = v3<T>(std::max(a.mx.x, b.mx.x), std::max(a.mx.y, b.mx.y), std::max(a.mx.z, b.mx.z));
        S.tree.add(u);
    }
    S.tree.build();
    /* diagonal in double from the root box (src/scene.cc:384-387) */
    const Box<T> &rb = S.tree.nodes[0].box;
    if (d->num_objects) {
        double dx = (double)rb.mx.x - (double)rb.mn.x, dy = (double)rb.mx.y - (double)rb.mn.y,
               dz = (double)rb.mx.z - (double)rb.mn.z;
        S.diag = std::sqrt(dx * dx + dy * dy + dz * dz);
    } else {
        S.diag = 0;
    }
}

/* ------------------------------------------------------------------ */
/* render_block + variance, src/render.cc:194-244, 261-270              */
/* ------------------------------------------------------------------ */
template <class T> struct Renderer {
    const Scene<T> &S;
    const sray_render_params *p;
    float *rgba;
    int W;
    std::atomic<uint64_t> cnt[6];
    Renderer(const Scene<T> &s, const sray_render_params *pp, float *out) : S(s), p(pp), rgba(out)
    {
        W = p->x1 - p->x0;
        for (auto &c : cnt) c = 0;
    }

    void block(int bx0, int by0, int bw, int bh)
    {
        int maxs = std::max(1, p->max_samples);
        std::vector<C4<T>> sub(maxs);
        uint64_t rc[5] = {0, 0, 0, 0, 0};
        uint64_t nsamp = 0;
        std::vector<int> nsm(bw);
        for (int y = by0; y < by0 + bh; y++) {
            float *row = rgba + ((size_t)(y - p->y0) * W + (bx0 - p->x0)) * 4;
            for (int x = bx0; x < bx0 + bw; x++) {
                float *px = row + (x - bx0) * 4;
                px[0] = px[1] = px[2] = px[3] = 0.0f;
                int i = 0;
                uint32_t pix = (uint32_t)y * (uint32_t)p->width + (uint32_t)x;
                while (i < maxs) {
                    Ray<T> r = S.primary(x, y, i, p->frame_time, p, p->seed);
                    r.iter = p->iter;
                    typename Scene<T>::ShadeCtx sc{p, p->seed, pix, (uint32_t)i, rc};
                    rc[0]++;
                    sub[i] = S.trace(r, sc, 1);
                    /* img[] += subpix[]: float framebuffer accumulation */
                    px[0] = (float)((T)px[0] + sub[i].x);
                    px[1] = (float)((T)px[1] + sub[i].y);
                    px[2] = (float)((T)px[2] + sub[i].z);
                    px[3] = (float)((T)px[3] + sub[i].w);
                    i++;
                    if (i >= p->min_samples && i > 1) {
                        float rcp = (float)((T)1 / (T)i);
                        C4<T> mean = c4<T>((T)px[0] * (T)rcp, (T)px[1] * (T)rcp, (T)px[2] * (T)rcp, (T)px[3] * (T)rcp);
                        float sq = 0.0f;
                        for (int k = 0; k < i; k++) {
                            C4<T> dd = mean - sub[k];
                            T l2 = dd.x * dd.x + dd.y * dd.y + dd.z * dd.z + dd.w * dd.w;
                            sq = (float)((T)sq + l2);
                        }
                        float v = sq * rcp;
                        if (v < p->max_var) break;
                    }
                }
                nsamp += (uint64_t)i;
                nsm[x - bx0] = i;
                if (!(p->flags & SRAY_RENDER_REF_Q1)) {
                    T rcp = (T)1 / (T)i;
                    for (int c = 0; c < 4; c++) px[c] = (float)((T)px[c] * rcp);
                }
            }
            if (p->flags & SRAY_RENDER_REF_Q1) {
                /* img[x] *= rcp_lut[i] (src/render.cc:240): float x of the
                 * block row, scaled by pixel x's own sample reciprocal */
                for (int x = 0; x < bw; x++) {
                    T rcp = (T)1 / (T)nsm[x];
                    row[x] = (float)((T)row[x] * rcp);
                }
            }
        }
        for (int k = 0; k < 5; k++) cnt[k] += rc[k];
        cnt[5] += nsamp;
    }

    /* start_frame: blk_sz tiles sorted by gaussian centre priority
     * (src/render.cc:144-192, src/block.c:83-88), run on a FIFO pool */
    void run(int nthreads)
    {
        const int bs = 48;
        struct Blk {
            int x, y, w, h, pri;
        };
        std::vector<Blk> blks;
        int xb = ((p->width << 8) / bs + 255) >> 8;
        int yb = ((p->height << 8) / bs + 255) >> 8;
        for (int i = 0; i < xb; i++)
            for (int j = 0; j < yb; j++) {
                Blk b;
                b.x = i * bs;
                b.y = j * bs;
                b.w = std::min(bs, p->width - b.x);
                b.h = std::min(bs, p->height - b.y);
                /* clip to the requested rectangle */
                int cx0 = std::max(b.x, p->x0), cy0 = std::max(b.y, p->y0);
                int cx1 = std::min(b.x + b.w, p->x1), cy1 = std::min(b.y + b.h, p->y1);
                double dx = ((double)b.x + (double)b.w / 2.0) - (double)p->width / 2.0;
                double dy = ((double)b.y + (double)b.h / 2.0) - (double)p->height / 2.0;
                dx /= (double)p->width * 4.0;
                dy /= (double)p->height * 4.0;
                double g = std::exp(-(dx * dx + dy * dy) / (2.0 * 0.01 * 0.01)) / (0.01 * std::sqrt(2.0 * M_PI));
                b.pri = (int)(g * 200.0);
                if (cx0 >= cx1 || cy0 >= cy1) continue;
                b.x = cx0;
                b.y = cy0;
                b.w = cx1 - cx0;
                b.h = cy1 - cy0;
                blks.push_back(b);
            }
        std::stable_sort(blks.begin(), blks.end(), [](const Blk &a, const Blk &b) { return a.pri > b.pri; });
        std::atomic<size_t> next{0};
        auto worker = [&]() {
            for (;;) {
                size_t k = next++;
                if (k >= blks.size()) break;
                block(blks[k].x, blks[k].y, blks[k].w, blks[k].h);
            }
        };
        nthreads = std::max(1, nthreads);
        std::vector<std::thread> th;
        for (int i = 1; i < nthreads; i++) th.emplace_back(worker);
        worker();
        for (auto &t : th) t.join();
    }
};

struct AnyScene {
    int prec;
    std::unique_ptr<Scene<float>> f;
    std::unique_ptr<Scene<double>> d;
};

/* SurfPoint of the closest hit of each ray (pos, normal, u, v); rays that
 * miss get NaN (src/sphere.cc:107-129, src/cylinder.cc:166-182,
 * src/mesh.cc:147-163, 293-306, interp_face :321-445) */
template <class T> static int do_surface(Scene<T> &S, const sray_ray *rays, uint32_t n, float *out)
{
    for (uint32_t i = 0; i < n; i++) {
        Ray<T> r;
        r.o = v3<T>(rays[i].origin[0], rays[i].origin[1], rays[i].origin[2]);
        r.d = v3<T>(rays[i].dir[0], rays[i].dir[1], rays[i].dir[2]);
        r.time = rays[i].time;
        Counters cnt;
        Best<T> b = S.cast(r, cnt, false);
        float *o = out + (size_t)i * 8;
        if (!b.found()) {
            for (int k = 0; k < 8; k++) o[k] = NAN;
            continue;
        }
        SurfPoint<T> sp = S.surf(r, (int)b.item, b.sub, b.t);
        o[0] = (float)sp.pos.x, o[1] = (float)sp.pos.y, o[2] = (float)sp.pos.z;
        o[3] = (float)sp.normal.x, o[4] = (float)sp.normal.y, o[5] = (float)sp.normal.z;
        o[6] = (float)sp.u, o[7] = (float)sp.v;
    }
    return 0;
}

/* the shading normal shade_phong uses at each closest hit (the normal
 * facing the ray, then the "normal" attribute's, src/shader.cc:88-96,
 * 137-138) and the surface tangent: 6 floats, NaN on a miss */
template <class T> static int do_shade_normal(Scene<T> &S, const sray_ray *rays, uint32_t n, float *out)
{
    for (uint32_t i = 0; i < n; i++) {
        Ray<T> r;
        r.o = v3<T>(rays[i].origin[0], rays[i].origin[1], rays[i].origin[2]);
        r.d = v3<T>(rays[i].dir[0], rays[i].dir[1], rays[i].dir[2]);
        r.time = rays[i].time;
        Counters cnt;
        Best<T> b = S.cast(r, cnt, false);
        float *o = out + (size_t)i * 6;
        if (!b.found()) {
            for (int k = 0; k < 6; k++) o[k] = NAN;
            continue;
        }
        SurfPoint<T> sp = S.surf(r, (int)b.item, b.sub, b.t);
        V3<T> nrm = dot(r.d, sp.normal) > (T)0 ? -sp.normal : sp.normal;
        const int m = S.objs[b.item].mat;
        if (m >= 0 && (S.mats[m].flags & SRAY_MAT_HAS_NORMAL)) nrm = S.bump_normal(S.mats[m], sp);
        o[0] = (float)nrm.x, o[1] = (float)nrm.y, o[2] = (float)nrm.z;
        o[3] = (float)sp.tangent.x, o[4] = (float)sp.tangent.y, o[5] = (float)sp.tangent.z;
    }
    return 0;
}

/* Texture::lookup of texture tex at n (u, v) pairs: RGBA */
template <class T> static int do_tex_lookup(Scene<T> &S, int tex, const float *uv, uint32_t n, float *out)
{
    if (tex < 0 || tex >= (int)S.texs.size()) return -1;
    for (uint32_t i = 0; i < n; i++) {
        C4<T> c = Scene<T>::tex_lookup(S.texs[tex], (T)uv[2 * i], (T)uv[2 * i + 1]);
        out[4 * i] = (float)c.x, out[4 * i + 1] = (float)c.y, out[4 * i + 2] = (float)c.z, out[4 * i + 3] = (float)c.w;
    }
    return 0;
}

template <class T> static int do_trace(Scene<T> &S, const sray_ray *rays, uint32_t n, sray_hit *hits,
                                       oracle_hit_ext *ext, int any, int nthreads)
{
    std::atomic<uint32_t> next{0};
    const uint32_t chunk = 256;
    auto worker = [&]() {
        for (;;) {
            uint32_t b = next.fetch_add(chunk);
            if (b >= n) break;
            uint32_t e = std::min(n, b + chunk);
            for (uint32_t i = b; i < e; i++) {
                Ray<T> r;
                r.o = v3<T>(rays[i].origin[0], rays[i].origin[1], rays[i].origin[2]);
                r.d = v3<T>(rays[i].dir[0], rays[i].dir[1], rays[i].dir[2]);
                r.time = rays[i].time;
                Counters cnt;
                Best<T> bst = S.cast(r, cnt, any != 0);
                sray_hit &h = hits[i];
                if (bst.found()) {
                    h.t = (float)bst.t;
                    h.obj = (int32_t)bst.item;
                    h.prim = (int32_t)bst.sub;
                    h.rank = bst.rank;
                } else {
                    h.t = FLT_MAX;
                    h.obj = -1;
                    h.prim = -1;
                    h.rank = 0;
                }
                if (ext) {
                    ext[i].tie = bst.tie;
                    ext[i].nodes = cnt.nodes;
                    ext[i].prims = cnt.prims;
                }
            }
        }
    };
    nthreads = std::max(1, nthreads);
    std::vector<std::thread> th;
    for (int i = 1; i < nthreads; i++) th.emplace_back(worker);
    worker();
    for (auto &t : th) t.join();
    return 0;
}

template <class T> static void tree_dump(const Octree<T> &tr, int n, std::vector<uint32_t> &items, std::vector<uint32_t> &counts)
{
    const OctNode<T> &nd = tr.nodes[n];
    if (!nd.items.empty()) {
        counts.push_back((uint32_t)nd.items.size());
        items.insert(items.end(), nd.items.begin(), nd.items.end());
    }
    for (int i = 0; i < 8; i++)
        if (nd.child[i] >= 0) tree_dump(tr, nd.child[i], items, counts);
}

template <class T> static int do_tree(Scene<T> &S, int level, oracle_tree_info *info, uint32_t *li, uint32_t *lc,
                                      uint32_t cap_items, uint32_t cap_leaves)
{
    const Octree<T> *tr;
    if (level < 0) tr = &S.tree;
    else {
        if ((size_t)level >= S.meshes.size()) return -1;
        tr = &S.meshes[level].tree;
    }
    std::vector<uint32_t> items, counts;
    if (!tr->nodes.empty()) tree_dump(*tr, 0, items, counts);
    info->num_nodes = (uint32_t)tr->nodes.size();
    info->num_leaves = (uint32_t)counts.size();
    info->num_item_refs = (uint32_t)items.size();
    info->height = tr->height;
    if (!tr->nodes.empty()) {
        const Box<T> &b = tr->nodes[0].box;
        info->root_min[0] = (double)b.mn.x;
        info->root_min[1] = (double)b.mn.y;
        info->root_min[2] = (double)b.mn.z;
        info->root_max[0] = (double)b.mx.x;
        info->root_max[1] = (double)b.mx.y;
        info->root_max[2] = (double)b.mx.z;
    }
    if (li && items.size() <= cap_items) std::copy(items.begin(), items.end(), li);
    if (lc && counts.size() <= cap_leaves) std::copy(counts.begin(), counts.end(), lc);
    return 0;
}

template <class T> static int do_render(Scene<T> &S, const sray_render_params *p, float *rgba, int nthreads, sray_stats *st)
{
    /* gather radius relative to the scene octree diagonal (Q19) */
    S.gather_dist_abs = S.diag * (double)p->gather_dist;
    for (int kind = 0; kind < 2; kind++) S.build_grid(kind, S.gather_dist_abs);
    Renderer<T> R(S, p, rgba);
    R.run(nthreads);
    if (st) {
        memset(st, 0, sizeof *st);
        st->rays_primary = R.cnt[0];
        st->rays_shadow = R.cnt[1];
        st->rays_reflect = R.cnt[2];
        st->rays_refract = R.cnt[3];
        st->rays_diffuse = R.cnt[4];
        st->samples = R.cnt[5];
    }
    return 0;
}

template <class T> static int do_primary(Scene<T> &S, const sray_render_params *p, int sample, sray_ray *rays)
{
    int W = p->x1 - p->x0;
    for (int y = p->y0; y < p->y1; y++)
        for (int x = p->x0; x < p->x1; x++) {
            Ray<T> r = S.primary(x, y, sample, p->frame_time, p, p->seed);
            sray_ray &o = rays[(size_t)(y - p->y0) * W + (x - p->x0)];
            o.origin[0] = (float)r.o.x;
            o.origin[1] = (float)r.o.y;
            o.origin[2] = (float)r.o.z;
            o.dir[0] = (float)r.d.x;
            o.dir[1] = (float)r.d.y;
            o.dir[2] = (float)r.d.z;
            o.time = r.time;
            o.pad_ = 0;
        }
    return 0;
}

template <class T> static int do_gather(Scene<T> &S, int kind, const sray_gather_query *q, uint32_t n, float radius,
                                        int k, int mode, float *rgb)
{
    S.build_grid(kind, (double)radius);
    /* the estimates read the caustic (irradiance) or global (radiance) map */
    const int slot = mode == SRAY_GATHER_IRRADIANCE ? SRAY_PMAP_CAUSTIC : SRAY_PMAP_GLOBAL;
    std::swap(S.grid[slot], S.grid[kind]);
    for (uint32_t i = 0; i < n; i++) {
        V3<T> pos = v3<T>(q[i].pos[0], q[i].pos[1], q[i].pos[2]);
        V3<T> nrm = v3<T>(q[i].norm[0], q[i].norm[1], q[i].norm[2]);
        C4<T> c;
        if (mode == SRAY_GATHER_IRRADIANCE) {
            /* the irradiance estimate reads the map of the given kind */
            std::swap(S.pm[SRAY_PMAP_CAUSTIC], S.pm[kind]);
            c = S.irradiance(pos, nrm, (double)radius, k);
            std::swap(S.pm[SRAY_PMAP_CAUSTIC], S.pm[kind]);
        } else {
            std::swap(S.pm[SRAY_PMAP_GLOBAL], S.pm[kind]);
            c = S.radiance(pos, nrm, (double)radius, k);
            std::swap(S.pm[SRAY_PMAP_GLOBAL], S.pm[kind]);
        }
        rgb[i * 3 + 0] = (float)c.x;
        rgb[i * 3 + 1] = (float)c.y;
        rgb[i * 3 + 2] = (float)c.z;
    }
    std::swap(S.grid[slot], S.grid[kind]);
    return 0;
}

template <class T>
static int do_shoot(Scene<T> &S, const sray_photon_params *pp, sray_photon *caust, uint32_t cap_c, uint32_t *n_c,
                    sray_photon *glob, uint32_t cap_g, uint32_t *n_g, uint64_t *segments)
{
    PhotonTracer<T> tr(S, pp);
    tr.build_maps(pp->t0, pp->t1);
    std::vector<sray_photon> c, g;
    tr.shoot(0, pp->caust_photons, c);
    tr.shoot(1, pp->gi_photons, g);
    if (n_c) *n_c = (uint32_t)c.size();
    if (n_g) *n_g = (uint32_t)g.size();
    if (caust && c.size() <= cap_c) std::copy(c.begin(), c.end(), caust);
    if (glob && g.size() <= cap_g) std::copy(g.begin(), g.end(), glob);
    if (segments) *segments = tr.segments;
    /* the maps also feed this scene's renders */
    S.pm[0].resize(c.size());
    S.pm[1].resize(g.size());
    for (size_t i = 0; i < c.size(); i++) memcpy(&S.pm[0][i], &c[i], sizeof(Photon));
    for (size_t i = 0; i < g.size(); i++) memcpy(&S.pm[1][i], &g[i], sizeof(Photon));
    S.photons_changed(0);
    S.photons_changed(1);
    return 0;
}

template <class T>
static int do_projmap(Scene<T> &S, int light, int kind, int t0, int t1, uint64_t seed, uint8_t *cells, double *cov)
{
    if (light < 0 || (size_t)light >= S.lights.size()) return -1;
    sray_photon_params pp;
    memset(&pp, 0, sizeof pp);
    pp.seed = seed;
    PhotonTracer<T> tr(S, &pp);
    tr.build_maps(t0, t1);
    const ProjMapT &m = kind == 0 ? tr.proj[light] : tr.spec[light];
    if (cells) std::copy(m.cells.begin(), m.cells.end(), cells);
    if (cov) *cov = m.get_coverage();
    return 0;
}

} // namespace orc

using namespace orc;

/* n sample points of one light (Scene::light_point, shadow sample si = 0,
 * path 0), sample index i for point i: the draws the renderer makes for
 * sample i of a pixel */
template <class T>
static int do_light_points(const Scene<T> &S, int light, uint64_t seed, uint32_t pixel, uint32_t n, double *out)
{
    if (light < 0 || (size_t)light >= S.lights.size()) return -1;
    typename Scene<T>::ShadeCtx sc{};
    sc.seed = seed;
    sc.pixel = pixel;
    for (uint32_t i = 0; i < n; i++) {
        sc.sample = i;
        V3<T> v = S.light_point(S.lights[(size_t)light], sc, 0u, (uint32_t)light, 0u, S.frame_t0);
        out[(size_t)i * 3] = (double)v.x;
        out[(size_t)i * 3 + 1] = (double)v.y;
        out[(size_t)i * 3 + 2] = (double)v.z;
    }
    return 0;
}

struct oracle_scene {
    AnyScene s;
};

extern "C" {

oracle_scene *oracle_scene_create(const sray_scene_desc *desc, int prec, uint32_t flags)
{
    if (!desc) return nullptr;
    oracle_scene *o = new oracle_scene;
    o->s.prec = prec;
    if (prec == 32) {
        o->s.f.reset(new Scene<float>);
        load(*o->s.f, desc, flags);
    } else {
        o->s.d.reset(new Scene<double>);
        load(*o->s.d, desc, flags);
    }
    return o;
}

void oracle_scene_destroy(oracle_scene *s) { delete s; }

int oracle_trace(oracle_scene *s, const sray_ray *rays, uint32_t n, sray_hit *hits, oracle_hit_ext *ext, int any,
                 int nthreads)
{
    if (!s || (n && (!rays || !hits))) return -1;
    if (s->s.prec == 32) return do_trace(*s->s.f, rays, n, hits, ext, any, nthreads);
    return do_trace(*s->s.d, rays, n, hits, ext, any, nthreads);
}

int oracle_tree(oracle_scene *s, int level, oracle_tree_info *info, uint32_t *leaf_items, uint32_t *leaf_counts,
                uint32_t cap_items, uint32_t cap_leaves)
{
    if (!s || !info) return -1;
    if (s->s.prec == 32) return do_tree(*s->s.f, level, info, leaf_items, leaf_counts, cap_items, cap_leaves);
    return do_tree(*s->s.d, level, info, leaf_items, leaf_counts, cap_items, cap_leaves);
}

int oracle_primary_rays(oracle_scene *s, const sray_render_params *p, int sample, sray_ray *rays)
{
    if (!s || !p || !rays) return -1;
    if (s->s.prec == 32) return do_primary(*s->s.f, p, sample, rays);
    return do_primary(*s->s.d, p, sample, rays);
}

int oracle_render(oracle_scene *s, const sray_render_params *p, float *rgba, int nthreads, sray_stats *st)
{
    if (!s || !p || !rgba) return -1;
    if (p->x0 < 0 || p->y0 < 0 || p->x1 > p->width || p->y1 > p->height || p->x0 >= p->x1 || p->y0 >= p->y1)
        return -1;
    if (s->s.prec == 32) return do_render(*s->s.f, p, rgba, nthreads, st);
    return do_render(*s->s.d, p, rgba, nthreads, st);
}

int oracle_set_photons(oracle_scene *s, int kind, const sray_photon *p, uint32_t n)
{
    if (!s || kind < 0 || kind > 1) return -1;
    std::vector<Photon> v(n);
    for (uint32_t i = 0; i < n; i++) memcpy(&v[i], &p[i], sizeof(Photon));
    if (s->s.prec == 32) {
        s->s.f->pm[kind] = v;
        s->s.f->photons_changed(kind);
    } else {
        s->s.d->pm[kind] = v;
        s->s.d->photons_changed(kind);
    }
    return 0;
}

int oracle_gather(oracle_scene *s, int kind, const sray_gather_query *q, uint32_t n, float radius, int k, int mode,
                  float *rgb)
{
    if (!s || kind < 0 || kind > 1) return -1;
    if (s->s.prec == 32) return do_gather(*s->s.f, kind, q, n, radius, k, mode, rgb);
    return do_gather(*s->s.d, kind, q, n, radius, k, mode, rgb);
}

int oracle_shoot_photons(oracle_scene *s, const sray_photon_params *pp, sray_photon *caust, uint32_t cap_c,
                         uint32_t *n_c, sray_photon *glob, uint32_t cap_g, uint32_t *n_g, uint64_t *segments)
{
    if (!s || !pp) return -1;
    if (s->s.prec == 32) return do_shoot(*s->s.f, pp, caust, cap_c, n_c, glob, cap_g, n_g, segments);
    return do_shoot(*s->s.d, pp, caust, cap_c, n_c, glob, cap_g, n_g, segments);
}

int oracle_projmap(oracle_scene *s, int light, int kind, int t0, int t1, uint64_t seed, uint8_t *cells,
                   double *coverage)
{
    if (!s) return -1;
    if (s->s.prec == 32) return do_projmap(*s->s.f, light, kind, t0, t1, seed, cells, coverage);
    return do_projmap(*s->s.d, light, kind, t0, t1, seed, cells, coverage);
}

int oracle_cell_from_dir(float x, float y, float z, int *uc, int *vc)
{
    int u, v;
    cell_from_dir(x, y, z, u, v);
    if (uc) *uc = u;
    if (vc) *vc = v;
    return 0;
}

int oracle_surface(oracle_scene *s, const sray_ray *rays, uint32_t n, float *out)
{
    if (!s || (n && (!rays || !out))) return -1;
    if (s->s.prec == 32) return do_surface(*s->s.f, rays, n, out);
    return do_surface(*s->s.d, rays, n, out);
}

int oracle_shade_normal(oracle_scene *s, const sray_ray *rays, uint32_t n, float *out)
{
    if (!s || (n && (!rays || !out))) return -1;
    if (s->s.prec == 32) return do_shade_normal(*s->s.f, rays, n, out);
    return do_shade_normal(*s->s.d, rays, n, out);
}

int oracle_tex_lookup(oracle_scene *s, int tex, const float *uv, uint32_t n, float *out)
{
    if (!s || (n && (!uv || !out))) return -1;
    if (s->s.prec == 32) return do_tex_lookup(*s->s.f, tex, uv, n, out);
    return do_tex_lookup(*s->s.d, tex, uv, n, out);
}

double oracle_scene_diag(oracle_scene *s)
{
    if (!s) return 0;
    return s->s.prec == 32 ? s->s.f->diag : s->s.d->diag;
}

int oracle_light_points(oracle_scene *s, int light, uint64_t seed, uint32_t pixel, uint32_t n, double *out)
{
    if (!s || (n && !out)) return -1;
    if (s->s.prec == 32) return do_light_points(*s->s.f, light, seed, pixel, n, out);
    return do_light_points(*s->s.d, light, seed, pixel, n, out);
}

uint32_t oracle_rng_u32(uint64_t seed, uint32_t a, uint32_t b, uint32_t c, uint32_t d)
{
    return rng_u32(seed, a, b, c, d);
}

uint64_t oracle_splitmix64(uint64_t x) { return splitmix64(x); }

} /* extern "C" */
