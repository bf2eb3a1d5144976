/*
 * accel.cc — octree build and flattening for the device traversal.
 *
 * Build rules are the reference's (Octree<T>::add/build/subdivide,
 * src/octree.inl:100-191): root box = union of item boxes; a node splits iff
 * it holds more than max_items items and its level is below max_depth;
 * children are the 8 octants around the box centre, numbered as in
 * src/octree.inl:161-169; an item goes to every child its box overlaps
 * (closed intervals, AABox::in_box, src/aabb.cc:53-65).  The tree is kept as
 * index lists (no std::list of pointers) and flattened into the two-float4
 * node arrays the kernels read; see host.h for the layout.
 */
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>

#include "host.h"

namespace sray {

namespace {

/* reference octant index -> (x | y<<1 | z<<2) bits, src/octree.inl:161-169 */
const int kRefToBits[8] = {0, 1, 3, 2, 4, 5, 7, 6};

struct BNode {
    float mn[3], mx[3];
    std::vector<uint32_t> items;
    int child[8];
    bool split = false;
};

struct Builder {
    const std::vector<float> &box;
    int max_items, max_depth;
    std::vector<BNode> nodes;
    uint32_t height = 0;
    /* budget: past these the device layout could not hold the tree anyway
     * (node indices are 23-bit), and a tiny leaf size with a large depth
     * would otherwise grow without bound where items share points */
    static constexpr size_t MAX_NODES = size_t(1) << 23, MAX_REFS = size_t(1) << 28;
    size_t refs = 0;
    bool overflow = false;

    Builder(const std::vector<float> &b, int mi, int md) : box(b), max_items(mi), max_depth(md) {}

    static bool overlap(const float *a, const float *bmn, const float *bmx)
    {
        /* a = item box (min xyz, max xyz) */
        for (int k = 0; k < 3; k++)
            if (a[k] > bmx[k] || bmn[k] > a[3 + k]) return false;
        return true;
    }

    int add_node()
    {
        nodes.emplace_back();
        BNode &n = nodes.back();
        for (int i = 0; i < 8; i++) n.child[i] = -1;
        return (int)nodes.size() - 1;
    }

    void subdivide(int ni, int lvl)
    {
        if (overflow || (int)nodes[ni].items.size() <= max_items || lvl >= max_depth) return;
        if (nodes.size() + 8 > MAX_NODES) {
            overflow = true;
            return;
        }
        float mn[3], mx[3], c[3];
        for (int k = 0; k < 3; k++) {
            mn[k] = nodes[ni].mn[k];
            mx[k] = nodes[ni].mx[k];
            c[k] = (mn[k] + mx[k]) * 0.5f;
        }
        if ((uint32_t)(lvl + 1) > height) height = (uint32_t)(lvl + 1);
        nodes[ni].split = true;
        std::vector<uint32_t> parent_items;
        parent_items.swap(nodes[ni].items);
        refs -= parent_items.size();
        for (int r = 0; r < 8; r++) {
            int bits = kRefToBits[r];
            float cmn[3], cmx[3];
            for (int k = 0; k < 3; k++) {
                bool hi = (bits >> k) & 1;
                cmn[k] = hi ? c[k] : mn[k];
                cmx[k] = hi ? mx[k] : c[k];
            }
            int ch = add_node();
            nodes[ni].child[r] = ch;
            std::memcpy(nodes[ch].mn, cmn, sizeof cmn);
            std::memcpy(nodes[ch].mx, cmx, sizeof cmx);
            std::vector<uint32_t> its;
            for (uint32_t it : parent_items)
                if (overlap(&box[(size_t)it * 6], cmn, cmx)) its.push_back(it);
            refs += its.size();
            if (refs > MAX_REFS) overflow = true;
            nodes[ch].items.swap(its);
            subdivide(ch, lvl + 1);
        }
    }

    void build()
    {
        size_t n = box.size() / 6;
        float mn[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, mx[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
        for (size_t i = 0; i < n; i++) {
            const float *b = &box[i * 6];
            for (int k = 0; k < 3; k++) {
                if (b[k] < mn[k]) mn[k] = b[k];
                if (b[3 + k] > mx[k]) mx[k] = b[3 + k];
            }
        }
        int root = add_node();
        std::memcpy(nodes[root].mn, mn, sizeof mn);
        std::memcpy(nodes[root].mx, mx, sizeof mx);
        nodes[root].items.resize(n);
        for (size_t i = 0; i < n; i++) nodes[root].items[i] = (uint32_t)i;
        refs = n;
        subdivide(root, 0);
    }

    bool stored(int ni) const { return nodes[ni].split || !nodes[ni].items.empty(); }

    /* write node ni at slot `at`; children blocks appended as visited;
     * items appended in depth-first reference order */
    void flatten(int ni, uint32_t at, uint32_t item_base, FlatTree &out, uint32_t &leaves)
    {
        const BNode &nd = nodes[ni];
        float *lo = &out.lo[(size_t)at * 4], *hi = &out.hi[(size_t)at * 4];
        for (int k = 0; k < 3; k++) {
            lo[k] = nd.mn[k];
            hi[k] = nd.mx[k];
        }
        uint32_t first, meta;
        if (!nd.split) {
            first = (uint32_t)out.items.size();
            meta = SR_LEAF | (uint32_t)nd.items.size();
            for (uint32_t it : nd.items) out.items.push_back(item_base + it);
            if (!nd.items.empty()) leaves++;
            std::memcpy(&lo[3], &first, 4);
            std::memcpy(&hi[3], &meta, 4);
            return;
        }
        uint32_t mask = 0;
        for (int r = 0; r < 8; r++)
            if (nd.child[r] >= 0 && stored(nd.child[r])) mask |= 1u << kRefToBits[r];
        /* eight slots, child with octant bits k at first + k; slots of
         * children that are not stored stay zero and are never visited */
        first = (uint32_t)(out.lo.size() / 4);
        out.lo.resize(out.lo.size() + 8 * 4, 0.0f);
        out.hi.resize(out.hi.size() + 8 * 4, 0.0f);
        lo = &out.lo[(size_t)at * 4];
        hi = &out.hi[(size_t)at * 4];
        meta = mask;
        std::memcpy(&lo[3], &first, 4);
        std::memcpy(&hi[3], &meta, 4);
        for (int r = 0; r < 8; r++) {
            int ch = nd.child[r];
            if (ch < 0 || !stored(ch)) continue;
            flatten(ch, first + (uint32_t)kRefToBits[r], item_base, out, leaves);
        }
    }
};

} // namespace

uint32_t build_flat_octree(const std::vector<float> &boxes, int max_items, int max_depth, uint32_t item_base,
                           FlatTree &out)
{
    Builder b(boxes, max_items, max_depth);
    b.build();
    if (b.overflow) {
        out.overflow = true;
        return 0;
    }
    uint32_t root = (uint32_t)(out.lo.size() / 4);
    out.lo.resize(out.lo.size() + 4);
    out.hi.resize(out.hi.size() + 4);
    uint32_t leaves = 0;
    b.flatten(0, root, item_base, out, leaves);
    out.height = b.height;
    out.num_leaves = leaves;
    for (int k = 0; k < 3; k++) {
        out.root_min[k] = b.nodes[0].mn[k];
        out.root_max[k] = b.nodes[0].mx[k];
    }
    return root;
}

/* ---------------------------------------------------------------- */
/* item boxes                                                         */
/* ---------------------------------------------------------------- */
namespace {

inline void xf_point(const float *m, float x, float y, float z, float out[3])
{
    out[0] = m[0] * x + m[1] * y + m[2] * z + m[3];
    out[1] = m[4] * x + m[5] * y + m[6] * z + m[7];
    out[2] = m[8] * x + m[9] * y + m[10] * z + m[11];
}

/* object matrix at time t: the keyframed row of that msec (the nearest row
 * outside the frame interval), or linear translation keys */
void xform_at(const sray_scene_desc *sd, const sray_object_desc &o, int time, float m[12])
{
    if (o.time_xforms) {
        const int n = sd->frame_t1 - sd->frame_t0 + 1;
        int k = time - sd->frame_t0;
        k = k < 0 ? 0 : k >= n ? n - 1 : k;
        std::memcpy(m, o.time_xforms + (size_t)k * 24, sizeof(float) * 12);
        return;
    }
    std::memcpy(m, o.xform, sizeof(float) * 12);
    bool moving = (o.motion[0] != 0 || o.motion[1] != 0 || o.motion[2] != 0) && o.motion_t1 > o.motion_t0;
    if (!moving) return;
    float s;
    if (time <= o.motion_t0) s = 0.0f;
    else if (time >= o.motion_t1) s = 1.0f;
    else s = (float)(time - o.motion_t0) / (float)(o.motion_t1 - o.motion_t0);
    m[3] = m[3] + o.motion[0] * s;
    m[7] = m[7] + o.motion[1] * s;
    m[11] = m[11] + o.motion[2] * s;
}

/* Object::calc_bounds: sphere src/sphere.cc:27-34, cylinder
 * src/cylinder.cc:63-86, mesh src/mesh.cc:169-188 */
void object_box(const sray_scene_desc *s, const sray_object_desc &o, int time, float b[6])
{
    float m[12];
    xform_at(s, o, time, m);
    if (o.type == SRAY_OBJ_SPHERE) {
        float p[3];
        xf_point(m, 0.0f, 0.0f, 0.0f, p);
        for (int k = 0; k < 3; k++) {
            b[k] = p[k] - o.radius;
            b[3 + k] = p[k] + o.radius;
        }
    } else if (o.type == SRAY_OBJ_CYLINDER) {
        float s1[3], s2[3];
        xf_point(m, 0.0f, o.start, 0.0f, s1);
        xf_point(m, 0.0f, o.end, 0.0f, s2);
        for (int k = 0; k < 3; k++) {
            float a0 = s1[k] - o.radius, a1 = s1[k] + o.radius;
            float b0 = s2[k] - o.radius, b1 = s2[k] + o.radius;
            b[k] = a0 < b0 ? a0 : b0;
            b[3 + k] = a1 > b1 ? a1 : b1;
        }
    } else {
        for (int k = 0; k < 3; k++) {
            b[k] = FLT_MAX;
            b[3 + k] = -FLT_MAX;
        }
        const sray_mesh_desc &md = s->meshes[o.mesh];
        for (uint32_t f = 0; f < md.tri_count; f++) {
            const float *v = s->tri_pos + (size_t)(md.tri_start + f) * 9;
            for (int c = 0; c < 3; c++) {
                float p[3];
                xf_point(m, v[c * 3], v[c * 3 + 1], v[c * 3 + 2], p);
                for (int k = 0; k < 3; k++) {
                    if (p[k] < b[k]) b[k] = p[k];
                    if (p[k] > b[3 + k]) b[3 + k] = p[k];
                }
            }
        }
    }
}

} // namespace

void object_box_at(const sray_scene_desc *s, const sray_object_desc &o, int time, float b[6])
{
    object_box(s, o, time, b);
}

/* Scene::build_tree, src/scene.cc:364-378: union of bounds at t0 and t1 */
void object_boxes(const sray_scene_desc *s, std::vector<float> &boxes)
{
    boxes.resize((size_t)s->num_objects * 6);
    for (uint32_t i = 0; i < s->num_objects; i++) {
        float a[6], c[6];
        object_box(s, s->objects[i], s->frame_t0, a);
        object_box(s, s->objects[i], s->frame_t1, c);
        float *b = &boxes[(size_t)i * 6];
        for (int k = 0; k < 3; k++) {
            b[k] = a[k] < c[k] ? a[k] : c[k];
            b[3 + k] = a[3 + k] > c[3 + k] ? a[3 + k] : c[3 + k];
        }
    }
}

/* Mesh::build_tree: Triangle::calc_bounds with the identity, src/mesh.cc:38-54, 210-228 */
void triangle_boxes(const sray_scene_desc *s, uint32_t mesh, std::vector<float> &boxes)
{
    const sray_mesh_desc &md = s->meshes[mesh];
    boxes.resize((size_t)md.tri_count * 6);
    for (uint32_t f = 0; f < md.tri_count; f++) {
        const float *v = s->tri_pos + (size_t)(md.tri_start + f) * 9;
        float *b = &boxes[(size_t)f * 6];
        for (int k = 0; k < 3; k++) {
            b[k] = FLT_MAX;
            b[3 + k] = -FLT_MAX;
        }
        for (int c = 0; c < 3; c++)
            for (int k = 0; k < 3; k++) {
                float p = v[c * 3 + k];
                if (p < b[k]) b[k] = p;
                if (p > b[3 + k]) b[3 + k] = p;
            }
    }
}

/* ---------------------------------------------------------------- */
/* matrices (double, host)                                            */
/* ---------------------------------------------------------------- */
void mat_identity(double m[12])
{
    for (int i = 0; i < 12; i++) m[i] = 0;
    m[0] = m[5] = m[10] = 1;
}

void mat_mul(const double a[12], const double b[12], double out[12])
{
    double r[12];
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 4; j++) {
            double v = a[i * 4 + 0] * b[0 * 4 + j] + a[i * 4 + 1] * b[1 * 4 + j] + a[i * 4 + 2] * b[2 * 4 + j];
            if (j == 3) v += a[i * 4 + 3];
            r[i * 4 + j] = v;
        }
    }
    std::memcpy(out, r, sizeof r);
}

bool mat_inverse(const double m[12], double out[12])
{
    /* inverse of an affine 3x4: [R t]^-1 = [R^-1, -R^-1 t] */
    double a = m[0], b = m[1], c = m[2], d = m[4], e = m[5], f = m[6], g = m[8], h = m[9], i = m[10];
    double A = e * i - f * h, B = -(d * i - f * g), C = d * h - e * g;
    double det = a * A + b * B + c * C;
    if (det == 0.0 || !std::isfinite(det)) return false;
    double id = 1.0 / det;
    double r[12];
    r[0] = A * id;
    r[1] = -(b * i - c * h) * id;
    r[2] = (b * f - c * e) * id;
    r[4] = B * id;
    r[5] = (a * i - c * g) * id;
    r[6] = -(a * f - c * d) * id;
    r[8] = C * id;
    r[9] = -(a * h - b * g) * id;
    r[10] = (a * e - b * d) * id;
    for (int row = 0; row < 3; row++)
        r[row * 4 + 3] = -(r[row * 4 + 0] * m[3] + r[row * 4 + 1] * m[7] + r[row * 4 + 2] * m[11]);
    std::memcpy(out, r, sizeof r);
    return true;
}

void mat_to_float(const double m[12], float out[12])
{
    for (int i = 0; i < 12; i++) out[i] = (float)m[i];
}

sray_camera_desc default_camera()
{
    sray_camera_desc c;
    std::memset(&c, 0, sizeof c);
    double m[12];
    mat_identity(m);
    mat_to_float(m, c.xform);
    c.vfov = M_PI / 4.0; /* QUARTER_PI */
    c.max_dist = (float)SRAY_RAY_MAGNITUDE;
    c.shutter = 0;
    return c;
}

const sray_scene_desc *HostScene::finish()
{
    std::memset(&desc, 0, sizeof desc);
    desc.num_objects = (uint32_t)objects.size();
    desc.objects = objects.data();
    desc.num_meshes = (uint32_t)meshes.size();
    desc.meshes = meshes.data();
    desc.num_triangles = (uint32_t)(tri_pos.size() / 9);
    desc.tri_pos = tri_pos.data();
    desc.tri_normal = has_normal ? tri_normal.data() : nullptr;
    desc.tri_tangent = has_tangent ? tri_tangent.data() : nullptr;
    desc.tri_texcoord = has_texcoord ? tri_texcoord.data() : nullptr;
    desc.num_materials = (uint32_t)materials.size();
    desc.materials = materials.data();
    desc.num_lights = (uint32_t)lights.size();
    desc.lights = lights.data();
    for (size_t i = 0; i < textures.size() && i < texels.size(); i++) textures[i].texels = texels[i].data();
    desc.num_textures = (uint32_t)textures.size();
    desc.textures = textures.empty() ? nullptr : textures.data();
    desc.camera = camera;
    std::memcpy(desc.env_ambient, env_ambient, sizeof env_ambient);
    std::memcpy(desc.env_color, env_color, sizeof env_color);
    desc.env_ior = env_ior;
    desc.scene_oct_depth = scene_oct_depth;
    desc.scene_oct_items = scene_oct_items;
    desc.mesh_oct_depth = mesh_oct_depth;
    desc.mesh_oct_items = mesh_oct_items;
    desc.frame_t0 = frame_t0;
    desc.frame_t1 = frame_t1;
    for (size_t i = 0; i < objects.size(); i++)
        objects[i].time_xforms = i < obj_tab.size() && !obj_tab[i].empty() ? obj_tab[i].data() : nullptr;
    for (size_t i = 0; i < lights.size(); i++)
        lights[i].time_pos = i < light_tab.size() && !light_tab[i].empty() ? light_tab[i].data() : nullptr;
    camera.time_xforms = cam_tab.empty() ? nullptr : cam_tab.data();
    desc.camera = camera;
    return &desc;
}

} // namespace sray
