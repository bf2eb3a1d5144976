/*
 * photon_host.cc — the host half of photon shooting: projection maps
 * (ProjMap, src/projmap.cc; Light::build_projmap / build_proj_intern,
 * src/light.cc:151-205, 392-432) and the per-light emission plan of
 * Scene::build_caustics_map / build_global_map (src/scene.cc:445-549).
 *
 * Projection maps are 64x32 cells over the sphere of directions from the
 * light; a cell is full when a corner direction, scaled to the distance of
 * an object's bounding sphere, lands inside that sphere, or when it holds
 * the direction to the sphere's centre.  They are small (2048 cells x
 * objects x light samples) and built once per shoot, in double precision,
 * so they stay on the host; the device only reads the finished cells.
 *
 * Differences from the reference, both deliberate:
 *   - cell_from_dir counts boundary cosines instead of calling atan2/acos
 *     (the cell differs only for a direction exactly on a cell border);
 *   - light sample points come from the counter-based generator of rng.h.
 */
#include <cfloat>
#include <cmath>
#include <cstring>

#include "host.h"
#include "rng.h"

namespace sray {

namespace {

struct CellTables {
    float cv[SRAY_PROJ_V + 1]; /* cos(k*pi/32) */
    float cu[SRAY_PROJ_U + 1]; /* cos(2*pi*k/64) */
    double corner[SRAY_PROJ_V + 1][SRAY_PROJ_U + 1][3]; /* grid directions (cell corners) */
    double area[SRAY_PROJ_U * SRAY_PROJ_V];              /* solid-angle proxy per cell */
    double total = 0;
    CellTables()
    {
        for (int k = 0; k <= SRAY_PROJ_V; k++) cv[k] = (float)std::cos((double)k * M_PI / SRAY_PROJ_V);
        for (int k = 0; k <= SRAY_PROJ_U; k++) cu[k] = (float)std::cos(2.0 * M_PI * (double)k / SRAY_PROJ_U);
        /* ProjMap::cell_corner_dir, src/projmap.cc:129-144 */
        for (int j = 0; j <= SRAY_PROJ_V; j++)
            for (int i = 0; i <= SRAY_PROJ_U; i++) {
                double u = (double)i * (1.0 / SRAY_PROJ_U);
                double v = (double)j * (1.0 / SRAY_PROJ_V);
                double th = u * M_PI * 2.0, ph = v * M_PI;
                corner[j][i][0] = std::sin(ph) * std::cos(th);
                corner[j][i][1] = std::cos(ph);
                corner[j][i][2] = std::sin(ph) * std::sin(th);
            }
        /* ProjMap::set_subdiv, src/projmap.cc:44-74: cell area = the two
         * triangles of its corner quad */
        for (int c = 0; c < SRAY_PROJ_U * SRAY_PROJ_V; c++) {
            int uc = c % SRAY_PROJ_U, vc = c / SRAY_PROJ_U;
            const double *q[4] = {corner[vc][uc], corner[vc][uc + 1], corner[vc + 1][uc + 1], corner[vc + 1][uc]};
            area[c] = tri_area(q[0], q[1], q[2]) + tri_area(q[0], q[2], q[3]);
            total += area[c];
        }
    }
    static double tri_area(const double *a, const double *b, const double *c)
    {
        double e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]}, e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
        double cx = e1[1] * e2[2] - e1[2] * e2[1], cy = e1[2] * e2[0] - e1[0] * e2[2], cz = e1[0] * e2[1] - e1[1] * e2[0];
        return std::sqrt(cx * cx + cy * cy + cz * cz) * 0.5;
    }
};

const CellTables &tables()
{
    static const CellTables t;
    return t;
}

/* the coverage of a map follows its cells: every change adds or removes the
 * cell's share of the total area (ProjMap::set_cell, src/projmap.cc:84-102) */
void set_cell(ProjMap &m, int c, bool val)
{
    const CellTables &T = tables();
    if ((m.cells[c] != 0) != val) m.coverage += val ? T.area[c] / T.total : -T.area[c] / T.total;
    m.cells[c] = val ? 1 : 0;
}

/* ::build_projmap, src/light.cc:392-432: a light inside a box fills the map
 * (and ends the build); otherwise each box's bounding sphere marks the cells
 * it covers as seen from the light */
void add_boxes(ProjMap &map, const double lpos[3], const std::vector<double> &boxes)
{
    const CellTables &T = tables();
    for (size_t o = 0; o + 6 <= boxes.size(); o += 6) {
        const double *b = &boxes[o];
        if (lpos[0] >= b[0] && lpos[1] >= b[1] && lpos[2] >= b[2] && lpos[0] < b[3] && lpos[1] < b[4] && lpos[2] < b[5]) {
            memset(map.cells, 1, sizeof map.cells);
            map.coverage = 1.0;
            return;
        }
        double c[3] = {(b[0] + b[3]) / 2.0, (b[1] + b[4]) / 2.0, (b[2] + b[5]) / 2.0};
        double r2 = (b[0] - c[0]) * (b[0] - c[0]) + (b[1] - c[1]) * (b[1] - c[1]) + (b[2] - c[2]) * (b[2] - c[2]);
        double v[3] = {c[0] - lpos[0], c[1] - lpos[1], c[2] - lpos[2]};
        double dist = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        int uc, vc;
        cell_from_dir((float)v[0], (float)v[1], (float)v[2], uc, vc);
        set_cell(map, vc * SRAY_PROJ_U + uc, true);
        static const int offs[4][2] = {{0, 0}, {1, 0}, {1, 1}, {0, 1}};
        for (int j = 0; j < SRAY_PROJ_U * SRAY_PROJ_V; j++) {
            int u = j % SRAY_PROJ_U, w = j / SRAY_PROJ_U;
            for (int k = 0; k < 4; k++) {
                const double *d = T.corner[w + offs[k][1]][u + offs[k][0]];
                double p[3] = {lpos[0] + d[0] * dist - c[0], lpos[1] + d[1] * dist - c[1],
                               lpos[2] + d[2] * dist - c[2]};
                if (p[0] * p[0] + p[1] * p[1] + p[2] * p[2] <= r2) {
                    set_cell(map, j, true);
                    break;
                }
            }
        }
    }
}

inline void xf_pointf(const float *m, float x, float y, float z, float out[3])
{
    out[0] = m[0] * x + m[1] * y + m[2] * z + m[3];
    out[1] = m[4] * x + m[5] * y + m[6] * z + m[7];
    out[2] = m[8] * x + m[9] * y + m[10] * z + m[11];
}

/* translation of an object's matrix at a time (linear keys, src/anim.inl:188-261) */
void translation_at(const PhotonObj &o, int time, float t[3])
{
    t[0] = o.xform[3], t[1] = o.xform[7], t[2] = o.xform[11];
    if (!o.moving) return;
    float s;
    if (time <= o.mt0) s = 0.0f;
    else if (time >= o.mt1) s = 1.0f;
    else s = (float)(time - o.mt0) / (float)(o.mt1 - o.mt0);
    t[0] = t[0] + o.motion[0] * s;
    t[1] = t[1] + o.motion[1] * s;
    t[2] = t[2] + o.motion[2] * s;
}

} // namespace

/* ProjMap::cell_from_dir (src/projmap.cc:146-163) without atan2/acos: the
 * polar cell counts the cos(k*pi/32) boundaries at or above y/|d|, the
 * azimuth cell the cos(2*pi*k/64) boundaries on the z half plane's side */
void cell_from_dir(float x, float y, float z, int &uc, int &vc)
{
    const CellTables &T = tables();
    float r = std::sqrt(x * x + y * y + z * z);
    float cy = y / r;
    int v = 0;
    for (int k = 1; k < SRAY_PROJ_V; k++) v += (cy <= T.cv[k]) ? 1 : 0;
    float rho = std::sqrt(x * x + z * z);
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
    uc = u < SRAY_PROJ_U - 1 ? u : SRAY_PROJ_U - 1;
    vc = v < SRAY_PROJ_V - 1 ? v : SRAY_PROJ_V - 1;
}

void cell_tables(float cv[SRAY_PROJ_V + 1], float cu[SRAY_PROJ_U + 1])
{
    memcpy(cv, tables().cv, sizeof(float) * (SRAY_PROJ_V + 1));
    memcpy(cu, tables().cu, sizeof(float) * (SRAY_PROJ_U + 1));
}

void photon_scene_from_desc(const sray_scene_desc *s, PhotonScene &ps)
{
    ps.objs.assign(s->num_objects, PhotonObj{});
    for (uint32_t i = 0; i < s->num_objects; i++) {
        const sray_object_desc &d = s->objects[i];
        PhotonObj &o = ps.objs[i];
        o.type = d.type;
        o.rad = d.radius;
        o.start = d.start;
        o.end = d.end;
        memcpy(o.xform, d.xform, sizeof o.xform);
        memcpy(o.motion, d.motion, sizeof o.motion);
        o.mt0 = d.motion_t0;
        o.mt1 = d.motion_t1;
        o.moving = (d.motion[0] != 0 || d.motion[1] != 0 || d.motion[2] != 0) && d.motion_t1 > d.motion_t0;
        if (d.time_xforms) {
            const int n = s->frame_t1 - s->frame_t0 + 1;
            o.box_rows.resize((size_t)n * 6);
            o.row_t0 = s->frame_t0;
            for (int t = 0; t < n; t++) object_box_at(s, d, s->frame_t0 + t, &o.box_rows[(size_t)t * 6]);
        }
        /* caustic casters: Material::get_value falls back to 1 for a missing
         * reflect/refract attribute (src/material.cc:164-168) */
        o.specular = false;
        if (d.material >= 0 && (uint32_t)d.material < s->num_materials) {
            const sray_material_desc &m = s->materials[d.material];
            float refl = (m.flags & SRAY_MAT_HAS_REFLECT) ? m.reflect : 1.0f;
            float refr = (m.flags & SRAY_MAT_HAS_REFRACT) ? m.refract : 1.0f;
            o.specular = refl > 0.0f || refr > 0.0f;
        }
        if (d.type == SRAY_OBJ_MESH) {
            /* vertex bounds of the linear part of the matrix: a point is
             * fl(L + t) with L the linear part and t the translation, and
             * rounding is monotonic, so the box at any time is
             * fl(min L + t) .. fl(max L + t) (Mesh::calc_bounds, src/mesh.cc:169-188) */
            float m0[12];
            memcpy(m0, d.xform, sizeof m0);
            m0[3] = m0[7] = m0[11] = 0.0f;
            for (int k = 0; k < 3; k++) {
                o.lin[k] = FLT_MAX;
                o.lin[3 + k] = -FLT_MAX;
            }
            const sray_mesh_desc &md = s->meshes[d.mesh];
            for (uint32_t f = 0; f < md.tri_count; f++) {
                const float *v = s->tri_pos + (size_t)(md.tri_start + f) * 9;
                for (int c = 0; c < 3; c++) {
                    float p[3];
                    xf_pointf(m0, v[c * 3], v[c * 3 + 1], v[c * 3 + 2], p);
                    for (int k = 0; k < 3; k++) {
                        if (p[k] < o.lin[k]) o.lin[k] = p[k];
                        if (p[k] > o.lin[3 + k]) o.lin[3 + k] = p[k];
                    }
                }
            }
        }
    }
    ps.lights.assign(s->lights, s->lights + s->num_lights);
    ps.frame_t0 = s->frame_t0;
    ps.nrow = s->frame_t1 - s->frame_t0 + 1;
    ps.light_rows.assign(s->num_lights, std::vector<float>());
    for (uint32_t i = 0; i < s->num_lights; i++) {
        if (s->lights[i].time_pos && s->lights[i].type != SRAY_LIGHT_BOX)
            ps.light_rows[i].assign(s->lights[i].time_pos, s->lights[i].time_pos + (size_t)ps.nrow * 3);
        ps.lights[i].time_pos = nullptr;
    }
}

void PhotonScene::light_pos(int li, int time, float out[3]) const
{
    const std::vector<float> &r = light_rows[(size_t)li];
    if (r.empty()) {
        for (int k = 0; k < 3; k++) out[k] = lights[(size_t)li].pos[k];
        return;
    }
    int k = time - frame_t0;
    k = k < 0 ? 0 : k >= nrow ? nrow - 1 : k;
    for (int a = 0; a < 3; a++) out[a] = r[(size_t)k * 3 + a];
}

/* Object::calc_bounds at a time (sphere src/sphere.cc:27-34, cylinder
 * src/cylinder.cc:63-86, mesh src/mesh.cc:169-188) */
void photon_object_box(const PhotonObj &o, int time, float b[6])
{
    if (!o.box_rows.empty()) {
        /* keyframed: the bounds at that msec (the nearest row outside) */
        const int n = (int)(o.box_rows.size() / 6);
        int k = time - o.row_t0;
        k = k < 0 ? 0 : k >= n ? n - 1 : k;
        memcpy(b, &o.box_rows[(size_t)k * 6], 6 * sizeof(float));
        return;
    }
    float m[12];
    memcpy(m, o.xform, sizeof m);
    float t[3];
    translation_at(o, time, t);
    m[3] = t[0], m[7] = t[1], m[11] = t[2];
    if (o.type == SRAY_OBJ_SPHERE) {
        float p[3];
        xf_pointf(m, 0.0f, 0.0f, 0.0f, p);
        for (int k = 0; k < 3; k++) {
            b[k] = p[k] - o.rad;
            b[3 + k] = p[k] + o.rad;
        }
    } else if (o.type == SRAY_OBJ_CYLINDER) {
        float s1[3], s2[3];
        xf_pointf(m, 0.0f, o.start, 0.0f, s1);
        xf_pointf(m, 0.0f, o.end, 0.0f, s2);
        for (int k = 0; k < 3; k++) {
            float a0 = s1[k] - o.rad, a1 = s1[k] + o.rad;
            float b0 = s2[k] - o.rad, b1 = s2[k] + o.rad;
            b[k] = a0 < b0 ? a0 : b0;
            b[3 + k] = a1 > b1 ? a1 : b1;
        }
    } else {
        for (int k = 0; k < 3; k++) {
            b[k] = o.lin[k] + t[k];
            b[3 + k] = o.lin[3 + k] + t[k];
        }
    }
}

/* sample point of a light for photon work: point light = its position,
 * sphere light = position + radius * uniform unit vector, box light = the
 * box matrix applied to a uniform point of the box (Light::get_point,
 * src/light.cc:106-109, 272-275, 346-354) */
void photon_light_point(const PhotonScene &ps, int li, int time, uint64_t seed, uint32_t a, uint32_t b, uint32_t c,
                        uint32_t purpose, float out[3])
{
    using namespace srd;
    const sray_light_desc &l = ps.lights[(size_t)li];
    if (l.type == SRAY_LIGHT_BOX) {
        float f[3];
        for (int k = 0; k < 3; k++) f[k] = prand(seed, a, b, c, purpose, 0, (uint32_t)k);
        xf_pointf(l.xform, f[0] * l.dim[0] - 0.5f * l.dim[0], f[1] * l.dim[1] - 0.5f * l.dim[1],
                  f[2] * l.dim[2] - 0.5f * l.dim[2], out);
        return;
    }
    float pos[3];
    ps.light_pos(li, time, pos);
    if (l.type == SRAY_LIGHT_SPHERE) {
        for (uint32_t at = 0; at < PHOTON_SPHERE_ATTEMPTS; at++) {
            float d[3];
            if (marsaglia(prand(seed, a, b, c, purpose, at, 0), prand(seed, a, b, c, purpose, at, 1), d)) {
                for (int k = 0; k < 3; k++) out[k] = pos[k] + d[k] * l.radius;
                return;
            }
        }
    }
    for (int k = 0; k < 3; k++) out[k] = pos[k];
}

/* Light::build_proj_intern (src/light.cc:151-205): the maps of one light
 * over the frame interval's end points, 8 light samples for a sphere light */
void build_projmaps(const PhotonScene &ps, int light, int t0, int t1, uint64_t seed, ProjMap &all, ProjMap &spec)
{
    memset(&all, 0, sizeof all);
    memset(&spec, 0, sizeof spec);
    const sray_light_desc &lt = ps.lights[(size_t)light];
    int tm[2] = {t0, t1};
    int ntm = t0 == t1 ? 1 : 2;
    int samples = lt.type == SRAY_LIGHT_SPHERE ? 8 : 1;
    std::vector<double> boxes, caust;
    for (int ti = 0; ti < ntm; ti++) {
        boxes.clear();
        caust.clear();
        for (const PhotonObj &o : ps.objs) {
            float b[6];
            photon_object_box(o, tm[ti], b);
            for (int k = 0; k < 6; k++) boxes.push_back((double)b[k]);
            if (o.specular)
                for (int k = 0; k < 6; k++) caust.push_back((double)b[k]);
        }
        for (int sm = 0; sm < samples; sm++) {
            float lpf[3];
            photon_light_point(ps, light, tm[ti], seed, 0xffffffffu, (uint32_t)light, (uint32_t)(ti << 8 | sm), srd::P_PROJ,
                               lpf);
            double lp[3] = {(double)lpf[0], (double)lpf[1], (double)lpf[2]};
            add_boxes(all, lp, boxes);
            add_boxes(spec, lp, caust);
        }
    }
    all.coverage = std::min(std::max(all.coverage, 0.0), 1.0);
    spec.coverage = std::min(std::max(spec.coverage, 0.0), 1.0);
}

/* photons per light: num * intensity share, rounded up
 * (Scene::build_caustics_map, src/scene.cc:445-500) */
void photon_counts(const PhotonScene &ps, int num, std::vector<uint32_t> &nphot)
{
    nphot.assign(ps.lights.size(), 0);
    if (num <= 0) return;
    double acc = 0;
    std::vector<double> inten(ps.lights.size());
    for (size_t i = 0; i < ps.lights.size(); i++) {
        const float *c = ps.lights[i].color;
        inten[i] = ((double)c[0] + (double)c[1] + (double)c[2]) / 3.0;
        acc += inten[i];
    }
    if (!(acc > 0.0)) return;
    for (size_t i = 0; i < ps.lights.size(); i++) {
        double n = std::ceil((double)num * (inten[i] / acc));
        nphot[i] = n > 0.0 ? (uint32_t)n : 0u;
    }
}

} // namespace sray
