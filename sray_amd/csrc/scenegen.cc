/*
 * scenegen.cc — seeded synthetic workloads in the reference's scene model.
 *
 * Every generator fills a flat scene whose numbers are all float and which
 * write_scene_xml() prints as reference-format XML with %.9g, so the same
 * workload exists as a scene file for the reference's loader.
 *
 *   config 0  spheres:   64 Phong spheres, one point light, 640x480 depth 5
 *   config 1  heightfield: 1000x500-quad fBm terrain (1M triangles) + 20
 *             glass + 20 mirror spheres, 1920x1080 depth 6, mesh octree 8/32
 *   config 2  photon room: 12 axis-aligned wall quads, 2 glass spheres, a
 *             displaced icosphere, one square box light
 *   config 3  batch scene: displaced icosphere of n triangles + random
 *             spheres (variant = triangle count), 10% moving objects
 *   variant (configs 0-2): scale factor for reduced sizes in tests
 *             (config 1: quads along x = 1000 / variant when variant > 1)
 */
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>

#include "host.h"

namespace sray {

namespace {

struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed) {}
    uint64_t next()
    {
        /* SplitMix64 sequence */
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double unit() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    float uni(double a, double b) { return (float)(a + (b - a) * unit()); }
};

uint64_t hash3(uint64_t seed, int64_t x, int64_t y, int64_t z)
{
    uint64_t h = seed;
    auto mix = [](uint64_t v) {
        v += 0x9E3779B97F4A7C15ull;
        v = (v ^ (v >> 30)) * 0xBF58476D1CE4E5B9ull;
        v = (v ^ (v >> 27)) * 0x94D049BB133111EBull;
        return v ^ (v >> 31);
    };
    h = mix(h ^ (uint64_t)x);
    h = mix(h ^ (uint64_t)y);
    h = mix(h ^ (uint64_t)z);
    return h;
}

/* lattice value in [-1,1] */
double lattice(uint64_t seed, int64_t x, int64_t y, int64_t z)
{
    return (double)(hash3(seed, x, y, z) >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

double smooth(double t) { return t * t * (3.0 - 2.0 * t); }

/* 2D value noise */
double vnoise2(uint64_t seed, double x, double z)
{
    double fx = floor(x), fz = floor(z);
    int64_t ix = (int64_t)fx, iz = (int64_t)fz;
    double tx = smooth(x - fx), tz = smooth(z - fz);
    double a = lattice(seed, ix, 0, iz), b = lattice(seed, ix + 1, 0, iz);
    double c = lattice(seed, ix, 0, iz + 1), d = lattice(seed, ix + 1, 0, iz + 1);
    double ab = a + (b - a) * tx, cd = c + (d - c) * tx;
    return ab + (cd - ab) * tz;
}

/* 3D value noise */
double vnoise3(uint64_t seed, double x, double y, double z)
{
    double fx = floor(x), fy = floor(y), fz = floor(z);
    int64_t ix = (int64_t)fx, iy = (int64_t)fy, iz = (int64_t)fz;
    double tx = smooth(x - fx), ty = smooth(y - fy), tz = smooth(z - fz);
    double v[2][2][2];
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++)
            for (int c = 0; c < 2; c++) v[a][b][c] = lattice(seed, ix + a, iy + b, iz + c);
    double x00 = v[0][0][0] + (v[1][0][0] - v[0][0][0]) * tx, x10 = v[0][1][0] + (v[1][1][0] - v[0][1][0]) * tx;
    double x01 = v[0][0][1] + (v[1][0][1] - v[0][0][1]) * tx, x11 = v[0][1][1] + (v[1][1][1] - v[0][1][1]) * tx;
    double y0 = x00 + (x10 - x00) * ty, y1 = x01 + (x11 - x01) * ty;
    return y0 + (y1 - y0) * tz;
}

/* fBm: 4 octaves, lacunarity 2, gain 0.5, normalised to [-1,1] */
double fbm2(uint64_t seed, double x, double z, int oct)
{
    double sum = 0, amp = 1, freq = 1, norm = 0;
    for (int o = 0; o < oct; o++) {
        sum += amp * vnoise2(seed + (uint64_t)o * 7919, x * freq, z * freq);
        norm += amp;
        amp *= 0.5;
        freq *= 2.0;
    }
    return sum / norm;
}

double fbm3(uint64_t seed, double x, double y, double z, int oct)
{
    double sum = 0, amp = 1, freq = 1, norm = 0;
    for (int o = 0; o < oct; o++) {
        sum += amp * vnoise3(seed + (uint64_t)o * 104729, x * freq, y * freq, z * freq);
        norm += amp;
        amp *= 0.5;
        freq *= 2.0;
    }
    return sum / norm;
}

int add_material(HostScene &hs, const std::string &name, const float kd[3], float ks, float shin, float refl,
                 float refr, float ior)
{
    sray_material_desc m;
    std::memset(&m, 0, sizeof m);
    m.diffuse[0] = kd[0];
    m.diffuse[1] = kd[1];
    m.diffuse[2] = kd[2];
    m.diffuse[3] = 1.0f;
    m.specular[0] = m.specular[1] = m.specular[2] = ks;
    m.specular[3] = 1.0f;
    m.shininess = shin;
    m.reflect = refl;
    m.refract = refr;
    m.ior = ior;
    m.flags = SRAY_MAT_HAS_SPECULAR | SRAY_MAT_HAS_REFLECT | SRAY_MAT_HAS_REFRACT | SRAY_MAT_PHONG;
    hs.materials.push_back(m);
    hs.material_names.push_back(name);
    return (int)hs.materials.size() - 1;
}

void set_translation(sray_object_desc &o, float x, float y, float z)
{
    double m[12], inv[12];
    mat_identity(m);
    m[3] = x;
    m[7] = y;
    m[11] = z;
    mat_inverse(m, inv);
    mat_to_float(m, o.xform);
    mat_to_float(inv, o.inv_xform);
}

void add_sphere(HostScene &hs, int mat, float x, float y, float z, float r)
{
    sray_object_desc o;
    std::memset(&o, 0, sizeof o);
    o.type = SRAY_OBJ_SPHERE;
    o.material = mat;
    o.mesh = -1;
    o.radius = r;
    o.end = 1.0f;
    set_translation(o, x, y, z);
    hs.objects.push_back(o);
}

int add_mesh_object(HostScene &hs, int mat, uint32_t tri_start, uint32_t tri_count, float tx, float ty, float tz)
{
    sray_mesh_desc md;
    md.tri_start = tri_start;
    md.tri_count = tri_count;
    hs.meshes.push_back(md);
    sray_object_desc o;
    std::memset(&o, 0, sizeof o);
    o.type = SRAY_OBJ_MESH;
    o.material = mat;
    o.mesh = (int32_t)hs.meshes.size() - 1;
    o.radius = 1.0f;
    o.end = 1.0f;
    set_translation(o, tx, ty, tz);
    hs.objects.push_back(o);
    return (int)hs.objects.size() - 1;
}

void add_point_light(HostScene &hs, float x, float y, float z, float c)
{
    sray_light_desc l;
    std::memset(&l, 0, sizeof l);
    l.type = SRAY_LIGHT_POINT;
    l.pos[0] = x;
    l.pos[1] = y;
    l.pos[2] = z;
    l.color[0] = l.color[1] = l.color[2] = c;
    l.color[3] = 1.0f;
    l.radius = 1.0f;
    l.dim[0] = l.dim[1] = l.dim[2] = 1.0f;
    double m[12];
    mat_identity(m);
    m[3] = x;
    m[7] = y;
    m[11] = z;
    mat_to_float(m, l.xform);
    hs.lights.push_back(l);
}

void set_free_camera(HostScene &hs, float x, float y, float z, double fov_deg)
{
    hs.camera = default_camera();
    hs.camera.vfov = (double)(float)fov_deg * M_PI / 180.0;
    hs.camera.xform[3] = x;
    hs.camera.xform[7] = y;
    hs.camera.xform[11] = z;
    hs.cam_has_target = false;
}

void set_target_camera(HostScene &hs, const float p[3], const float t[3], double fov_deg)
{
    hs.camera = default_camera();
    hs.camera.vfov = (double)(float)fov_deg * M_PI / 180.0;
    double P[3] = {p[0], p[1], p[2]}, T[3] = {t[0], t[1], t[2]}, m[12];
    target_camera_matrix(P, T, 0.0, m);
    mat_to_float(m, hs.camera.xform);
    hs.cam_has_target = true;
    for (int k = 0; k < 3; k++) {
        hs.cam_pos[k] = p[k];
        hs.cam_target[k] = t[k];
    }
}

void push_tri(HostScene &hs, const float *a, const float *b, const float *c)
{
    for (int k = 0; k < 3; k++) hs.tri_pos.push_back(a[k]);
    for (int k = 0; k < 3; k++) hs.tri_pos.push_back(b[k]);
    for (int k = 0; k < 3; k++) hs.tri_pos.push_back(c[k]);
}

/* displaced class-I geodesic sphere of frequency f (20*f^2 triangles), unit
 * radius displaced by amp * fBm(direction), in local space */
uint32_t add_icosphere(HostScene &hs, int f, double amp, uint64_t seed, double scale)
{
    const double t = (1.0 + sqrt(5.0)) / 2.0;
    double V[12][3] = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
                       {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
    static const int F[20][3] = {{0, 11, 5}, {0, 5, 1}, {0, 1, 7}, {0, 7, 10}, {0, 10, 11}, {1, 5, 9}, {5, 11, 4},
                                 {11, 10, 2}, {10, 7, 6}, {7, 1, 8}, {3, 9, 4}, {3, 4, 2}, {3, 2, 6}, {3, 6, 8},
                                 {3, 8, 9}, {4, 9, 5}, {2, 4, 11}, {6, 2, 10}, {8, 6, 7}, {9, 8, 1}};
    uint32_t start = (uint32_t)(hs.tri_pos.size() / 9);
    /* canonical point on face (A,B,C) at barycentric (i,j,k)/f: points on an
     * edge are computed from the edge's endpoints in index order, so faces
     * sharing an edge produce bit-identical vertices (a closed mesh) */
    auto point = [&](const int *fc, int i, int j, int k, float out[3]) {
        int idx[3] = {fc[0], fc[1], fc[2]};
        int w[3] = {i, j, k};
        /* sort corners by vertex index so the combination is canonical */
        for (int a = 0; a < 2; a++)
            for (int b = 0; b < 2 - a; b++)
                if (idx[b] > idx[b + 1]) {
                    std::swap(idx[b], idx[b + 1]);
                    std::swap(w[b], w[b + 1]);
                }
        double p[3] = {0, 0, 0};
        for (int c = 0; c < 3; c++)
            for (int d = 0; d < 3; d++) p[d] += V[idx[c]][d] * (double)w[c];
        double l = sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
        double dx = p[0] / l, dy = p[1] / l, dz = p[2] / l;
        double r = scale * (1.0 + amp * fbm3(seed, dx * 2.0 + 11.0, dy * 2.0 + 17.0, dz * 2.0 + 23.0, 4));
        out[0] = (float)(dx * r);
        out[1] = (float)(dy * r);
        out[2] = (float)(dz * r);
    };
    for (int fi = 0; fi < 20; fi++) {
        const int *fc = F[fi];
        for (int i = 0; i < f; i++) {
            for (int j = 0; j < f - i; j++) {
                /* point (a,b): weights (f-a-b, a, b) over corners (A,B,C) */
                float p00[3], p10[3], p01[3], p11[3];
                point(fc, f - i - j, i, j, p00);
                point(fc, f - i - 1 - j, i + 1, j, p10);
                point(fc, f - i - j - 1, i, j + 1, p01);
                push_tri(hs, p00, p10, p01);
                if (j < f - i - 1) {
                    point(fc, f - i - 1 - j - 1, i + 1, j + 1, p11);
                    push_tri(hs, p10, p11, p01);
                }
            }
        }
    }
    return (uint32_t)(hs.tri_pos.size() / 9) - start;
}

bool gen_spheres(uint64_t seed, HostScene &hs)
{
    Rng r(seed);
    hs.env_ambient[0] = hs.env_ambient[1] = hs.env_ambient[2] = 0.05f;
    hs.env_ambient[3] = 1.0f;
    /* exactly 16 of 64 spheres are reflective (seeded choice) */
    int idx[64];
    for (int i = 0; i < 64; i++) idx[i] = i;
    for (int i = 63; i > 0; i--) {
        int j = (int)(r.next() % (uint64_t)(i + 1));
        std::swap(idx[i], idx[j]);
    }
    bool refl[64] = {false};
    for (int i = 0; i < 16; i++) refl[idx[i]] = true;
    for (int i = 0; i < 64; i++) {
        float kd[3] = {r.uni(0, 1), r.uni(0, 1), r.uni(0, 1)};
        int m = add_material(hs, "sph" + std::to_string(i), kd, 0.5f, 40.0f, refl[i] ? 0.3f : 0.0f, 0.0f, 1.5f);
        float x = r.uni(-10, 10), y = r.uni(-10, 10), z = r.uni(-10, 10);
        float rad = r.uni(0.3, 1.5);
        add_sphere(hs, m, x, y, z, rad);
    }
    add_point_light(hs, 0.0f, 20.0f, 0.0f, 1.0f);
    set_free_camera(hs, 0.0f, 5.0f, -25.0f, 45.0);
    return true;
}

bool gen_heightfield(uint64_t seed, int variant, HostScene &hs)
{
    Rng r(seed);
    int div = variant > 1 ? variant : 1;
    int nx = 1000 / div, nz = 500 / div;
    if (nx < 2) nx = 2;
    if (nz < 1) nz = 1;
    const double x0 = -50, x1 = 50, z0 = -25, z1 = 25, amp = 2.0;
    uint64_t nseed = r.next();
    std::vector<float> h((size_t)(nx + 1) * (nz + 1));
    auto height = [&](double x, double z) { return amp * fbm2(nseed, x * 0.08, z * 0.08, 4); };
    for (int j = 0; j <= nz; j++)
        for (int i = 0; i <= nx; i++) {
            double x = x0 + (x1 - x0) * i / nx, z = z0 + (z1 - z0) * j / nz;
            h[(size_t)j * (nx + 1) + i] = (float)height(x, z);
        }
    hs.env_ambient[0] = hs.env_ambient[1] = hs.env_ambient[2] = 0.05f;
    hs.env_ambient[3] = 1.0f;
    hs.env_color[0] = 0.25f;
    hs.env_color[1] = 0.35f;
    hs.env_color[2] = 0.5f;
    hs.env_color[3] = 1.0f;
    float tkd[3] = {0.55f, 0.5f, 0.4f};
    int tm = add_material(hs, "terrain", tkd, 0.1f, 10.0f, 0.0f, 0.0f, 1.5f);
    float gkd[3] = {0.05f, 0.05f, 0.05f};
    int gm = add_material(hs, "glass", gkd, 1.0f, 60.0f, 0.1f, 0.9f, 1.5f);
    float mkd[3] = {0.05f, 0.05f, 0.05f};
    int mm = add_material(hs, "mirror", mkd, 1.0f, 60.0f, 1.0f, 0.0f, 1.5f);
    uint32_t start = 0;
    for (int j = 0; j < nz; j++)
        for (int i = 0; i < nx; i++) {
            auto P = [&](int a, int b, float out[3]) {
                out[0] = (float)(x0 + (x1 - x0) * a / nx);
                out[1] = h[(size_t)b * (nx + 1) + a];
                out[2] = (float)(z0 + (z1 - z0) * b / nz);
            };
            float v00[3], v10[3], v01[3], v11[3];
            P(i, j, v00);
            P(i + 1, j, v10);
            P(i, j + 1, v01);
            P(i + 1, j + 1, v11);
            push_tri(hs, v00, v01, v11);
            push_tri(hs, v00, v11, v10);
        }
    add_mesh_object(hs, tm, start, (uint32_t)(hs.tri_pos.size() / 9), 0.0f, 0.0f, 0.0f);
    for (int s = 0; s < 40; s++) {
        float rad = r.uni(0.5, 2.0);
        float x = r.uni(-45, 45), z = r.uni(-22, 22);
        float y = (float)(height(x, z) + rad + r.uni(0.5, 3.0));
        add_sphere(hs, s < 20 ? gm : mm, x, y, z, rad);
    }
    add_point_light(hs, 10.0f, 40.0f, -20.0f, 1.0f);
    float cp[3] = {0.0f, 18.0f, -55.0f}, ct[3] = {0.0f, 0.0f, 0.0f};
    set_target_camera(hs, cp, ct, 45.0);
    hs.mesh_oct_depth = 8;
    hs.mesh_oct_items = 32;
    return true;
}

/* quad as a 2-triangle mesh object */
void add_quad(HostScene &hs, int mat, const float a[3], const float b[3], const float c[3], const float d[3])
{
    uint32_t start = (uint32_t)(hs.tri_pos.size() / 9);
    push_tri(hs, a, b, c);
    push_tri(hs, a, c, d);
    add_mesh_object(hs, mat, start, 2, 0.0f, 0.0f, 0.0f);
}

bool gen_room(uint64_t seed, int variant, HostScene &hs)
{
    Rng r(seed);
    hs.env_ambient[0] = hs.env_ambient[1] = hs.env_ambient[2] = 0.02f;
    hs.env_ambient[3] = 1.0f;
    float grey[3] = {0.7f, 0.7f, 0.7f}, red[3] = {0.7f, 0.1f, 0.1f}, green[3] = {0.1f, 0.7f, 0.1f};
    int mg = add_material(hs, "grey", grey, 0.0f, 1.0f, 0.0f, 0.0f, 1.5f);
    int mr = add_material(hs, "red", red, 0.0f, 1.0f, 0.0f, 0.0f, 1.5f);
    int mgr = add_material(hs, "green", green, 0.0f, 1.0f, 0.0f, 0.0f, 1.5f);
    float gkd[3] = {0.0f, 0.0f, 0.0f};
    int mglass = add_material(hs, "glass", gkd, 1.0f, 80.0f, 0.05f, 0.95f, 1.5f);
    float ikd[3] = {0.6f, 0.55f, 0.45f};
    int mico = add_material(hs, "stone", ikd, 0.2f, 20.0f, 0.0f, 0.0f, 1.5f);
    /* room [-5,5] x [0,8] x [-5,5]; each of the 6 walls is 2 quads (12) */
    const float X0 = -5, X1 = 5, Y0 = 0, Y1 = 8, Z0 = -5, Z1 = 5;
    auto P = [](float x, float y, float z, float o[3]) {
        o[0] = x;
        o[1] = y;
        o[2] = z;
    };
    float a[3], b[3], c[3], d[3];
    for (int k = 0; k < 2; k++) {
        float zl = k == 0 ? Z0 : 0.0f, zh = k == 0 ? 0.0f : Z1;
        /* floor, ceiling */
        P(X0, Y0, zl, a), P(X1, Y0, zl, b), P(X1, Y0, zh, c), P(X0, Y0, zh, d);
        add_quad(hs, mg, a, b, c, d);
        P(X0, Y1, zl, a), P(X0, Y1, zh, b), P(X1, Y1, zh, c), P(X1, Y1, zl, d);
        add_quad(hs, mg, a, b, c, d);
        /* left (red) and right (green) walls */
        P(X0, Y0, zl, a), P(X0, Y0, zh, b), P(X0, Y1, zh, c), P(X0, Y1, zl, d);
        add_quad(hs, mr, a, b, c, d);
        P(X1, Y0, zl, a), P(X1, Y1, zl, b), P(X1, Y1, zh, c), P(X1, Y0, zh, d);
        add_quad(hs, mgr, a, b, c, d);
    }
    for (int k = 0; k < 2; k++) {
        float xl = k == 0 ? X0 : 0.0f, xh = k == 0 ? 0.0f : X1;
        /* back wall z = Z1, front wall z = Z0 (behind the camera) */
        P(xl, Y0, Z1, a), P(xh, Y0, Z1, b), P(xh, Y1, Z1, c), P(xl, Y1, Z1, d);
        add_quad(hs, mg, a, b, c, d);
        P(xl, Y0, Z0, a), P(xl, Y1, Z0, b), P(xh, Y1, Z0, c), P(xh, Y0, Z0, d);
        add_quad(hs, mg, a, b, c, d);
    }
    add_sphere(hs, mglass, -2.0f, 1.0f, 1.5f, 1.0f);
    add_sphere(hs, mglass, 1.8f, 1.0f, -0.5f, 1.0f);
    int freq = variant > 1 ? std::max(2, 50 / variant) : 50;
    uint32_t start = (uint32_t)(hs.tri_pos.size() / 9);
    uint32_t n = add_icosphere(hs, freq, 0.15, r.next(), 1.2);
    add_mesh_object(hs, mico, start, n, 0.5f, 1.6f, 3.0f);
    /* square area light: a box light with a flat box under the ceiling */
    sray_light_desc l;
    std::memset(&l, 0, sizeof l);
    l.type = SRAY_LIGHT_BOX;
    l.pos[0] = 0.0f;
    l.pos[1] = 7.9f;
    l.pos[2] = 0.0f;
    l.color[0] = l.color[1] = l.color[2] = 1.0f;
    l.color[3] = 1.0f;
    l.radius = 1.0f;
    l.dim[0] = 2.0f;
    l.dim[1] = 0.0f;
    l.dim[2] = 2.0f;
    double m[12];
    mat_identity(m);
    m[7] = 7.9;
    mat_to_float(m, l.xform);
    hs.lights.push_back(l);
    float cp[3] = {0.0f, 4.0f, -4.8f}, ct[3] = {0.0f, 3.0f, 0.0f};
    set_target_camera(hs, cp, ct, 60.0);
    return true;
}

/* one scene of the batch workload: displaced icosphere with about ntri
 * triangles plus 8-32 spheres; 10% of objects move linearly (<= 0.5) */
bool gen_batch_scene(uint64_t seed, int ntri, HostScene &hs)
{
    Rng r(seed);
    hs.env_ambient[0] = hs.env_ambient[1] = hs.env_ambient[2] = 0.05f;
    hs.env_ambient[3] = 1.0f;
    hs.env_color[0] = 0.2f;
    hs.env_color[1] = 0.2f;
    hs.env_color[2] = 0.25f;
    hs.env_color[3] = 1.0f;
    int freq = std::max(1, (int)lround(sqrt((double)std::max(ntri, 20) / 20.0)));
    float ikd[3] = {r.uni(0.3, 0.9), r.uni(0.3, 0.9), r.uni(0.3, 0.9)};
    int mi = add_material(hs, "ico", ikd, 0.3f, 30.0f, 0.0f, 0.0f, 1.5f);
    uint32_t start = (uint32_t)(hs.tri_pos.size() / 9);
    uint32_t n = add_icosphere(hs, freq, 0.2, r.next(), 2.0);
    add_mesh_object(hs, mi, start, n, 0.0f, 0.0f, 0.0f);
    /* vertex normals (radial) and spherical texture coordinates, so the
     * batch exercises interpolated shading normals (interp_face) */
    hs.has_normal = hs.has_texcoord = true;
    hs.tri_normal.resize(hs.tri_pos.size());
    hs.tri_texcoord.resize(hs.tri_pos.size() / 9 * 6);
    for (size_t v = 0; v < hs.tri_pos.size() / 3; v++) {
        const float *p = &hs.tri_pos[v * 3];
        double l = sqrt((double)p[0] * p[0] + (double)p[1] * p[1] + (double)p[2] * p[2]);
        for (int k = 0; k < 3; k++) hs.tri_normal[v * 3 + k] = (float)(p[k] / l);
        hs.tri_texcoord[v * 2] = (float)(atan2((double)p[2], (double)p[0]) / (2.0 * M_PI) + 0.5);
        hs.tri_texcoord[v * 2 + 1] = (float)(acos(std::max(-1.0, std::min(1.0, p[1] / l))) / M_PI);
    }
    int ns = 8 + (int)(r.next() % 25);
    for (int s = 0; s < ns; s++) {
        float kd[3] = {r.uni(0, 1), r.uni(0, 1), r.uni(0, 1)};
        bool mirror = (r.next() % 4) == 0;
        int m = add_material(hs, "s" + std::to_string(s), kd, 0.5f, 40.0f, mirror ? 0.5f : 0.0f, 0.0f, 1.5f);
        float x, y, z;
        do {
            x = r.uni(-6, 6);
            y = r.uni(-6, 6);
            z = r.uni(-6, 6);
        } while (x * x + y * y + z * z < 9.0f);
        add_sphere(hs, m, x, y, z, r.uni(0.3, 1.2));
    }
    for (sray_object_desc &o : hs.objects) {
        if (r.unit() < 0.1) {
            double dx = r.unit() - 0.5, dy = r.unit() - 0.5, dz = r.unit() - 0.5;
            double l = sqrt(dx * dx + dy * dy + dz * dz);
            double len = 0.5 * r.unit();
            /* end position rounded to float first, so the scene file
             * (two position keys) reproduces the motion exactly */
            double m[3] = {dx / l * len, dy / l * len, dz / l * len};
            const float p0[3] = {o.xform[3], o.xform[7], o.xform[11]};
            for (int k = 0; k < 3; k++) {
                float p1 = (float)((double)p0[k] + m[k]);
                o.motion[k] = (float)((double)p1 - (double)p0[k]);
            }
            o.motion_t0 = 0;
            o.motion_t1 = 40;
        }
    }
    hs.frame_t0 = 0;
    hs.frame_t1 = 40;
    add_point_light(hs, 5.0f, 15.0f, -10.0f, 1.0f);
    float cp[3] = {0.0f, 3.0f, -16.0f}, ct[3] = {0.0f, 0.0f, 0.0f};
    set_target_camera(hs, cp, ct, 50.0);
    return true;
}

} // namespace

bool generate_scene(int config, uint64_t seed, int variant, HostScene &hs, std::string &err)
{
    hs = HostScene();
    hs.camera = default_camera();
    bool ok;
    switch (config) {
    case 0:
        ok = gen_spheres(seed, hs);
        break;
    case 1:
        ok = gen_heightfield(seed, variant, hs);
        break;
    case 2:
        ok = gen_room(seed, variant, hs);
        break;
    case 3:
        ok = gen_batch_scene(seed, variant > 0 ? variant : 1000, hs);
        break;
    default:
        err = "unknown config";
        return false;
    }
    hs.finish();
    return ok;
}

} // namespace sray
