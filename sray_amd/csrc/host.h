/*
 * host.h — host-side C++ of the S-ray MI355X path (internal).
 *
 * What the reference keeps on the host stays here, in C++: the scene XML
 * loader (Scene::load, src/scene.cc:75-183), the octree build of both levels
 * (src/octree.inl:100-191, src/scene.cc:353-390, src/mesh.cc:210-228) and its
 * flattening into the device node arrays, plus the synthetic scene
 * generators used by tests and the benchmark.
 */
#ifndef SRAY_HOST_H_
#define SRAY_HOST_H_

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/sray_gpu.h"

namespace sray {

/* ------------------------------------------------------------------ */
/* keyframe tracks of an XFormNode (anim_host.cc)                        */
/* ------------------------------------------------------------------ */
enum { SR_INTERP_STEP, SR_INTERP_LINEAR, SR_INTERP_CUBIC };
enum { SR_EXTRAP_CLAMP, SR_EXTRAP_REPEAT, SR_EXTRAP_PINGPONG };
struct TrackKey3 {
    int time;
    double v[3];
};
struct TrackKeyQ {
    int time;
    double q[4];     /* s, x, y, z (normalised, or from an axis-angle) */
    float raw[4];    /* the XML values, for writing the scene back */
    int axisrot;     /* raw is "angle(deg) axis" rather than "s x y z" */
};
struct XFormTracks {
    std::vector<TrackKey3> pos, scale;
    std::vector<TrackKeyQ> rot;
    double pivot[3] = {0, 0, 0};
    int interp = SR_INTERP_LINEAR, extrap = SR_EXTRAP_CLAMP;
    void set_pos(int t, double x, double y, double z);
    void set_scale(int t, double x, double y, double z);
    void set_rot(int t, const double q[4], const float raw[4], int axisrot);
    bool animated() const { return pos.size() > 1 || rot.size() > 1 || scale.size() > 1; }
    /* two position keys, linear, clamped, no rotation or scale keys beyond
     * one, no pivot: a translation the device interpolates itself */
    bool linear_translation() const;
    /* Track::operator() of a vector track (false: no keys, def returned) */
    bool vec(const std::vector<TrackKey3> &tr, int time, const double def[3], double out[3]) const;
    void position(int time, double out[3]) const;
    void scaling(int time, double out[3]) const;
    void rotation(int time, double q[4]) const;
    /* XFormNode::get_xform_matrix: pivot * T * R * S * pivot^-1 */
    void matrix(int time, double m[12]) const;
};
void quat_from_axis_angle(double ax, double ay, double az, double angle, double q[4]);
void quat_to_matrix(const double q[4], double m[12]);

/* ------------------------------------------------------------------ */
/* an owning flat scene: the vectors behind an sray_scene_desc           */
/* ------------------------------------------------------------------ */
struct HostScene {
    std::vector<sray_object_desc> objects;
    std::vector<sray_mesh_desc> meshes;
    std::vector<float> tri_pos, tri_normal, tri_tangent, tri_texcoord;
    bool has_normal = false, has_tangent = false, has_texcoord = false;
    std::vector<sray_material_desc> materials;
    std::vector<std::string> material_names;
    /* textures of the "map" attributes: descriptors, their texels (RGBA,
     * top row first), the file each came from and the name the scene
     * gave it */
    std::vector<sray_texture_desc> textures;
    std::vector<std::vector<float>> texels;
    std::vector<std::string> texture_paths, texture_names;
    std::vector<sray_light_desc> lights;
    sray_camera_desc camera{};
    bool cam_has_target = false; /* target camera: position + target point */
    float cam_pos[3] = {0, 0, 0}, cam_target[3] = {0, 0, 0};
    float env_ambient[4] = {0, 0, 0, 0};
    float env_color[4] = {0, 0, 0, 0};
    float env_ior = 1.0f;
    int scene_oct_depth = 5, scene_oct_items = 5;
    int mesh_oct_depth = 7, mesh_oct_items = 14;
    int frame_t0 = 0, frame_t1 = 0;
    sray_scene_desc desc{};

    /* keyframe tracks of a scene loaded from XML (one per object / light;
     * the camera's, and a target camera's target) and the frame they were
     * last evaluated at */
    bool has_tracks = false;
    std::vector<XFormTracks> obj_tracks, light_tracks;
    XFormTracks cam_tracks, cam_target_tracks;
    int frame_msec = 0;
    /* per-millisecond rows over [frame_t0, frame_t1] of the animated ones:
     * object xform + inverse (24 floats), light position (3), camera
     * xform (12); empty = not animated */
    std::vector<std::vector<float>> obj_tab, light_tab;
    std::vector<float> cam_tab;

    /* refresh desc's pointers/counts from the vectors */
    const sray_scene_desc *finish();
    /* evaluates the tracks for render(msec) (src/render.cc:34-55): the
     * frame interval msec - shutter/2 .. + shutter, the matrices at msec
     * and the rows of the animated transforms over the interval */
    bool set_frame(int msec, std::string &err);
};

/* TargetCamera::get_matrix (src/camera.cc:216-234), double precision */
void target_camera_matrix(const double pos[3], const double target[3], double roll, double m[12]);

/* default camera: Camera() (src/camera.cc:29-35) with the identity matrix */
sray_camera_desc default_camera();

/* row-major 3x4 helpers (double precision, host) */
void mat_identity(double m[12]);
void mat_mul(const double a[12], const double b[12], double out[12]);
bool mat_inverse(const double m[12], double out[12]);
void mat_to_float(const double m[12], float out[12]);

/* ------------------------------------------------------------------ */
/* octree: build (reference rules) + flatten (device layout)             */
/* ------------------------------------------------------------------ */
/* Device node layout, two float4 arrays ("lo" and "hi"):
 *   lo = (min.x, min.y, min.z, bits(first))   hi = (max.x, max.y, max.z, bits(meta))
 *   inner node: first = index of an 8-slot child block, meta = child mask
 *               (bit k = the child whose octant bits (x | y<<1 | z<<2) equal
 *               k is stored, at slot first + k; other slots are unused)
 *   leaf:       first = first item index, meta = SR_LEAF | item count
 * Item ranges are assigned in the reference's depth-first order (children
 * 0..7 of src/octree.inl:161-169), so an item's position in the item array
 * is its occurrence rank — the tie-break key of src/octree.inl:211-234.
 * Empty leaves are not stored. */
constexpr uint32_t SR_LEAF = 0x80000000u;

struct FlatTree {
    std::vector<float> lo, hi; /* 4 floats per node */
    std::vector<uint32_t> items;
    uint32_t height = 0;
    uint32_t num_leaves = 0;
    bool overflow = false; /* the build passed 2^23 nodes or 2^28 item references */
    float root_min[3] = {0, 0, 0}, root_max[3] = {0, 0, 0};
};

/* Builds an octree over item boxes (6 floats per item: min xyz, max xyz)
 * with the reference rules and appends its flattened form to out (node
 * indices offset by out's current node count, item values = item_base +
 * local item index).  Returns the root node index; sets out.overflow (and
 * appends nothing) when the tree would pass 2^23 nodes or 2^28 item
 * references. */
uint32_t build_flat_octree(const std::vector<float> &boxes, int max_items, int max_depth, uint32_t item_base,
                           FlatTree &out);

/* item boxes in float, same arithmetic as the device:
 * scene level = union of the object's bounds at t0 and t1 */
void object_boxes(const sray_scene_desc *s, std::vector<float> &boxes);
/* one object's bounds at a time (Object::calc_bounds) */
void object_box_at(const sray_scene_desc *s, const sray_object_desc &o, int time, float b[6]);
void triangle_boxes(const sray_scene_desc *s, uint32_t mesh, std::vector<float> &boxes);

/* ------------------------------------------------------------------ */
/* photon shooting, host half (photon_host.cc)                           */
/* ------------------------------------------------------------------ */
struct PhotonObj {
    int32_t type = 0;
    float rad = 0, start = 0, end = 0;
    float xform[12] = {};
    float motion[3] = {};
    int32_t mt0 = 0, mt1 = 0;
    bool moving = false;
    bool specular = false; /* caustic caster (reflect or refract attribute > 0) */
    float lin[6] = {};     /* mesh: vertex bounds under the matrix without translation */
    /* keyframed object: its bounds (6 floats) per msec from the scene's
     * frame_t0 (empty otherwise) */
    std::vector<float> box_rows;
    int row_t0 = 0;
};
/* what the projection maps need of a scene, kept after upload */
struct PhotonScene {
    std::vector<PhotonObj> objs;
    std::vector<sray_light_desc> lights; /* time_pos cleared: see light_rows */
    std::vector<std::vector<float>> light_rows; /* a light's position per msec, or empty */
    int frame_t0 = 0, nrow = 0;
    /* Light::get_position(msec) of light li */
    void light_pos(int li, int time, float out[3]) const;
};
struct ProjMap {
    uint8_t cells[SRAY_PROJ_U * SRAY_PROJ_V]; /* v-major */
    double coverage;
};
void photon_scene_from_desc(const sray_scene_desc *s, PhotonScene &ps);
void photon_object_box(const PhotonObj &o, int time, float b[6]);
void cell_from_dir(float x, float y, float z, int &uc, int &vc);
void cell_tables(float cv[SRAY_PROJ_V + 1], float cu[SRAY_PROJ_U + 1]);
void photon_light_point(const PhotonScene &ps, int li, int time, uint64_t seed, uint32_t a, uint32_t b, uint32_t c,
                        uint32_t purpose, float out[3]);
/* all = map of every object, spec = map of the caustic casters */
void build_projmaps(const PhotonScene &ps, int light, int t0, int t1, uint64_t seed, ProjMap &all, ProjMap &spec);
void photon_counts(const PhotonScene &ps, int num, std::vector<uint32_t> &nphot);

/* ------------------------------------------------------------------ */
/* XML                                                                  */
/* ------------------------------------------------------------------ */
/* attribute typing of src/xmltree.c:51-99 */
enum { ATYPE_STR, ATYPE_INT, ATYPE_FLT, ATYPE_VEC };
struct XmlAttr {
    std::string name, str;
    int type = ATYPE_STR;
    int ival = 0;
    float fval = 0;
    float vval[4] = {0, 0, 0, 0};
};
struct XmlNode {
    std::string name;
    std::vector<XmlAttr> attr;
    std::vector<XmlNode> chld;
    const XmlAttr *get(const char *n) const;
};
bool xml_parse(const std::string &text, XmlNode &root, std::string &err);
XmlAttr xml_make_attr(const std::string &name, const std::string &val);

bool load_scene_xml(const std::string &text, HostScene &hs, std::string &err, const std::string &scene_dir = "");
/* Image::load (src/img.cc:64-78): a PNG, PPM/PGM or PFM file as RGBA
 * floats, row-major, top row first */
bool read_image_rgbaf(const std::string &path, int &w, int &h, std::vector<float> &px, std::string &err);
std::string write_scene_xml(const HostScene &hs);

/* ------------------------------------------------------------------ */
/* synthetic workloads (seeded)                                         */
/* ------------------------------------------------------------------ */
bool generate_scene(int config, uint64_t seed, int variant, HostScene &hs, std::string &err);

} // namespace sray

#endif
