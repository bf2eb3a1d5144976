/*
 * capi_host.cc — host-only part of the C ABI (no device needed): scene XML
 * loading (Scene::load, src/scene.cc:75-183), octree inspection and the
 * synthetic workload generators.
 */
#include <cstdio>
#include <cstring>
#include <new>
#include <fstream>
#include <sstream>
#include <string>

#include "host.h"

struct sray_host_scene {
    sray::HostScene hs;
};

static thread_local std::string g_host_err;

extern "C" {

const char *sray_host_last_error(void) { return g_host_err.c_str(); }

int sray_scene_load_xml(const char *path, const char *xml_text, sray_host_scene **out)
{
    if (!out || (!path && !xml_text)) {
        g_host_err = "sray_scene_load_xml: need a path or XML text";
        return SRAY_ERR_ARG;
    }
    *out = nullptr;
    std::string text;
    if (xml_text) {
        text = xml_text;
    } else {
        std::ifstream f(path, std::ios::binary);
        if (!f) {
            g_host_err = std::string("failed to open file: ") + path;
            return SRAY_ERR_PARSE;
        }
        std::ostringstream ss;
        ss << f.rdbuf();
        text = ss.str();
    }
    /* texture maps are searched for in the scene file's directory first */
    std::string dir;
    if (path) {
        const char *slash = strrchr(path, '/');
        if (slash) dir.assign(path, (size_t)(slash - path));
        if (slash == path) dir = "/";
    }
    auto *h = new sray_host_scene;
    std::string err;
    if (!sray::load_scene_xml(text, h->hs, err, dir)) {
        g_host_err = "failed to load scene: " + err;
        delete h;
        return SRAY_ERR_PARSE;
    }
    *out = h;
    return SRAY_OK;
}

const sray_scene_desc *sray_scene_desc_of(const sray_host_scene *h) { return h ? &h->hs.desc : nullptr; }

int sray_scene_set_frame(sray_host_scene *h, int32_t msec)
{
    if (!h) return SRAY_ERR_ARG;
    std::string err;
    if (!h->hs.set_frame(msec, err)) {
        g_host_err = err;
        return SRAY_ERR_ARG;
    }
    return SRAY_OK;
}

void sray_scene_free(sray_host_scene *h) { delete h; }

int sray_scene_generate(int config, uint64_t seed, int variant, sray_host_scene **out)
{
    if (!out) return SRAY_ERR_ARG;
    *out = nullptr;
    auto *h = new sray_host_scene;
    std::string err;
    if (!sray::generate_scene(config, seed, variant, h->hs, err)) {
        g_host_err = "generate: " + err;
        delete h;
        return SRAY_ERR_ARG;
    }
    *out = h;
    return SRAY_OK;
}

int sray_scene_write_xml(const sray_host_scene *h, char *buf, size_t cap, size_t *len)
{
    if (!h || !len) return SRAY_ERR_ARG;
    std::string s = sray::write_scene_xml(h->hs);
    *len = s.size() + 1;
    if (buf) {
        if (cap < s.size() + 1) return SRAY_ERR_ARG;
        memcpy(buf, s.c_str(), s.size() + 1);
    }
    return SRAY_OK;
}

static int octree_build_host(const sray_scene_desc *s, int level, sray_octree_info *info, float *node_lo,
                             float *node_hi, uint32_t *items, uint32_t cap_nodes, uint32_t cap_items);

int sray_octree_build_host(const sray_scene_desc *s, int level, sray_octree_info *info, float *node_lo,
                           float *node_hi, uint32_t *items, uint32_t cap_nodes, uint32_t cap_items)
{
    try {
        return octree_build_host(s, level, info, node_lo, node_hi, items, cap_nodes, cap_items);
    } catch (const std::bad_alloc &) {
        return SRAY_ERR_NOMEM;
    }
}

static int octree_build_host(const sray_scene_desc *s, int level, sray_octree_info *info, float *node_lo,
                             float *node_hi, uint32_t *items, uint32_t cap_nodes, uint32_t cap_items)
{
    if (!s || !info) return SRAY_ERR_ARG;
    std::vector<float> boxes;
    sray::FlatTree ft;
    if (level < 0) {
        sray::object_boxes(s, boxes);
        sray::build_flat_octree(boxes, s->scene_oct_items, s->scene_oct_depth, 0, ft);
    } else {
        if ((uint32_t)level >= s->num_meshes) return SRAY_ERR_ARG;
        sray::triangle_boxes(s, (uint32_t)level, boxes);
        sray::build_flat_octree(boxes, s->mesh_oct_items, s->mesh_oct_depth, 0, ft);
    }
    if (ft.overflow) return SRAY_ERR_LIMIT;
    info->num_nodes = (uint32_t)(ft.lo.size() / 4);
    info->num_leaves = ft.num_leaves;
    info->num_item_refs = (uint32_t)ft.items.size();
    info->height = ft.height;
    for (int k = 0; k < 3; k++) {
        info->root_min[k] = ft.root_min[k];
        info->root_max[k] = ft.root_max[k];
    }
    /* (an empty tree has nothing to copy, and memcpy takes no null source) */
    if (node_lo && node_hi && info->num_nodes <= cap_nodes && !ft.lo.empty()) {
        memcpy(node_lo, ft.lo.data(), ft.lo.size() * 4);
        memcpy(node_hi, ft.hi.data(), ft.hi.size() * 4);
    }
    if (items && info->num_item_refs <= cap_items && !ft.items.empty())
        memcpy(items, ft.items.data(), ft.items.size() * 4);
    return SRAY_OK;
}

int sray_projmap_build(const sray_scene_desc *s, int light, int kind, int t0, int t1, uint64_t seed, uint8_t *cells,
                       double *coverage)
{
    if (!s || !cells || !coverage || light < 0 || (uint32_t)light >= s->num_lights || (kind != 0 && kind != 1)) {
        g_host_err = "sray_projmap_build: bad arguments";
        return SRAY_ERR_ARG;
    }
    sray::PhotonScene ps;
    sray::photon_scene_from_desc(s, ps);
    sray::ProjMap all, spec;
    sray::build_projmaps(ps, light, t0, t1, seed, all, spec);
    const sray::ProjMap &m = kind == 0 ? all : spec;
    memcpy(cells, m.cells, sizeof m.cells);
    *coverage = m.coverage;
    return SRAY_OK;
}

int sray_projmap_cell(float x, float y, float z, int *uc, int *vc)
{
    if (!uc || !vc) return SRAY_ERR_ARG;
    sray::cell_from_dir(x, y, z, *uc, *vc);
    return SRAY_OK;
}

} /* extern "C" */

extern "C" int sray_abi_sizes(uint32_t *out, int n)
{
    const uint32_t sz[] = {(uint32_t)sizeof(sray_object_desc),   (uint32_t)sizeof(sray_mesh_desc),
                           (uint32_t)sizeof(sray_material_desc), (uint32_t)sizeof(sray_texture_desc),
                           (uint32_t)sizeof(sray_light_desc),
                           (uint32_t)sizeof(sray_camera_desc),   (uint32_t)sizeof(sray_scene_desc),
                           (uint32_t)sizeof(sray_ray),           (uint32_t)sizeof(sray_hit),
                           (uint32_t)sizeof(sray_trace_counts),  (uint32_t)sizeof(sray_render_params),
                           (uint32_t)sizeof(sray_stats),         (uint32_t)sizeof(sray_photon),
                           (uint32_t)sizeof(sray_gather_query),  (uint32_t)sizeof(sray_request),
                           (uint32_t)sizeof(sray_result),        (uint32_t)sizeof(sray_octree_info),
                           (uint32_t)sizeof(sray_kernel_stat),   (uint32_t)sizeof(sray_photon_params),
                           (uint32_t)sizeof(sray_photon_stats),  (uint32_t)sizeof(sray_tuning)};
    int k = 0;
    for (; k < n && k < (int)(sizeof sz / sizeof sz[0]); k++) out[k] = sz[k];
    return k;
}
