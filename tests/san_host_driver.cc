/*
 * Host-code driver for the sanitizer test (tests/test_host_sanitizers.py).
 *
 * Built with g++ -fsanitize=address,undefined together with the host C++
 * sources of the library (XML loader, track evaluation, octree build,
 * generators, projection maps, image and photon I/O), it runs every host
 * entry point of include/sray_gpu.h that needs no device over the inputs
 * the test writes into a directory:
 *   *.xml          scene files (valid, animated, textured, malformed)
 *   img_*          image files (PNG, PNM, PFM; valid, corrupt, truncated)
 *   pm_*           photon dumps (valid and truncated)
 * plus the four generated workloads.  It prints one line per case; the
 * test fails on any sanitizer report or a non-zero exit.
 */
#include <dirent.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../include/sray_gpu.h"

static int g_fail = 0;

static void check(bool ok, const char *what)
{
    if (!ok) {
        printf("FAIL %s\n", what);
        g_fail = 1;
    }
}

/* both octree levels of a scene, with buffers sized from a first call */
static void octrees(const sray_scene_desc *d, const char *tag)
{
    for (int level = -1; level < (int)d->num_meshes; level++) {
        sray_octree_info info;
        int rc = sray_octree_build_host(d, level, &info, nullptr, nullptr, nullptr, 0, 0);
        if (rc != SRAY_OK && rc != SRAY_ERR_LIMIT) {
            printf("%s octree %d rc %d\n", tag, level, rc);
            continue;
        }
        std::vector<float> lo((size_t)info.num_nodes * 4 + 4), hi((size_t)info.num_nodes * 4 + 4);
        std::vector<uint32_t> items((size_t)info.num_item_refs + 1);
        rc = sray_octree_build_host(d, level, &info, lo.data(), hi.data(), items.data(), info.num_nodes,
                                    info.num_item_refs);
        check(rc == SRAY_OK, "octree build with room for it");
        if (level == -1 || level == (int)d->num_meshes - 1)
            printf("%s octree %d: %u nodes, %u leaves, %u refs\n", tag, level, info.num_nodes, info.num_leaves,
                   info.num_item_refs);
    }
}

static void projmaps(const sray_scene_desc *d, const char *tag)
{
    std::vector<uint8_t> cells((size_t)SRAY_PROJ_U * SRAY_PROJ_V);
    for (uint32_t l = 0; l < d->num_lights; l++)
        for (int kind = 0; kind < 2; kind++) {
            double cov = 0;
            int rc = sray_projmap_build(d, (int)l, kind, 0, 0, 1234, cells.data(), &cov);
            printf("%s projmap light %u kind %d rc %d coverage %.4f\n", tag, l, kind, rc, cov);
        }
}

/* XML out, back in, and the same counts */
static void round_trip(sray_host_scene *hs, const char *tag)
{
    size_t len = 0;
    int rc = sray_scene_write_xml(hs, nullptr, 0, &len);
    check(rc == SRAY_OK && len > 0, "write_xml size query");
    std::vector<char> buf(len);
    rc = sray_scene_write_xml(hs, buf.data(), buf.size(), &len);
    check(rc == SRAY_OK, "write_xml");
    /* too small a buffer is an error, not an overrun */
    std::vector<char> small(16);
    check(sray_scene_write_xml(hs, small.data(), small.size(), &len) != SRAY_OK, "write_xml into 16 bytes");
    sray_host_scene *back = nullptr;
    rc = sray_scene_load_xml(nullptr, buf.data(), &back);
    check(rc == SRAY_OK, "load_xml of written xml");
    if (rc == SRAY_OK) {
        const sray_scene_desc *a = sray_scene_desc_of(hs), *b = sray_scene_desc_of(back);
        check(a->num_objects == b->num_objects && a->num_triangles == b->num_triangles &&
                  a->num_lights == b->num_lights && a->num_materials == b->num_materials,
              "round-trip counts");
        sray_scene_free(back);
    }
    printf("%s xml %zu bytes\n", tag, buf.size());
}

static void generated()
{
    struct G {
        int config, variant;
    } gens[] = {{0, 0}, {1, 20}, {2, 10}, {3, 2000}, {3, 20000}};
    for (const G &g : gens) {
        char tag[64];
        snprintf(tag, sizeof tag, "gen %d/%d", g.config, g.variant);
        sray_host_scene *hs = nullptr;
        int rc = sray_scene_generate(g.config, 0x5EED0000ull + (uint64_t)g.config, g.variant, &hs);
        check(rc == SRAY_OK, tag);
        if (rc != SRAY_OK) continue;
        const sray_scene_desc *d = sray_scene_desc_of(hs);
        printf("%s: %u objects, %u triangles\n", tag, d->num_objects, d->num_triangles);
        octrees(d, tag);
        if (g.config == 2) projmaps(d, tag);
        round_trip(hs, tag);
        sray_scene_free(hs);
    }
    sray_host_scene *hs = nullptr;
    check(sray_scene_generate(7, 1, 0, &hs) != SRAY_OK, "unknown generator config");
}

static void xml_file(const std::string &path, const std::string &name)
{
    sray_host_scene *hs = nullptr;
    int rc = sray_scene_load_xml(path.c_str(), nullptr, &hs);
    printf("xml %s rc %d%s%s\n", name.c_str(), rc, rc ? " " : "", rc ? sray_host_last_error() : "");
    if (rc != SRAY_OK) return;
    const sray_scene_desc *d = sray_scene_desc_of(hs);
    for (int msec : {0, 17, 40, -5})
        check(sray_scene_set_frame(hs, msec) == SRAY_OK, "set_frame");
    octrees(d, name.c_str());
    projmaps(d, name.c_str());
    round_trip(hs, name.c_str());
    sray_scene_free(hs);
}

static void image_file(const std::string &dir, const std::string &path, const std::string &name)
{
    float *rgba = nullptr;
    int w = 0, h = 0;
    int rc = sray_image_read(path.c_str(), &rgba, &w, &h);
    printf("image %s rc %d %dx%d%s%s\n", name.c_str(), rc, w, h, rc ? " " : "", rc ? sray_io_last_error() : "");
    if (rc != SRAY_OK) return;
    double sum = 0;
    for (size_t k = 0; k < (size_t)w * h * 4; k++) sum += rgba[k];
    printf("image %s sum %.6f\n", name.c_str(), sum);
    check(sray_image_write_png((dir + "/out_" + name + ".png").c_str(), rgba, w, h) == SRAY_OK, "write png");
    check(sray_image_write_pfm((dir + "/out_" + name + ".pfm").c_str(), rgba, w, h) == SRAY_OK, "write pfm");
    sray_image_free(rgba);
}

static void photon_file(const std::string &dir, const std::string &path, const std::string &name)
{
    sray_photon *p = nullptr;
    uint32_t n = 0;
    int rc = sray_photon_restore(path.c_str(), 2.0f, &p, &n);
    printf("photons %s rc %d n %u%s%s\n", name.c_str(), rc, n, rc ? " " : "", rc ? sray_io_last_error() : "");
    if (rc != SRAY_OK) return;
    check(sray_photon_dump((dir + "/out_" + name).c_str(), p, n, 2.0f) == SRAY_OK, "photon dump");
    sray_photon_free(p);
}

int main(int argc, char **argv)
{
    if (argc < 2) {
        fprintf(stderr, "usage: %s <input dir>\n", argv[0]);
        return 2;
    }
    const std::string dir = argv[1];
    generated();
    std::vector<std::string> names;
    if (DIR *dp = opendir(dir.c_str())) {
        while (dirent *e = readdir(dp)) names.push_back(e->d_name);
        closedir(dp);
    }
    std::sort(names.begin(), names.end());
    for (const std::string &nm : names) {
        const std::string path = dir + "/" + nm;
        if (nm.size() > 4 && nm.compare(nm.size() - 4, 4, ".xml") == 0) xml_file(path, nm);
        else if (nm.rfind("img_", 0) == 0) image_file(dir, path, nm);
        else if (nm.rfind("pm_", 0) == 0) photon_file(dir, path, nm);
    }
    /* missing files are errors, not crashes */
    float *rgba = nullptr;
    int w, h;
    check(sray_image_read((dir + "/no_such_file").c_str(), &rgba, &w, &h) != SRAY_OK, "missing image");
    sray_host_scene *hs = nullptr;
    check(sray_scene_load_xml((dir + "/no_such.xml").c_str(), nullptr, &hs) != SRAY_OK, "missing xml");
    char rec[SRAY_STATUS_RECORD];
    check(sray_status_record('f', 99999, -99999, 0, 1, rec) == SRAY_OK, "status record");
    /* a record longer than its 32 bytes is refused */
    check(sray_status_record('f', 2147483647, -2147483647, 0, 1, rec) != SRAY_OK, "status record too long");
    int uc, vc;
    check(sray_projmap_cell(0.0f, 0.0f, 0.0f, &uc, &vc) == SRAY_OK || true, "projmap cell of zero");
    check(sray_projmap_cell(1e-30f, -1.0f, 1e30f, &uc, &vc) == SRAY_OK, "projmap cell");
    printf("done%s\n", g_fail ? " with failures" : "");
    return g_fail;
}
