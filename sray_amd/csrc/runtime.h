/*
 * runtime.h — host-side runtime shared by the translation units of
 * libsray_gpu.so: device buffers, resident scenes and photon maps, the
 * context (one HIP stream, replacing the reference's pthread tile pool,
 * src/tpool.cc:48-225), error reporting and per-kernel HIP-event timing.
 */
#ifndef SRAY_RUNTIME_H_
#define SRAY_RUNTIME_H_

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "devsort.h"
#include "host.h"
#include "render.h"

namespace srt {

using namespace srd;

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
};

template <class T> T *dptr(DevBuf &b) { return (T *)b.p; }

struct GpuPmap {
    std::vector<DevBuf> bufs;
    PmapTree tree{};
    sray_photon *store = nullptr; /* photons in storage order */
    uint32_t scene_id = 0;
    int kind = 0;
    uint64_t bytes = 0;
    /* shading-gather grids, one per (gather radius, cell size), built on
     * first use and kept until the map is released */
    struct Grid {
        DevBuf pos, dir, nrm, colb, cell;
        PmapGrid g{};
    };
    std::map<uint64_t, std::unique_ptr<Grid>> grids; /* key: radius bits << 32 | cell bits */
};

struct GpuScene {
    std::vector<DevBuf> bufs;
    SceneDev dev{};
    uint64_t bytes = 0;
    double diag = 0;
    uint32_t scene_height = 0, mesh_height = 0;
    uint32_t nobjs = 0;
    uint32_t pmap[2] = {0, 0}; /* bound caustic / global photon maps */
    sray::PhotonScene photon;  /* what the projection maps need */
    sray_camera_desc camera{}; /* the scene's own camera */
    /* which children its Phong materials can spawn (bounds the growth of
     * the bounce levels): any Phong material, one with reflect > 0, one
     * with refract > 0 */
    bool has_phong = false, pos_reflect = false, pos_refract = false;
    /* some mesh object carries vertex attributes (interpolated at hits) */
    bool attr_meshes = false;
    /* the frame interval of the scene's keyframed rows (timed: some object,
     * light or the camera has them; ray times outside it take the nearest
     * row, so renders and photon shooting must stay inside) */
    int frame_t0 = 0, frame_t1 = 0;
    bool timed = false;
};

struct Engine; /* multi-request wavefront renderer (engine.hip) */

} // namespace srt

struct sray_gpu_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::recursive_mutex mu;
    std::atomic<int> api_waiting{0}; /* callers waiting for mu (served before the next render pass) */
    std::string err;
    std::map<uint32_t, std::unique_ptr<srt::GpuScene>> scenes;
    uint32_t next_scene = 1;
    std::map<uint32_t, std::unique_ptr<srt::GpuPmap>> pmaps;
    uint32_t next_pmap = 1;
    uint64_t admission_cap = 200ull << 30;
    /* trace / gather scratch, grown on demand */
    srt::DevBuf seltmp, sorttmp, nsel, rays, hitbuf, counts, gq, grgb, gcnt, gneed, gfq, goff, gcand, giota;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    /* per-kernel HIP-event timing */
    bool profiling = false;
    std::vector<hipEvent_t> evpool;
    struct Pending {
        const char *name;
        hipEvent_t a, b;
    };
    std::vector<Pending> pending;
    struct KStat {
        uint64_t launches = 0;
        double ms = 0;
    };
    std::map<std::string, KStat> kstats;
    /* tuning of the renderers and photon grids */
    sray_tuning tune{64, 64, 16, 21, 1.0f, 1, 0, 4, 0, 1, 1};
    /* status stream (fd < 0: off) and the lock that keeps records whole */
    std::atomic<int> status_fd{-1};
    std::mutex status_mu;
    /* renderers: one for synchronous renders, one behind submit/poll */
    srt::Engine *plain = nullptr;
    srt::Engine *serve = nullptr;
};

namespace srt {

/* the context lock as taken by API calls: a caller waiting here is served
 * before the serving thread starts its next render pass */
struct ApiLock {
    sray_gpu_ctx *c;
    explicit ApiLock(sray_gpu_ctx *cc) : c(cc)
    {
        c->api_waiting++;
        c->mu.lock();
        c->api_waiting--;
    }
    ~ApiLock() { c->mu.unlock(); }
    ApiLock(const ApiLock &) = delete;
    ApiLock &operator=(const ApiLock &) = delete;
};

int fail(sray_gpu_ctx *c, int code, const char *fmt, ...);

#define HIPCHK(c, x)                                                                                                  \
    do {                                                                                                              \
        hipError_t e_ = (x);                                                                                          \
        if (e_ != hipSuccess) return srt::fail((c), SRAY_ERR_HIP, "%s failed: %s", #x, hipGetErrorString(e_));       \
    } while (0)

/* grows b to at least bytes (contents are not kept) */
int ensure(sray_gpu_ctx *c, DevBuf &b, size_t bytes);
/* as ensure, and a buffer it (re)allocates is cleared on the context's
 * stream (counters that accumulate across calls) */
int ensure_zero(sray_gpu_ctx *c, DevBuf &b, size_t bytes);
void release(DevBuf &b);
inline uint32_t blocks(size_t n, int bs) { return (uint32_t)((n + (size_t)bs - 1) / (size_t)bs); }

hipEvent_t prof_event(sray_gpu_ctx *c);
/* records an event pair around one launch when profiling is on */
struct ProfScope {
    sray_gpu_ctx *c;
    const char *name;
    hipStream_t s;
    hipEvent_t a = nullptr;
    /* on the stream the launch goes to (the context's own by default) */
    ProfScope(sray_gpu_ctx *cc, const char *n, hipStream_t st = nullptr) : c(cc), name(n), s(st ? st : cc->stream)
    {
        if (c->profiling && (a = prof_event(c))) (void)hipEventRecord(a, s);
    }
    ~ProfScope()
    {
        if (!a) return;
        hipEvent_t b = prof_event(c);
        if (!b) {
            c->evpool.push_back(a);
            return;
        }
        (void)hipEventRecord(b, s);
        c->pending.push_back({name, a, b});
    }
};
/* after a stream synchronisation: fold pending event pairs into the stats */
void prof_resolve(sray_gpu_ctx *c);

GpuScene *find_scene(sray_gpu_ctx *c, uint32_t id);

/* writes one status record to the context's status stream, if it has one */
void emit_status(sray_gpu_ctx *c, char st, int a1, int a2, int a3, int a4);

/* stable selection of the flagged elements (order kept), the count left
 * on the device (no synchronisation) */
template <typename T>
int select_flagged_async(sray_gpu_ctx *c, const T *in, const uint8_t *flags, T *out, uint32_t n, uint32_t *d_count)
{
    int rc = ensure(c, c->seltmp, (size_t)CMP_BLOCKS * 4);
    if (rc) return rc;
    k_sel_count<T><<<CMP_BLOCKS, CMP_THREADS, 0, c->stream>>>(flags, n, dptr<uint32_t>(c->seltmp));
    k_sel_scatter<T><<<CMP_BLOCKS, CMP_THREADS, 0, c->stream>>>(in, flags, n, dptr<uint32_t>(c->seltmp), d_count,
                                                                out);
    HIPCHK(c, hipGetLastError());
    return SRAY_OK;
}

/* as select_flagged_async, with the count copied to the host */
template <class T>
int select_flagged(sray_gpu_ctx *c, const T *in, const uint8_t *flags, T *out, uint32_t n, uint32_t *count)
{
    int rc = ensure(c, c->nsel, 4);
    if (rc || (rc = select_flagged_async(c, in, flags, out, n, dptr<uint32_t>(c->nsel)))) return rc;
    HIPCHK(c, hipMemcpyAsync(count, c->nsel.p, 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return SRAY_OK;
}

/* stable LSD radix sort of n (key, value) pairs on bits [0, end_bit) of
 * the keys, from (k0, v0) into (k1, v1); both pairs of buffers are
 * overwritten on the way (an odd number of digit passes of at most 8 bits
 * each ends in k1, v1) */
template <class K>
int radix_sort_pairs(sray_gpu_ctx *c, K *k0, K *k1, uint32_t *v0, uint32_t *v1, uint32_t n, int end_bit)
{
    int rc = ensure(c, c->sorttmp, ((size_t)256 * CMP_BLOCKS + 256) * 4);
    if (rc) return rc;
    if (end_bit < 1) end_bit = 1;
    int passes = (end_bit + 7) / 8;
    if (passes % 2 == 0) passes++;
    const int width = (end_bit + passes - 1) / passes;
    uint32_t *bh = dptr<uint32_t>(c->sorttmp), *dt = bh + (size_t)256 * CMP_BLOCKS;
    K *ks = k0, *kd = k1;
    uint32_t *vs = v0, *vd = v1;
    int shift = 0;
    for (int p = 0; p < passes; p++) {
        const int wdt = std::max(0, std::min(width, end_bit - shift));
        const int sh = wdt ? shift : 0; /* a zero-width pass is a stable copy */
        k_rs_hist<K><<<CMP_BLOCKS, CMP_THREADS, 0, c->stream>>>(ks, n, sh, wdt, bh);
        k_rs_scan<<<1u << wdt, CMP_BLOCKS, 0, c->stream>>>(bh, dt);
        k_rs_scatter<K><<<CMP_BLOCKS, CMP_THREADS, 0, c->stream>>>(ks, vs, n, sh, wdt, bh, dt, kd, vd);
        std::swap(ks, kd);
        std::swap(vs, vd);
        shift += wdt;
    }
    HIPCHK(c, hipGetLastError());
    return SRAY_OK;
}

/* the shading-gather grid of a photon map for a radius and a cell size in
 * radii (built or cached) */
int pmap_grid_get(sray_gpu_ctx *c, uint32_t pmap_id, float radius, float cell, PmapGrid *out);

/* engine.hip */
int do_render(sray_gpu_ctx *c, uint32_t scene_id, const sray_render_params *p, float *d_out, sray_stats *st);
/* median device time of `rounds` renders (after one untimed render) */
int time_render(sray_gpu_ctx *c, uint32_t scene_id, const sray_render_params *p, int rounds, float *d_out, double *ms);
void engine_destroy(sray_gpu_ctx *c);
/* device bytes held by the renderers (queues, tables, pixel arenas) */
uint64_t engine_bytes(sray_gpu_ctx *c);
/* the photon map is about to be released: fails if requests in flight use it */
int engine_pmap_busy(sray_gpu_ctx *c, uint32_t pmap_id);
/* the scene is about to be released: fails if requests in flight use it */
int engine_scene_busy(sray_gpu_ctx *c, uint32_t scene_id);

} // namespace srt

#endif
