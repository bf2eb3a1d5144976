/*
 * runtime.h — host-side runtime shared by the translation units of
 * libsray_gpu.so: device buffers, resident scenes and photon maps, the
 * context (one HIP stream, replacing the reference's pthread tile pool,
 * src/tpool.cc:48-225), error reporting and per-kernel HIP-event timing.
 */
#ifndef SRAY_RUNTIME_H_
#define SRAY_RUNTIME_H_

#include <hip/hip_runtime.h>
#include <hipcub/device/device_select.hpp>

#include <atomic>
#include <condition_variable>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

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
    srt::DevBuf cubtmp, nsel, rays, hitbuf, counts, gq, grgb, gcnt, gneed, gfq, goff, gcand, giota;
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
    sray_tuning tune{128, 64, 16, 21, 0.5f, 1, 0, 4, 64, 1};
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

/* stable selection of the flagged elements (order kept), count to the host */
template <class T>
int select_flagged(sray_gpu_ctx *c, const T *in, const uint8_t *flags, T *out, uint32_t n, uint32_t *count)
{
    size_t tmp = 0;
    HIPCHK(c, hipcub::DeviceSelect::Flagged(nullptr, tmp, in, flags, out, (uint32_t *)c->nsel.p, (int)n, c->stream));
    int rc = ensure(c, c->cubtmp, tmp);
    if (rc) return rc;
    tmp = c->cubtmp.bytes;
    HIPCHK(c, hipcub::DeviceSelect::Flagged(c->cubtmp.p, tmp, in, flags, out, (uint32_t *)c->nsel.p, (int)n, c->stream));
    HIPCHK(c, hipMemcpyAsync(count, c->nsel.p, 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return SRAY_OK;
}

/* as select_flagged, with the count left on the device (no synchronisation) */
template <typename T>
int select_flagged_async(sray_gpu_ctx *c, const T *in, const uint8_t *flags, T *out, uint32_t n, uint32_t *d_count)
{
    size_t tmp = 0;
    HIPCHK(c, hipcub::DeviceSelect::Flagged(nullptr, tmp, in, flags, out, d_count, (int)n, c->stream));
    int rc = ensure(c, c->cubtmp, tmp);
    if (rc) return rc;
    tmp = c->cubtmp.bytes;
    HIPCHK(c, hipcub::DeviceSelect::Flagged(c->cubtmp.p, tmp, in, flags, out, d_count, (int)n, c->stream));
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
