"""Benchmark: Mrays/s of the S-ray ray-tracing hot path on one MI355X.

Default run (`python bench.py`): one JSON line.  Its value is the
traversal-bound heightfield frame; a step is one full frame (1000x500-quad
fBm terrain = 1,000,000 triangles + 20 glass + 20 mirror spheres, mesh
octree depth 8 with 32 items per leaf, 1920x1080, 1 sample per pixel,
recursion depth 6: primary, shadow, reflect and refract rays through the
wavefront path; scene and octrees resident in HBM before the timed region).
The line also carries:
  roofline      the closest-hit traversal kernel k_trace_q: algorithmic bytes
                from the library's visit counters over its HIP-event-timed
                launch duration;
  batched       8 heightfield frames of other sample seeds in flight
                together (one submit_many), against the plain frames;
  cpu_baseline  the CPU restatement of the reference renderer (double
                precision, -O3 -ffast-math, reference tiles and mesh
                re-traversal) on whole frames, on the host cores this job
                may use, with the RMSE between its image and the GPU's;
  legs          the other workloads, each with its own GPU-vs-GPU or
                GPU-vs-CPU figure:
                  spheres  64 spheres, 640x480, depth 5: plain, and 64 frames
                           in flight, against the CPU on whole frames;
                  room     photon-mapped room, 1280x720, adaptive 1-16 spp,
                           2M + 500k photons: 8 frames one after another
                           against the same 8 frames in flight;
                  batch    256 scenes: one after another, all in flight, and
                           arriving as a Poisson stream (latency p50/p99);
                  serve    32 resident rooms, a Poisson stream of tile
                           requests at a rate below the measured capacity;
  targets       the target ratios (GPU against CPU, batched against plain,
                latencies and image error), from the legs above.
`--workload NAME` runs one of them alone and prints its own line.
"""
import argparse
import ctypes
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
WORKLOADS = {
    # name: (config, variant, width, height, iter, samples)
    "heightfield": (1, 0, 1920, 1080, 6, (1, 1)),
    "spheres": (0, 0, 640, 480, 5, (1, 1)),
    "room": (2, 0, 1280, 720, 5, (1, 16)),
    "batch": (3, 0, 0, 0, 4, (8, 8)),
    "serve": (2, 0, 0, 0, 5, (1, 16)),
}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="all", choices=["all"] + sorted(WORKLOADS))
    ap.add_argument("--cpu-seconds", type=float, default=10.0, help="CPU baseline sample budget per workload")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default tuning (no untimed search)")
    ap.add_argument("--tuning", default=None,
                    help="set the tuning instead of searching, e.g. "
                         "trace_block=64,tile_w=64,chunk_log2=21,graphs=1,persistent=0")
    ap.add_argument("--octree", default=None,
                    help="heightfield: mesh octree 'items,depth' (default: the configured 32,8)")
    ap.add_argument("--octree-autotune", action="store_true",
                    help="heightfield: search the mesh octree among 32/8, 48/8, 64/7, 64/8 (untimed)")
    ap.add_argument("--frames", type=int, default=None, help="heightfield / room / spheres: frames in flight (8 / 8 / 64)")
    ap.add_argument("--scenes", type=int, default=None, help="batch: scenes (256); serve: resident scenes (32)")
    ap.add_argument("--requests", type=int, default=2000, help="serve: requests in the Poisson stream")
    ap.add_argument("--load", type=float, default=0.7,
                    help="serve / batch: Poisson rate as a fraction of the measured capacity")
    ap.add_argument("--rate", type=float, default=None, help="serve: fixed Poisson rate (req/s) instead")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 256 MiB device to device three times (scales the HBM byte counters)")
    ap.add_argument("--gather-k", type=int, default=200,
                    help="room / serve: photons per gather estimate (0: every photon within the radius)")
    ap.add_argument("--quiet", action="store_true")
    return ap.parse_args()


def log(args, *a):
    if not args.quiet:
        print(*a, file=sys.stderr, flush=True)


def cpu_share():
    """(threads, description): the host cores this job may use.  The
    reference starts one worker per online core (src/tpool.cc:273,
    sysconf(_SC_NPROCESSORS_ONLN)); here that count is capped by the
    process's affinity mask, a cgroup CPU quota and OMP_NUM_THREADS (the
    job's CPU share on a shared host), so the baseline does not count
    threads that would only time-slice the same cores."""
    online = os.cpu_count() or 1
    n, why = online, "online cores"
    aff = len(os.sched_getaffinity(0))
    if aff < n:
        n, why = aff, "affinity mask"
    try:
        quota, period = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if quota != "max":
            q = max(1, int(int(quota) // int(period)))
            if q < n:
                n, why = q, "cgroup CPU quota"
    except (OSError, ValueError):
        pass
    omp = os.environ.get("OMP_NUM_THREADS")
    if omp and omp.isdigit() and 0 < int(omp) < n:
        n, why = int(omp), "OMP_NUM_THREADS (the job's CPU share)"
    return n, {"online_cores": online, "affinity_cores": aff, "threads": n, "limited_by": why}


def read_pmc(workload):
    """the committed rocprofv3 --pmc summary of k_trace_q (key: value lines)"""
    path = os.path.join(ROOT, "profiles", f"pmc_k_trace_q_{workload}.txt")
    vals = {}
    if not os.path.exists(path):
        return vals
    with open(path) as f:
        for line in f:
            if ":" in line and not line.startswith("#"):
                k, v = line.split(":", 1)
                try:
                    vals[k.strip()] = float(v.split()[0])
                except (ValueError, IndexError):
                    pass
    return vals


def read_traffic(workload):
    """HBM bytes per k_trace_q launch from the committed rocprofv3 --pmc summary"""
    return read_pmc(workload).get("hbm_bytes_per_dispatch")


def read_valu(workload):
    """VALU issue busy of k_trace_q from the same summary (all levels, level 0, levels 1-6)"""
    v = read_pmc(workload)
    keys = ("valu_busy_all_levels", "valu_busy_level0", "valu_busy_levels1_6")
    if not all(k in v for k in keys):
        return None
    return {"all_levels": v[keys[0]], "level0": v[keys[1]], "levels1_6": v[keys[2]],
            "source": f"profiles/pmc_k_trace_q_{workload}.txt"}


def rmse(a, b):
    return float(np.sqrt(np.mean((a.astype(np.float64) - b.astype(np.float64)) ** 2)))


def cpu_checker(desc_ptr, photons=None):
    """the CPU restatement at the reference's precision and work pattern
    (double, -O3 -ffast-math build, a mesh traversed again in every scene
    leaf), with the device's photon maps when the workload has them"""
    import oracle
    ref = oracle.OracleScene(desc_ptr, 64, oracle.Q9_RETRAVERSE, fast=True)
    for kind, ph in enumerate(photons or ()):
        if ph is not None and len(ph):
            ref.set_photons(kind, ph)
    return ref


def cpu_frames(args, ref, p, label):
    """whole frames on the CPU until about --cpu-seconds of work: (baseline
    dict, first frame's image)"""
    threads, share = cpu_share()
    frames, rays, elapsed, first = 0, 0, 0.0, None
    while frames == 0 or elapsed < args.cpu_seconds:
        t0 = time.perf_counter()
        img, st = ref.render(p, nthreads=threads)
        elapsed += time.perf_counter() - t0
        rays += st.rays()
        frames += 1
        if first is None:
            first = img
    log(args, f"cpu baseline {label}: {frames} frames, {rays} rays, {elapsed:.1f} s, {threads} threads")
    return ({"value": round(rays / elapsed / 1e6, 4), "unit": "Mrays/s", "cores": threads, "kind": "port",
             "host": share,
             "sample": f"{frames} whole {p.x1 - p.x0}x{p.y1 - p.y0} frames of the {label} workload, {rays} ray "
                       f"queries, {elapsed:.1f} s; double precision, -O3 -ffast-math, 48x48 tiles, a mesh "
                       f"traversed again in every scene leaf as the reference"}, first)


def cpu_rows(args, ref, p, label):
    """bands of 16 rows from the centre out until about --cpu-seconds of work"""
    threads, share = cpu_share()
    W, H = p.x1 - p.x0, p.y1 - p.y0
    rows, rays, elapsed, y0 = 0, 0, 0.0, max(0, H // 2 - 16)
    while rows < H and elapsed < args.cpu_seconds:
        y1 = min(H, y0 + 16)
        q = type(p)()
        ctypes.memmove(ctypes.byref(q), ctypes.byref(p), ctypes.sizeof(p))
        q.y0, q.y1 = y0, y1
        t0 = time.perf_counter()
        _, st = ref.render(q, nthreads=threads)
        elapsed += time.perf_counter() - t0
        rays += st.rays()
        rows += y1 - y0
        y0 = y1 if y1 < H else 0
    log(args, f"cpu baseline {label}: {rows} rows, {rays} rays, {elapsed:.1f} s, {threads} threads")
    return {"value": round(rays / elapsed / 1e6, 4), "unit": "Mrays/s", "cores": threads, "kind": "port",
            "host": share,
            "sample": f"{rows} rows x {W} px of the {W}x{H} {label} frame, {rays} ray queries, {elapsed:.1f} s; "
                      f"double precision, -O3 -ffast-math, the device's photon maps"}


def pct(v, q):
    return round(float(np.percentile(np.asarray(v, np.float64), q)), 3) if len(v) else None


def tuning_dict(t):
    return {"trace_block": t.trace_block, "tile_w": t.tile_w, "chunk_log2": t.chunk_log2, "graphs": t.graphs,
            "persistent": t.persistent, "lanes": t.lanes, "sparse_from": t.sparse_from,
            "sample_runs": t.sample_runs}


def tuning_arg(args, ctx):
    """--tuning applied to the context; the tuning as a dict, or None"""
    if not args.tuning:
        return None
    kw = {k: (float(v) if k == "gather_cell" else int(v))
          for k, v in (item.split("=") for item in args.tuning.split(","))}
    return tuning_dict(ctx.set_tuning(**kw))


def look_at(pos, target, vfov_deg, shutter=0):
    """camera -> world matrix of a camera at pos looking at target (+z forward, +y up)"""
    import sray_amd as S
    f = np.asarray(target, np.float64) - np.asarray(pos, np.float64)
    f /= np.linalg.norm(f)
    r = np.cross([0.0, 1.0, 0.0], f)
    r /= np.linalg.norm(r)
    u = np.cross(f, r)
    cam = S.abi.CameraDesc()
    cam.xform[:] = np.array([[r[0], u[0], f[0], pos[0]], [r[1], u[1], f[1], pos[1]],
                             [r[2], u[2], f[2], pos[2]]], np.float32).ravel()
    cam.vfov = math.radians(vfov_deg)
    cam.max_dist = 10000.0
    cam.shutter = shutter
    return cam


def batched_leg(ctx, jobs, warm=True, timed_rounds=3):
    """all jobs ((scene, params, camera or None)) in flight together: queued
    with one submit_many call and polled, each writing its image to a
    device buffer of its own (as the plain legs do); one untimed round
    first grows the engine's arenas to the batch, then timed_rounds timed
    rounds, of which the median wall time counts.  Returns (Mrays/s, wall
    s, per-request latencies ms of that round)."""
    outs = [ctx.malloc((p.x1 - p.x0) * (p.y1 - p.y0) * 16) for _, p, _ in jobs]
    runs = []
    for k in range((1 if warm else 0) + timed_rounds):
        t0 = time.perf_counter()
        tickets = ctx.submit_many([(sid, p, cam, o) for (sid, p, cam), o in zip(jobs, outs)])
        res = [ctx.poll(t)[1] for t in tickets]
        wall = time.perf_counter() - t0
        if k >= (1 if warm else 0):
            runs.append((wall, res))
    for o in outs:
        ctx.free(o)
    runs.sort(key=lambda x: x[0])
    wall, res = runs[len(runs) // 2]
    rays = sum(r.rays for r in res)
    return rays / wall / 1e6, wall, [r.latency_ms for r in res]


def poisson_leg(ctx, jobs, rate, rng):
    """jobs ((scene, params, camera or None)) submitted one by one at
    Poisson arrival times of the given rate (req/s), each to a device
    buffer; returns (wall s from the first arrival to the last result,
    results)"""
    outs = [ctx.malloc((p.x1 - p.x0) * (p.y1 - p.y0) * 16) for _, p, _ in jobs]
    arrivals = np.cumsum(rng.exponential(1.0 / rate, len(jobs)))
    arrivals -= arrivals[0]
    tickets = []
    t0 = time.perf_counter()
    for k, ((sid, p, cam), o) in enumerate(zip(jobs, outs)):
        wait = arrivals[k] - (time.perf_counter() - t0)
        if wait > 0:
            time.sleep(wait)
        tickets.append(ctx.submit(sid, p, cam, out=o))
    res = [ctx.poll(t)[1] for t in tickets]
    wall = time.perf_counter() - t0
    for o in outs:
        ctx.free(o)
    return wall, res


def latency(res):
    lat = [r.latency_ms for r in res]
    q = [r.queue_ms for r in res]
    return {"p50_ms": pct(lat, 50), "p99_ms": pct(lat, 99), "max_ms": round(max(lat), 3),
            "queue_p50_ms": pct(q, 50), "queue_p99_ms": pct(q, 99)}


def latency_by_size(jobs, res):
    """latency percentiles per request size (width x height)"""
    groups = {}
    for (_, p, _), r in zip(jobs, res):
        groups.setdefault(f"{p.x1 - p.x0}x{p.y1 - p.y0}", []).append(r.latency_ms)
    return {k: {"n": len(v), "p50_ms": pct(v, 50), "p99_ms": pct(v, 99)} for k, v in sorted(groups.items())}


# ------------------------------------------------------------------ legs
def run_heightfield(args):
    """the traversal-bound frame: value, roofline and CPU baseline"""
    import sray_amd as S
    config, variant, W, H, it, samples = WORKLOADS["heightfield"]
    t0 = time.perf_counter()
    hs = S.HostScene.generate(config, variant=variant)
    ctx = S.Context(0)
    p = S.render_params(W, H, samples=samples, iter=it)
    desc = S.abi.SceneDesc()
    ctypes.memmove(ctypes.byref(desc), hs.desc_ptr(), ctypes.sizeof(desc))
    if args.octree:
        desc.mesh_oct_items, desc.mesh_oct_depth = (int(v) for v in args.octree.split(","))
    elif args.octree_autotune:
        cands = [(32, 8), (48, 8), (64, 7), (64, 8)]
        best, oms = ctx.autotune_octree(ctypes.pointer(desc), p, cands, rounds=2)
        desc.mesh_oct_items, desc.mesh_oct_depth = cands[best]
        log(args, f"octree autotune: {cands[best]}, {oms:.2f} ms/frame")
    sid = ctx.upload(ctypes.pointer(desc))
    octree = {"mesh_items": desc.mesh_oct_items, "mesh_depth": desc.mesh_oct_depth}
    upload_s = time.perf_counter() - t0
    log(args, f"heightfield: {hs.desc.num_triangles} triangles, {hs.desc.num_objects} objects, upload+octree "
              f"build {upload_s:.2f} s, {ctx.scene_bytes(sid) / 2**20:.1f} MiB resident, octree {octree}")
    out = ctx.malloc(W * H * 4 * 4)
    tuning = tuning_arg(args, ctx)
    if tuning is None and not args.no_autotune:
        t, tms = ctx.autotune(sid, p, rounds=5)
        tuning = tuning_dict(t)
        log(args, f"autotune: {tuning}, {tms:.2f} ms/frame")
    if tuning is None:
        tuning = tuning_dict(ctx.tuning())
    for _ in range(args.warmup):
        ctx.render_device(sid, p, out)
    rays, dev_ms = 0, 0.0
    ctx.synchronize()   # device-wide barrier before and after the timed steps (one GPU)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        st = ctx.render_device(sid, p, out)
        rays += st.rays()
        dev_ms += st.render_ms
    ctx.synchronize()
    wall = time.perf_counter() - t0
    img = ctx.to_host(out, W * H * 16, np.float32).reshape(H, W, 4)
    # K frames of other sample seeds in flight together (one submit_many):
    # one frame's deep bounce levels hold few waves, and the other frames'
    # levels fill the GPU beside them
    # The frames take 2^21-ray chunks: a bounce level may outgrow its chunk
    # (two children per ray), and each chunk that may is finished with a
    # read-back of its level counts, a host synchronisation that holds the
    # other lanes' next chunks back, so smaller chunks (which the one-frame
    # autotune may pick within its noise) cost the frames in flight half
    # their rate.
    K = args.frames or 8
    frames = [S.render_params(W, H, samples=samples, iter=it, seed=0x5EED + k) for k in range(K)]
    plain_chunk = ctx.tuning().chunk_log2
    ctx.set_tuning(chunk_log2=21)
    batched, bwall, _ = batched_leg(ctx, [(sid, q, None) for q in frames])
    ctx.set_tuning(chunk_log2=plain_chunk)
    log(args, f"heightfield: plain {rays / wall / 1e6:.1f} Mrays/s, {K} frames in flight {batched:.1f} Mrays/s")
    # the same steps again with every launch bracketed by HIP events on the
    # stream it runs on (per-kernel durations for the roofline; kernels are
    # launched one by one while events are recorded).  This pass renders the
    # frame as one chunk on one stream, so that each k_trace_q launch covers
    # a whole bounce level and no launch overlaps another (overlapping
    # launches would each be charged the time they share).
    timed_tuning = ctx.tuning()
    ctx.set_tuning(chunk_log2=21, lanes=1)
    ctx.set_profiling(True)
    ctx.kernel_stats(reset=True)
    for _ in range(args.steps):
        ctx.render_device(sid, p, out)
    ctx.synchronize()
    ctx.set_profiling(False)
    ks = ctx.kernel_stats(reset=True)
    ctx.set_tuning(chunk_log2=timed_tuning.chunk_log2, lanes=timed_tuning.lanes)
    # visit counters of the same frame (separate, untimed run)
    pc = S.render_params(W, H, samples=samples, iter=it, flags=S.abi.SRAY_RENDER_COUNT)
    cst = ctx.render_device(sid, pc, out)
    cl = cst.closest
    closest_rays = cst.rays_primary + cst.rays_reflect + cst.rays_refract + cst.rays_diffuse
    # per ray 32 B in + 16 B hit out; per fetched node 32 B; per triangle its
    # three vertices with the face normal (48 B); per sphere / cylinder the
    # inverse matrix and radius (+ extent) (52 / 60 B); per scene item its
    # index; per mesh leaf item box tested its 8 B of codes
    alg_bytes = (32 + 16) * closest_rays + 32 * cl.nodes + 48 * cl.triangles + 52 * cl.spheres + \
        60 * cl.cylinders + 4 * (cl.item_refs - cl.triangles) + 8 * cl.item_boxes
    launches, tq_ms = ks.get("k_trace_q", (0, 0.0))
    per_frame_launches = launches / args.steps
    avg_launch_ms = tq_ms / max(launches, 1)
    bytes_per_launch = alg_bytes / max(per_frame_launches, 1)
    achieved = bytes_per_launch / (avg_launch_ms * 1e-3) / 1e9 if avg_launch_ms > 0 else 0.0
    value = rays / wall / 1e6
    kernels = {k: round(v[1] / args.steps, 4) for k, v in ks.items()}
    log(args, "per-kernel ms/step: " + ", ".join(f"{k} {v:.3f}" for k, v in kernels.items()))
    log(args, f"closest-hit visits/frame: rays {closest_rays} nodes {cl.nodes} tris {cl.triangles} "
              f"spheres {cl.spheres} idx {cl.item_refs} boxes {cl.item_boxes}; alg bytes/frame {alg_bytes / 1e6:.1f} MB")
    ctx.free(out)
    ctx.close()
    cpu, image = None, None
    if not args.no_cpu:
        cpu, cimg = cpu_frames(args, cpu_checker(ctypes.pointer(desc)), p, "heightfield")
        cpu["gpu_vs_cpu"] = round(value / cpu["value"], 1)
        image = image_check(img, ctypes.pointer(desc), p, cimg)
    line = {
        "metric": "Mrays/s on 1 GPU",
        "value": round(value, 3),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"heightfield: {hs.desc.num_triangles} triangles + {hs.desc.num_objects - 1} "
                               f"spheres, {W}x{H}, {samples[0]} spp, depth {it}",
                   "rays_per_step": rays // args.steps, "device_ms_per_step": round(dev_ms / args.steps, 3),
                   "tuning": tuning, "octree": octree, "kernel_ms_per_step": kernels},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": read_traffic("heightfield"),
                     "valu_busy": read_valu("heightfield"),
                     "kernel": "k_trace_q", "avg_launch_ms": round(avg_launch_ms, 4),
                     "launches_per_step": per_frame_launches, "alg_bytes_per_launch": int(bytes_per_launch),
                     "pass": "one chunk of 2^21 rays per bounce level, one stream"},
        "batched": {"value": round(batched, 3), "frames_in_flight": K, "wall_s": round(bwall, 4),
                    "chunk_log2": 21, "batched_vs_plain": round(batched / value, 3)},
        "cpu_baseline": cpu,
        "image": image,
    }
    return line


def image_check(img, desc_ptr, p, fast_img, photons=None):
    """RMSE of the GPU image against the reference's semantics at its own
    precision: the strict double restatement (IEEE, no -ffast-math), one
    untimed frame.  Also against the -ffast-math baseline's frame: that
    build's finite-math assumptions change some box and shadow decisions,
    as the reference's own -ffast-math build may, depending on its
    compiler."""
    import oracle
    threads, _ = cpu_share()
    ref = oracle.OracleScene(desc_ptr, 64)
    for kind, ph in enumerate(photons or ()):
        if ph is not None and len(ph):
            ref.set_photons(kind, ph)
    dimg, _ = ref.render(p, nthreads=threads)
    return {"rmse_vs_cpu": rmse(img, dimg), "tolerance": 1e-3,
            "rmse_vs_cpu_fast_math": rmse(img, fast_img) if fast_img is not None else None,
            "note": "rmse_vs_cpu: against the strict double-precision CPU render of the same frame"}


def base_line(metric, value, steps, warmup, ms, workload, extra):
    line = {"metric": metric, "value": round(value, 3), "unit": "Mrays/s", "n_gpus": 1, "steps": steps,
            "warmup": warmup, "ms_per_step": round(ms, 3), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f32", "data": "synthetic", "config": {"workload": workload}}
    line.update(extra)
    return line


def run_spheres(args):
    """64 spheres, 640x480, 1 spp, depth 5: plain GPU path, and many frames
    in flight, against the CPU renderer on whole frames"""
    import sray_amd as S
    config, variant, W, H, it, samples = WORKLOADS["spheres"]
    hs = S.HostScene.generate(config)
    ctx = S.Context(0)
    tuning_arg(args, ctx)
    sid = ctx.upload(hs)
    p = S.render_params(W, H, samples=samples, iter=it)
    out = ctx.malloc(W * H * 16)
    for _ in range(args.warmup):
        ctx.render_device(sid, p, out)
    ctx.synchronize()
    rays, t0 = 0, time.perf_counter()
    for _ in range(args.steps):
        rays += ctx.render_device(sid, p, out).rays()
    ctx.synchronize()
    wall = time.perf_counter() - t0
    value = rays / wall / 1e6
    img = ctx.to_host(out, W * H * 16, np.float32).reshape(H, W, 4)
    K = args.frames or 64
    frames = [S.render_params(W, H, samples=samples, iter=it, seed=0x5EED + k) for k in range(K)]
    batched, bwall, _ = batched_leg(ctx, [(sid, q, None) for q in frames])
    log(args, f"spheres: plain {value:.1f} Mrays/s, {K} frames in flight {batched:.1f} Mrays/s")
    ctx.free(out)
    ctx.close()
    cpu, rm = None, None
    if not args.no_cpu:
        cpu, cimg = cpu_frames(args, cpu_checker(hs.desc_ptr()), p, "spheres")
        rm = image_check(img, hs.desc_ptr(), p, cimg)
    return base_line("Mrays/s on 1 GPU", value, args.steps, args.warmup, wall * 1e3 / args.steps,
                     f"spheres: 64 spheres, {W}x{H}, 1 spp, depth {it}",
                     {"batched": {"value": round(batched, 3), "frames_in_flight": K, "wall_s": round(bwall, 4)},
                      "cpu_baseline": cpu,
                      "plain_vs_cpu": round(value / cpu["value"], 1) if cpu else None,
                      "batched_vs_cpu": round(batched / cpu["value"], 1) if cpu else None,
                      "image": rm})


def room_params(W, H, diag, seed, gather_k=200):
    import sray_amd as S
    return S.render_params(W, H, samples=(1, 16), max_var=0.05, iter=5, shadow_samples=4, diffuse_samples=1,
                           caust_photons=500_000, gi_photons=2_000_000, gather_dist=0.1 / diag, seed=seed,
                           gather_k=gather_k)


def run_room(args):
    """photon-mapped room, 1280x720, adaptive 1-16 spp, 2M global + 500k
    caustic photons shot on the device, gathers of the 200 nearest within
    0.1: K frames one after another
    (plain) against the same K frames in flight together (batched)"""
    import sray_amd as S
    config, variant, W, H, it, samples = WORKLOADS["room"]
    hs = S.HostScene.generate(config, variant=variant)
    ctx = S.Context(0)
    tuning_arg(args, ctx)
    sid = ctx.upload(hs)
    diag = S.scene_diag(hs.desc_ptr())
    t0 = time.perf_counter()
    ids, pst = ctx.shoot_photons(sid, 500_000, 2_000_000, 300.0, 5, 0, 0, 0x5EED)
    shoot_wall = time.perf_counter() - t0
    log(args, f"photons: shot {pst.shot[0]}+{pst.shot[1]}, stored {pst.stored[0]}+{pst.stored[1]}, "
              f"{pst.segments} segments, trace {pst.trace_ms:.1f} ms, kd build {pst.build_ms:.1f} ms")
    K = args.frames or 8
    frames = [room_params(W, H, diag, 100 + k, args.gather_k) for k in range(K)]
    out = ctx.malloc(W * H * 16)
    ctx.render_device(sid, frames[0], out)  # warm-up (allocations)
    ctx.synchronize()
    rays, t0 = 0, time.perf_counter()
    for p in frames:
        rays += ctx.render_device(sid, p, out).rays()
    ctx.synchronize()
    plain_wall = time.perf_counter() - t0
    plain = rays / plain_wall / 1e6
    batched, batch_wall, lat = batched_leg(ctx, [(sid, p, None) for p in frames])
    log(args, f"room: plain {plain:.2f} Mrays/s ({plain_wall:.2f} s), batched {batched:.2f} Mrays/s "
              f"({batch_wall:.2f} s), rays/frame {rays // K}")
    photons = [ctx.pmap_download(i) if i else None for i in ids] if not args.no_cpu else None
    ctx.free(out)
    ctx.close()
    cpu = None
    if not args.no_cpu:
        cpu = cpu_rows(args, cpu_checker(hs.desc_ptr(), photons), frames[0], "room")
    return base_line("Mrays/s on 1 GPU (batched)", batched, K, 1, batch_wall * 1e3 / K,
                     f"room: 12 quads + 2 glass spheres + 50k-triangle icosphere, box light, {W}x{H}, 1-16 spp "
                     f"(variance 0.05), 2M+500k photons, k={args.gather_k} nearest within 0.1, {K} frames",
                     {"batched": {"value": round(batched, 3), "frames_in_flight": K, "wall_s": round(batch_wall, 3),
                                  "p50_ms": pct(lat, 50), "p99_ms": pct(lat, 99)},
                      "plain": {"value": round(plain, 3), "wall_s": round(plain_wall, 3)},
                      "batched_vs_plain": round(batched / plain, 3),
                      "photons": {"shot": [int(pst.shot[0]), int(pst.shot[1])],
                                  "stored": [int(pst.stored[0]), int(pst.stored[1])],
                                  "segments": int(pst.segments), "trace_ms": round(pst.trace_ms, 2),
                                  "kd_build_ms": round(pst.build_ms, 2),
                                  "Msegments_per_s": round(pst.segments / (pst.trace_ms * 1e-3) / 1e6, 2),
                                  "wall_s": round(shoot_wall, 3)},
                      "cpu_baseline": cpu,
                      "batched_vs_cpu": round(batched / cpu["value"], 1) if cpu else None})


def run_batch(args):
    """many independent scenes (triangle counts log-uniform in [1k, 200k]
    plus 8-32 spheres, resolutions 64^2..512^2, half with photon maps of
    100k-500k photons, 8 motion-blurred time samples per pixel): one after
    another (plain), all in flight (batched), and arriving as a Poisson
    stream below the measured capacity (latency)"""
    import sray_amd as S
    rng = np.random.default_rng(0x5EED0004)
    n = args.scenes or 256
    sids, params = [], []
    t0 = time.perf_counter()
    ctx = S.Context(0)
    tuning_arg(args, ctx)
    keep = []
    for k in range(n):
        ntri = int(math.exp(rng.uniform(math.log(1000), math.log(200_000))))
        hs = S.HostScene.generate(3, seed=0x5EED0400 + k, variant=ntri)
        d = S.abi.SceneDesc()
        ctypes.memmove(ctypes.byref(d), hs.desc_ptr(), ctypes.sizeof(d))
        d.camera.shutter = 40
        sid = ctx.upload(ctypes.pointer(d))
        keep.append((hs, d))
        res = 64 * int(rng.integers(1, 9))
        kw = {}
        if k % 2 == 0:
            nph = int(rng.integers(100_000, 500_001))
            ctx.shoot_photons(sid, nph // 5, nph - nph // 5, 300.0, 5, 0, 40, 0x5EED + k)
            kw = dict(caust_photons=1, gi_photons=1, gather_dist=0.1 / S.scene_diag(ctypes.pointer(d)))
        p = S.render_params(res, res, samples=(8, 8), iter=4, mblur=1, frame_time=20, seed=k, **kw)
        sids.append(sid)
        params.append(p)
    setup = time.perf_counter() - t0
    log(args, f"batch: {n} scenes set up in {setup:.1f} s")
    out = ctx.malloc(512 * 512 * 16)
    ctx.render_device(sids[0], params[0], out)
    ctx.synchronize()
    rays, t0 = 0, time.perf_counter()
    for sid, p in zip(sids, params):
        rays += ctx.render_device(sid, p, out).rays()
    ctx.synchronize()
    plain_wall = time.perf_counter() - t0
    plain = rays / plain_wall / 1e6
    ctx.free(out)
    jobs = [(sid, p, None) for sid, p in zip(sids, params)]
    batched, batch_wall, lat = batched_leg(ctx, jobs)
    capacity = n / batch_wall
    rate = args.load * capacity
    order = rng.permutation(n)
    pwall, pres = poisson_leg(ctx, [jobs[i] for i in order], rate, rng)
    prays = sum(r.rays for r in pres)
    log(args, f"batch: plain {plain:.2f} Mrays/s ({plain_wall:.2f} s), batched {batched:.2f} Mrays/s "
              f"({batch_wall:.3f} s), Poisson {rate:.0f} req/s: {prays / pwall / 1e6:.1f} Mrays/s, "
              f"p50 {pct([r.latency_ms for r in pres], 50)} ms p99 {pct([r.latency_ms for r in pres], 99)} ms")
    ctx.close()
    return base_line("Mrays/s on 1 GPU (batched)", batched, n, 1, batch_wall * 1e3 / n,
                     f"batch: {n} scenes (1k-200k triangles + 8-32 spheres, 64^2-512^2, half with 100k-500k "
                     f"photons, 8 motion-blur samples/pixel)",
                     {"batched": {"value": round(batched, 3), "wall_s": round(batch_wall, 3),
                                  "p50_ms": pct(lat, 50), "p99_ms": pct(lat, 99),
                                  "note": "all submitted at once: every request finishes on the last pass"},
                      "poisson": dict({"offered_rate_per_s": round(rate, 1), "load": args.load,
                                       "requests": n, "wall_s": round(pwall, 3),
                                       "value": round(prays / pwall / 1e6, 3)}, **latency(pres)),
                      "plain": {"value": round(plain, 3), "wall_s": round(plain_wall, 3)},
                      "batched_vs_plain": round(batched / plain, 3), "setup_s": round(setup, 1)})


def run_serve(args):
    """Poisson request stream over resident photon-mapped rooms: tiles of
    64x64 / 256x256 / 1920x1080 (p = .6/.3/.1) through random cameras on a
    shell around the room centre, at a rate below the capacity measured
    with a burst of the same requests"""
    import sray_amd as S
    rng = np.random.default_rng(0x5EED0005)
    ctx = S.Context(0)
    tuning_arg(args, ctx)
    sids, diags = [], []
    t0 = time.perf_counter()
    for k in range(args.scenes or 32):
        hs = S.HostScene.generate(2, seed=k)
        sid = ctx.upload(hs)
        ctx.shoot_photons(sid, 500_000, 2_000_000, 300.0, 5, 0, 0, 0x5EED + k)
        diags.append(S.scene_diag(hs.desc_ptr()))
        sids.append(sid)
    setup = time.perf_counter() - t0
    sizes = [(64, 64), (256, 256), (1920, 1080)]

    def make(k):
        sc = int(rng.integers(0, len(sids)))
        w, h = sizes[int(rng.choice(3, p=[0.6, 0.3, 0.1]))]
        v = rng.normal(size=3)
        v /= np.linalg.norm(v)
        cam = look_at(np.array([0.0, 4.0, 0.0]) + 3.5 * v, (0.0, 4.0, 0.0), 60.0)
        p = S.render_params(w, h, samples=(1, 16), max_var=0.05, iter=5, shadow_samples=4, caust_photons=1,
                            gi_photons=1, gather_dist=0.1 / diags[sc], seed=k, gather_k=args.gather_k)
        return (sids[sc], p, cam)

    # capacity: the completion rate of a stream of the same mix offered
    # faster than it can be served (a burst, with every request in flight
    # at once, batches better than a stream ever does and overstates it)
    burst = [make(100000 + k) for k in range(200)]
    _, bwall, _ = batched_leg(ctx, burst, warm=True, timed_rounds=1)
    probe = [make(200000 + k) for k in range(300)]
    pwall, _ = poisson_leg(ctx, probe, 2.0 * len(burst) / bwall, rng)
    capacity = len(probe) / pwall
    rate = args.rate if args.rate else args.load * capacity
    reqs = [make(k) for k in range(args.requests)]
    wall, res = poisson_leg(ctx, reqs, rate, rng)
    rays = sum(r.rays for r in res)
    value = rays / wall / 1e6
    lat = latency(res)
    log(args, f"serve: capacity {capacity:.1f} req/s; {len(reqs)} requests at {rate:.1f} req/s in {wall:.2f} s, "
              f"p50 {lat['p50_ms']} ms p99 {lat['p99_ms']} ms, {value:.1f} Mrays/s")
    ctx.close()
    return base_line("Mrays/s on 1 GPU (served)", value, len(reqs), 1, wall * 1e3 / len(reqs),
                     f"serve: Poisson stream over {len(sids)} resident photon-mapped rooms (2M+500k photons "
                     f"each), tiles 64^2/256^2/1920x1080 at .6/.3/.1, 1-16 spp",
                     {"latency_ms": lat, "latency_by_tile": latency_by_size(reqs, res),
                      "requests": len(reqs), "offered_rate_per_s": round(rate, 2),
                      "capacity_req_per_s": round(capacity, 2), "burst_req_per_s": round(len(burst) / bwall, 2),
                      "load": round(rate / capacity, 3),
                      "completed_rate_per_s": round(len(reqs) / wall, 2), "wall_s": round(wall, 2),
                      "setup_s": round(setup, 1)})


def leg_summary(line):
    """a leg's line without the fields the headline line already carries"""
    drop = {"n_gpus", "higher_is_better", "scaling", "vs_baseline", "dtype", "data"}
    return {k: v for k, v in line.items() if k not in drop}


def run_all(args):
    t_all = time.perf_counter()
    line = run_heightfield(args)
    legs = {}
    for name, fn in (("spheres", run_spheres), ("room", run_room), ("batch", run_batch), ("serve", run_serve)):
        t0 = time.perf_counter()
        legs[name] = leg_summary(fn(args))
        legs[name]["leg_wall_s"] = round(time.perf_counter() - t0, 1)
        log(args, f"leg {name} done in {legs[name]['leg_wall_s']} s")
    sph, room, bat, srv = legs["spheres"], legs["room"], legs["batch"], legs["serve"]
    line["legs"] = legs
    line["targets"] = {
        "spheres_batched_vs_cpu": sph.get("batched_vs_cpu"),
        "spheres_plain_vs_cpu": sph.get("plain_vs_cpu"),
        "spheres_gpu_ge_100x_cpu": (sph["batched_vs_cpu"] >= 100) if sph.get("batched_vs_cpu") else None,
        "heightfield_batched_vs_plain": line["batched"]["batched_vs_plain"],
        "room_batched_vs_plain": room["batched_vs_plain"],
        "room_batched_ge_3x_plain": room["batched_vs_plain"] >= 3,
        "batch_poisson_p99_ms": bat["poisson"]["p99_ms"],
        "serve_p50_ms": srv["latency_ms"]["p50_ms"],
        "serve_p99_ms": srv["latency_ms"]["p99_ms"],
        "image_rmse_vs_cpu": line["image"]["rmse_vs_cpu"] if line.get("image") else None,
        "image_rmse_le_1e-3": (line["image"]["rmse_vs_cpu"] <= 1e-3) if line.get("image") else None,
    }
    line["bench_wall_s"] = round(time.perf_counter() - t_all, 1)
    return line


def main():
    args = parse()
    if args.gpus > 1:
        sys.exit("bench.py measures one GPU: the path runs on a single MI355X, and sharding over several "
                 "GPUs is a design only (run with --gpus 1)")
    if args.copy_calibration:
        import sray_amd as S
        ctx = S.Context(0)
        n = 256 << 20
        a, b = ctx.malloc(n), ctx.malloc(n)
        for _ in range(3):
            ctx.memcpy_device(b, a, n)
        ctx.synchronize()
        ctx.free(a)
        ctx.free(b)
        ctx.close()
        print(json.dumps({"copy_bytes": n, "copies": 3}), flush=True)
        return
    fn = {"all": run_all, "heightfield": run_heightfield, "spheres": run_spheres, "room": run_room,
          "batch": run_batch, "serve": run_serve}[args.workload]
    print(json.dumps(fn(args)), flush=True)


if __name__ == "__main__":
    main()
