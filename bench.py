"""Benchmark: Mrays/s of the S-ray hot path on one MI355X.

A step is one full frame of the traversal-bound heightfield workload
(1000x500-quad fBm terrain = 1,000,000 triangles + 20 glass + 20 mirror
spheres; 1920x1080, 1 sample/pixel, recursion depth 6; mesh octree leaf size
and depth picked by the octree autotuner among 32/8 (the reference's), 48/8,
64/7 and 64/8): primary, shadow, reflect and refract rays through the
wavefront path, scene and octrees resident in HBM before the timed region.

Prints one JSON line with the value, a roofline object for the dominant
kernel (the closest-hit traversal k_trace_q: algorithmic bytes from the
library's visit counters / its HIP-event-timed launch duration) and a CPU
baseline (the CPU restatement of the reference renderer, double precision,
-O3 -ffast-math, reference work pattern, on a bounded sample of rows).
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

# A render lane uses two HIP streams (4 lanes by default).  With HIP's
# default of 4 hardware queues per process, streams share queues and wait
# behind each other's event waits; 8 queues give each stream its own.  Set
# before the HIP runtime starts (it starts when the library is loaded).
if int(os.environ.get("GPU_MAX_HW_QUEUES", "4")) < 8:
    os.environ["GPU_MAX_HW_QUEUES"] = "8"

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
    ap.add_argument("--workload", default="heightfield", choices=sorted(WORKLOADS))
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="CPU baseline sample budget")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default tuning (no untimed search)")
    ap.add_argument("--tuning", default=None,
                    help="set the tuning (heightfield: instead of searching), e.g. "
                         "trace_block=64,tile_w=64,chunk_log2=21,graphs=1,persistent=0")
    ap.add_argument("--octree", default=None,
                    help="heightfield: mesh octree 'items,depth' instead of searching (the reference's is 32,8)")
    ap.add_argument("--frames", type=int, default=None, help="room / spheres: frames in flight (8 / 64)")
    ap.add_argument("--scenes", type=int, default=None, help="batch: scenes (256); serve: resident scenes (32)")
    ap.add_argument("--rate", type=float, default=2000.0, help="serve: Poisson arrival rate (req/s)")
    ap.add_argument("--duration", type=float, default=0.25, help="serve: arrival window (s)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 256 MiB device to device three times (scales the HBM byte counters)")
    ap.add_argument("--quiet", action="store_true")
    return ap.parse_args()


def log(args, *a):
    if not args.quiet:
        print(*a, file=sys.stderr, flush=True)


def read_traffic(workload):
    """HBM bytes per k_trace_q launch from the committed rocprofv3 --pmc summary"""
    path = os.path.join(ROOT, "profiles", f"pmc_k_trace_q_{workload}.txt")
    if not os.path.exists(path):
        return None
    vals = {}
    with open(path) as f:
        for line in f:
            if ":" in line and not line.startswith("#"):
                k, v = line.split(":", 1)
                try:
                    vals[k.strip()] = float(v.split()[0])
                except (ValueError, IndexError):
                    pass
    if "hbm_bytes_per_dispatch" in vals:
        return vals["hbm_bytes_per_dispatch"]
    return None


def cpu_baseline(args, hs, wl, p=None, full=False):
    """the CPU restatement (reference semantics) timed on this host's cores:
    the whole frame (full) or bands of rows around the centre until about
    --cpu-seconds of work"""
    import oracle
    import sray_amd as S
    config, variant, W, H, it, samples = WORKLOADS[wl]
    nthreads = min(os.cpu_count() or 1, 16)
    ref = oracle.OracleScene(hs.desc_ptr(), 64, oracle.Q9_RETRAVERSE, fast=True)
    band = H if full else 16
    rows, rays, elapsed = 0, 0, 0.0
    y0 = 0 if full else max(0, H // 2 - band)
    # whole frames (full) until about --cpu-seconds of work, or bands of rows
    # until the frame is covered or the budget is spent
    while (elapsed < args.cpu_seconds) if full else (rows < H and elapsed < args.cpu_seconds):
        y1 = min(H, y0 + band)
        q = S.render_params(W, H, rect=(0, y0, W, y1), samples=samples, iter=it)
        if p is not None:
            for f in ("max_var", "shadow_samples", "diffuse_samples", "caust_photons", "gi_photons",
                      "gather_dist", "seed"):
                setattr(q, f, getattr(p, f))
        t0 = time.perf_counter()
        _, st = ref.render(q, nthreads=nthreads)
        elapsed += time.perf_counter() - t0
        rays += st.rays()
        rows += y1 - y0
        y0 = y1 if y1 < H else 0
    log(args, f"cpu baseline: {rows} rows, {rays} rays, {elapsed:.1f} s, {nthreads} threads")
    what = (f"{rows // H} whole frames" if full else
            f"{rows} rows x {W} px ({rows * W} pixels) of the {W}x{H} frame")
    return {"value": rays / elapsed / 1e6, "unit": "Mrays/s", "cores": nthreads, "kind": "port",
            "sample": f"{what}, {rays} ray queries, double precision, -O3 -ffast-math, "
                      f"mesh re-traversal per scene leaf as the reference"}


def pct(v, q):
    return float(np.percentile(np.asarray(v, np.float64), q)) if len(v) else None


def base_line(metric, value, steps, warmup, ms, dtype, workload, extra):
    line = {"metric": metric, "value": round(value, 3), "unit": "Mrays/s", "n_gpus": 1, "steps": steps,
            "warmup": warmup, "ms_per_step": round(ms, 3), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": dtype, "data": "synthetic", "config": {"workload": workload}}
    line.update(extra)
    return line


def tuning_arg(args, ctx):
    """--tuning applied to the context; the tuning as a dict, or None"""
    if not args.tuning:
        return None
    kw = {k: (float(v) if k == "gather_cell" else int(v))
          for k, v in (item.split("=") for item in args.tuning.split(","))}
    t = ctx.set_tuning(**kw)
    return {"trace_block": t.trace_block, "tile_w": t.tile_w, "chunk_log2": t.chunk_log2,
            "graphs": t.graphs, "persistent": t.persistent, "lanes": t.lanes, "sparse_from": t.sparse_from}


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


def run_spheres(args):
    """64 spheres, 640x480, 1 spp, depth 5: plain GPU path against the CPU
    renderer on the whole frame"""
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
    # the same frame, many in flight (different seeds): the serving engine
    K = args.frames
    frames = [S.render_params(W, H, samples=samples, iter=it, seed=0x5EED + k) for k in range(K)]
    batched, bwall, _ = batched_leg(ctx, [(sid, q, None) for q in frames])
    log(args, f"spheres: plain {value:.1f} Mrays/s, {K} frames in flight {batched:.1f} Mrays/s")
    cpu = None if args.no_cpu else cpu_baseline(args, hs, "spheres", full=True)
    extra = {"batched": {"value": round(batched, 3), "frames_in_flight": K, "wall_s": round(bwall, 4)},
             "cpu_baseline": cpu, "gpu_vs_cpu": round(value / cpu["value"], 1) if cpu else None,
             "batched_vs_cpu": round(batched / cpu["value"], 1) if cpu else None}
    print(json.dumps(base_line("Mrays/s on 1 GPU", value, args.steps, args.warmup, wall * 1e3 / args.steps, "f32",
                               f"spheres: 64 spheres, {W}x{H}, 1 spp, depth {it}", extra)), flush=True)
    ctx.free(out)
    ctx.close()


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


def room_params(W, H, diag, seed):
    import sray_amd as S
    return S.render_params(W, H, samples=(1, 16), max_var=0.05, iter=5, shadow_samples=4, diffuse_samples=1,
                           caust_photons=500_000, gi_photons=2_000_000, gather_dist=0.1 / diag, seed=seed)


def run_room(args):
    """photon-mapped room, 1280x720, adaptive 1-16 spp, 2M global + 500k
    caustic photons shot on the device: K frames one after another
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
    K = args.frames
    frames = [room_params(W, H, diag, 100 + k) for k in range(K)]
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
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(args, hs, "room", p=frames[0])
    extra = {"batched": {"value": round(batched, 3), "frames_in_flight": K, "wall_s": round(batch_wall, 3),
                         "p50_ms": pct(lat, 50), "p99_ms": pct(lat, 99)},
             "plain": {"value": round(plain, 3), "wall_s": round(plain_wall, 3)},
             "batched_vs_plain": round(batched / plain, 3),
             "photons": {"shot": [int(pst.shot[0]), int(pst.shot[1])],
                         "stored": [int(pst.stored[0]), int(pst.stored[1])],
                         "segments": int(pst.segments), "trace_ms": round(pst.trace_ms, 2),
                         "kd_build_ms": round(pst.build_ms, 2),
                         "Msegments_per_s": round(pst.segments / (pst.trace_ms * 1e-3) / 1e6, 2),
                         "wall_s": round(shoot_wall, 3)},
             "cpu_baseline": cpu}
    print(json.dumps(base_line("Mrays/s on 1 GPU (batched)", batched, K, 1, batch_wall * 1e3 / K, "f32",
                               f"room: 12 quads + 2 glass spheres + 50k-triangle icosphere, box light, "
                               f"{W}x{H}, 1-16 spp (variance 0.05), 2M+500k photons, {K} frames", extra)),
          flush=True)
    ctx.free(out)
    ctx.close()


def run_batch(args):
    """many independent scenes (triangle counts log-uniform in [1k, 200k]
    plus 8-32 spheres, resolutions 64^2..512^2, half with photon maps of
    100k-500k photons, 8 motion-blurred time samples per pixel): one after
    another (plain) against all in flight (batched)"""
    import sray_amd as S
    rng = np.random.default_rng(0x5EED0004)
    n = args.scenes
    scenes, sids, params = [], [], []
    t0 = time.perf_counter()
    ctx = S.Context(0)
    tuning_arg(args, ctx)
    for k in range(n):
        ntri = int(math.exp(rng.uniform(math.log(1000), math.log(200_000))))
        hs = S.HostScene.generate(3, seed=0x5EED0400 + k, variant=ntri)
        d = S.abi.SceneDesc()
        ctypes.memmove(ctypes.byref(d), hs.desc_ptr(), ctypes.sizeof(d))
        d.camera.shutter = 40
        sid = ctx.upload(ctypes.pointer(d))
        res = 64 * int(rng.integers(1, 9))
        photons = k % 2 == 0
        kw = {}
        if photons:
            nph = int(rng.integers(100_000, 500_001))
            ctx.shoot_photons(sid, nph // 5, nph - nph // 5, 300.0, 5, 0, 40, 0x5EED + k)
            diag = S.scene_diag(ctypes.pointer(d))
            kw = dict(caust_photons=1, gi_photons=1, gather_dist=0.1 / diag)
        p = S.render_params(res, res, samples=(8, 8), iter=4, mblur=1, frame_time=20, seed=k, **kw)
        scenes.append((hs, d))
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
    batched, batch_wall, lat = batched_leg(ctx, [(sid, p, None) for sid, p in zip(sids, params)])
    log(args, f"batch: plain {plain:.2f} Mrays/s ({plain_wall:.2f} s) batched {batched:.2f} Mrays/s "
              f"({batch_wall:.2f} s)")
    extra = {"batched": {"value": round(batched, 3), "wall_s": round(batch_wall, 3), "p50_ms": pct(lat, 50),
                         "p99_ms": pct(lat, 99)},
             "plain": {"value": round(plain, 3), "wall_s": round(plain_wall, 3)},
             "batched_vs_plain": round(batched / plain, 3), "setup_s": round(setup, 1)}
    print(json.dumps(base_line("Mrays/s on 1 GPU (batched)", batched, n, 1, batch_wall * 1e3 / n, "f32",
                               f"batch: {n} scenes (1k-200k triangles + 8-32 spheres, 64^2-512^2, half with "
                               f"100k-500k photons, 8 motion-blur samples/pixel)", extra)), flush=True)
    ctx.free(out)
    ctx.close()


def run_serve(args):
    """Poisson request stream over resident photon-mapped rooms: tiles of
    64x64 / 256x256 / 1920x1080 (p = .6/.3/.1) through random cameras on a
    shell around the room centre"""
    import sray_amd as S
    rng = np.random.default_rng(0x5EED0005)
    ctx = S.Context(0)
    tuning_arg(args, ctx)
    sids, diags = [], []
    t0 = time.perf_counter()
    for k in range(args.scenes):
        hs = S.HostScene.generate(2, seed=k)
        sid = ctx.upload(hs)
        ctx.shoot_photons(sid, 500_000, 2_000_000, 300.0, 5, 0, 0, 0x5EED + k)
        diags.append(S.scene_diag(hs.desc_ptr()))
        sids.append(sid)
    setup = time.perf_counter() - t0
    sizes = [(64, 64), (256, 256), (1920, 1080)]
    n = int(args.rate * args.duration)
    arrivals = np.cumsum(rng.exponential(1.0 / args.rate, n))
    reqs = []
    for k in range(n):
        sc = int(rng.integers(0, len(sids)))
        w, h = sizes[int(rng.choice(3, p=[0.6, 0.3, 0.1]))]
        v = rng.normal(size=3)
        v /= np.linalg.norm(v)
        pos = np.array([0.0, 4.0, 0.0]) + 3.5 * v
        cam = look_at(pos, (0.0, 4.0, 0.0), 60.0)
        p = S.render_params(w, h, samples=(1, 16), max_var=0.05, iter=5, shadow_samples=4, caust_photons=1,
                            gi_photons=1, gather_dist=0.1 / diags[sc], seed=k)
        reqs.append((sids[sc], p, cam))
    ctx.poll(ctx.submit(*reqs[0]))  # warm-up
    tickets = []
    t0 = time.perf_counter()
    for k, r in enumerate(reqs):
        wait = arrivals[k] - (time.perf_counter() - t0)
        if wait > 0:
            time.sleep(wait)
        tickets.append(ctx.submit(*r))
    res = [ctx.poll(t)[1] for t in tickets]
    wall = time.perf_counter() - t0
    rays = sum(r.rays for r in res)
    lat = [r.latency_ms for r in res]
    value = rays / wall / 1e6
    log(args, f"serve: {n} requests in {wall:.2f} s, p50 {pct(lat, 50):.1f} ms p99 {pct(lat, 99):.1f} ms")
    extra = {"latency_ms": {"p50": pct(lat, 50), "p99": pct(lat, 99), "max": max(lat)},
             "queue_ms": {"p50": pct([r.queue_ms for r in res], 50), "p99": pct([r.queue_ms for r in res], 99)},
             "requests": n, "offered_rate_per_s": args.rate, "arrival_window_s": args.duration,
             "completed_rate_per_s": round(n / wall, 1), "setup_s": round(setup, 1)}
    print(json.dumps(base_line("Mrays/s on 1 GPU (served)", value, n, 1, wall * 1e3 / n, "f32",
                               f"serve: Poisson {args.rate:g} req/s for {args.duration:g} s over "
                               f"{args.scenes} resident photon-mapped rooms (2M+500k photons each), tiles "
                               f"64^2/256^2/1920x1080 at .6/.3/.1, 1-16 spp", extra)), flush=True)
    ctx.close()


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
    if args.scenes is None:
        args.scenes = 256 if args.workload == "batch" else 32
    if args.frames is None:
        args.frames = 64 if args.workload == "spheres" else 8
    if args.workload != "heightfield":
        return {"spheres": run_spheres, "room": run_room, "batch": run_batch, "serve": run_serve}[args.workload](args)
    import sray_amd as S
    wl = args.workload
    config, variant, W, H, it, samples = WORKLOADS[wl]
    t0 = time.perf_counter()
    hs = S.HostScene.generate(config, variant=variant)
    ctx = S.Context(0)
    p = S.render_params(W, H, samples=samples, iter=it)
    # the mesh octree's leaf size and depth: given, searched (untimed; the
    # candidates change only which of two exactly equal hit distances
    # wins), or the reference's
    desc = S.abi.SceneDesc()
    ctypes.memmove(ctypes.byref(desc), hs.desc_ptr(), ctypes.sizeof(desc))
    if args.octree:
        desc.mesh_oct_items, desc.mesh_oct_depth = (int(v) for v in args.octree.split(","))
    elif args.tuning is None and not args.no_autotune:
        cands = [(32, 8), (48, 8), (64, 7), (64, 8)]
        best, oms = ctx.autotune_octree(ctypes.pointer(desc), p, cands, rounds=2)
        desc.mesh_oct_items, desc.mesh_oct_depth = cands[best]
        log(args, f"octree autotune: {cands[best]}, {oms:.2f} ms/frame")
    sid = ctx.upload(ctypes.pointer(desc))
    octree = {"mesh_items": desc.mesh_oct_items, "mesh_depth": desc.mesh_oct_depth}
    upload_s = time.perf_counter() - t0
    log(args, f"scene {wl}: {hs.desc.num_triangles} triangles, {hs.desc.num_objects} objects, "
              f"upload+octree build {upload_s:.2f} s, {ctx.scene_bytes(sid) / 2**20:.1f} MiB resident, octree {octree}")
    out = ctx.malloc(W * H * 4 * 4)
    tuning = tuning_arg(args, ctx)
    if tuning is None and not args.no_autotune:
        t, tms = ctx.autotune(sid, p, rounds=2)
        tuning = {"trace_block": t.trace_block, "tile_w": t.tile_w, "chunk_log2": t.chunk_log2,
                  "graphs": t.graphs, "persistent": t.persistent, "lanes": t.lanes, "sparse_from": t.sparse_from}
        log(args, f"autotune: {tuning}, {tms:.2f} ms/frame")
    for _ in range(args.warmup):
        ctx.render_device(sid, p, out)
    ctx.synchronize()
    rays = 0
    dev_ms = 0.0
    ctx.synchronize()   # device-wide barrier before and after the timed steps (one GPU)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        st = ctx.render_device(sid, p, out)
        rays += st.rays()
        dev_ms += st.render_ms
    ctx.synchronize()
    wall = time.perf_counter() - t0
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
    # inverse matrix and radius (+ extent) (52 / 60 B); per scene item its index
    alg_bytes = (32 + 16) * closest_rays + 32 * cl.nodes + 48 * cl.triangles + 52 * cl.spheres + \
        60 * cl.cylinders + 4 * (cl.item_refs - cl.triangles)
    launches, tq_ms = ks.get("k_trace_q", (0, 0.0))
    per_frame_launches = launches / args.steps
    avg_launch_ms = tq_ms / max(launches, 1)
    bytes_per_launch = alg_bytes / max(per_frame_launches, 1)
    achieved = bytes_per_launch / (avg_launch_ms * 1e-3) / 1e9 if avg_launch_ms > 0 else 0.0
    traffic = read_traffic(wl)
    value = rays / wall / 1e6
    kernels = {k: {"launches": v[0], "ms_per_step": v[1] / args.steps} for k, v in ks.items()}
    log(args, f"per-kernel ms/step: " + ", ".join(f"{k} {v['ms_per_step']:.3f}" for k, v in kernels.items()))
    log(args, f"closest-hit visits/frame: rays {closest_rays} nodes {cl.nodes} tris {cl.triangles} "
              f"spheres {cl.spheres} idx {cl.item_refs}; alg bytes/frame {alg_bytes / 1e6:.1f} MB")
    sh = cst.shadow
    log(args, f"shadow visits/frame: rays {cst.rays_shadow} nodes {sh.nodes} tris {sh.triangles} "
              f"spheres {sh.spheres} idx {sh.item_refs} mesh enters {sh.mesh_enters}")
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(args, hs, wl)
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
        "config": {"workload": f"{wl}: {hs.desc.num_triangles} triangles + {hs.desc.num_objects - 1} spheres, "
                               f"{W}x{H}, {samples[0]}-{samples[1]} spp, depth {it}",
                   "rays_per_step": rays // args.steps, "device_ms_per_step": round(dev_ms / args.steps, 3),
                   "tuning": tuning, "octree": octree},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                     "kernel": "k_trace_q", "avg_launch_ms": round(avg_launch_ms, 4),
                     "launches_per_step": per_frame_launches, "alg_bytes_per_launch": int(bytes_per_launch),
                     "pass": "one chunk of 2^21 rays per bounce level, one stream"},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    ctx.free(out)
    ctx.close()


if __name__ == "__main__":
    main()
