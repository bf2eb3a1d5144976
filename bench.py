"""Benchmark: Mrays/s of the S-ray hot path on one MI355X.

A step is one full frame of the traversal-bound heightfield workload
(1000x500-quad fBm terrain = 1,000,000 triangles + 20 glass + 20 mirror
spheres; 1920x1080, 1 sample/pixel, recursion depth 6, mesh octree depth 8 /
32 items): primary, shadow, reflect and refract rays through the wavefront
path, scene and octrees resident in HBM before the timed region.

Prints one JSON line with the value, a roofline object for the dominant
kernel (the closest-hit traversal k_trace_q: algorithmic bytes from the
library's visit counters / its HIP-event-timed launch duration) and a CPU
baseline (the CPU restatement of the reference renderer, double precision,
-O3 -ffast-math, reference work pattern, on a bounded sample of rows).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
WORKLOADS = {
    # name: (config, variant, width, height, iter, samples)
    "heightfield": (1, 0, 1920, 1080, 6, (1, 1)),
    "spheres": (0, 0, 640, 480, 5, (1, 1)),
}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="heightfield", choices=sorted(WORKLOADS))
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="CPU baseline sample budget")
    ap.add_argument("--no-cpu", action="store_true")
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


def cpu_baseline(args, hs, wl):
    """the CPU restatement (reference semantics) timed on this host's cores"""
    import oracle
    import sray_amd as S
    config, variant, W, H, it, samples = WORKLOADS[wl]
    nthreads = min(os.cpu_count() or 1, 16)
    ref = oracle.OracleScene(hs.desc_ptr(), 64, oracle.Q9_RETRAVERSE, fast=True)
    rows, y, rays, elapsed = 0, H // 2 - 24, 0, 0.0
    band = 16
    while elapsed < args.cpu_seconds and rows < H:
        y0 = max(0, min(H - band, y))
        p = S.render_params(W, H, rect=(0, y0, W, y0 + band), samples=samples, iter=it)
        t0 = time.perf_counter()
        _, st = ref.render(p, nthreads=nthreads)
        elapsed += time.perf_counter() - t0
        rays += st.rays()
        rows += band
        y += band
    log(args, f"cpu baseline: {rows} rows, {rays} rays, {elapsed:.1f} s, {nthreads} threads")
    return {"value": rays / elapsed / 1e6, "unit": "Mrays/s", "cores": nthreads, "kind": "port",
            "sample": f"{rows} rows x {W} px ({rows * W} pixels, {rays} ray queries) of the {W}x{H} frame, "
                      f"double precision, -O3 -ffast-math, mesh re-traversal per scene leaf as the reference"}


def main():
    args = parse()
    if args.gpus > 1:
        sys.exit("bench.py measures one GPU only: every box here has one MI355X and no multi-GPU "
                 "figure is produced (run with --gpus 1)")
    import sray_amd as S
    wl = args.workload
    config, variant, W, H, it, samples = WORKLOADS[wl]
    t0 = time.perf_counter()
    hs = S.HostScene.generate(config, variant=variant)
    ctx = S.Context(0)
    sid = ctx.upload(hs)
    upload_s = time.perf_counter() - t0
    log(args, f"scene {wl}: {hs.desc.num_triangles} triangles, {hs.desc.num_objects} objects, "
              f"upload+octree build {upload_s:.2f} s, {ctx.scene_bytes(sid) / 2**20:.1f} MiB resident")
    p = S.render_params(W, H, samples=samples, iter=it)
    out = ctx.malloc(W * H * 4 * 4)
    for _ in range(args.warmup):
        ctx.render_device(sid, p, out)
    ctx.synchronize()
    ctx.set_profiling(True)
    ctx.kernel_stats(reset=True)
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
    ctx.set_profiling(False)
    ks = ctx.kernel_stats(reset=True)
    # visit counters of the same frame (separate, untimed run)
    pc = S.render_params(W, H, samples=samples, iter=it, flags=S.abi.SRAY_RENDER_COUNT)
    cst = ctx.render_device(sid, pc, out)
    cl = cst.closest
    closest_rays = cst.rays_primary + cst.rays_reflect + cst.rays_refract + cst.rays_diffuse
    alg_bytes = (32 + 16) * closest_rays + 32 * cl.nodes + 36 * cl.triangles + 16 * cl.spheres + \
        28 * cl.cylinders + 4 * cl.item_refs
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
                   "rays_per_step": rays // args.steps, "device_ms_per_step": round(dev_ms / args.steps, 3)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                     "kernel": "k_trace_q", "avg_launch_ms": round(avg_launch_ms, 4),
                     "launches_per_step": per_frame_launches, "alg_bytes_per_launch": int(bytes_per_launch)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    ctx.free(out)
    ctx.close()


if __name__ == "__main__":
    main()
