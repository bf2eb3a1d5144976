"""Device traversal parity: sray_gpu_trace (C ABI) against the CPU checker.

Bar: hit object / triangle IDs bit-exact and t bit-exact against the strict
float restatement (same operation order), on every ray, including rays
where distinct items tie at exactly the same distance (the checker flags
those; their count is reported).  Agreement with the double restatement
(the reference's own precision) is reported and must stay above 99.9%.
"""
import numpy as np
import pytest

import oracle
import sray_amd as S

pytestmark = pytest.mark.gpu


def _compare(hits, ohits, ext, label):
    tie = ext["tie"] != 0
    bad_obj = np.flatnonzero(hits["obj"] != ohits["obj"])
    bad_prim = np.flatnonzero(hits["prim"] != ohits["prim"])
    hit = ohits["obj"] >= 0
    bad_t = np.flatnonzero(hit & (hits["t"].view(np.uint32) != ohits["t"].view(np.uint32)))
    msg = (f"{label}: {len(hits)} rays, {hit.sum()} hits, {tie.sum()} ties; "
           f"obj mismatches {len(bad_obj)} prim {len(bad_prim)} t {len(bad_t)}")
    print(msg)
    if len(bad_obj) or len(bad_prim) or len(bad_t):
        i = (list(bad_obj) + list(bad_prim) + list(bad_t))[0]
        raise AssertionError(msg + f"; first ray {i}: gpu {hits[i]} cpu {ohits[i]}")


def _scene(config, variant=0):
    return S.HostScene.generate(config, variant=variant)


def test_primary_spheres(ctx):
    hs = _scene(0)
    sid = ctx.upload(hs)
    ref = oracle.OracleScene(hs.desc_ptr(), 32)
    p = S.render_params(640, 480, iter=5)
    rays = ref.primary_rays(p)
    hits = ctx.trace(sid, rays)
    ohits, ext = ref.trace(rays)
    _compare(hits, ohits, ext, "spheres primary")
    r64 = oracle.OracleScene(hs.desc_ptr(), 64)
    h64, _ = r64.trace(rays)
    agree = (h64["obj"] == hits["obj"]).mean()
    print(f"agreement with the double restatement: {agree:.6f}")
    assert agree > 0.999
    ctx.release(sid)


def test_device_raygen_matches_checker(ctx):
    """raygen kernel == Camera::get_primary_ray restatement, bit for bit
    (sample 0 and a jittered sample)"""
    hs = _scene(0)
    sid = ctx.upload(hs)
    ref = oracle.OracleScene(hs.desc_ptr(), 32)
    p = S.render_params(320, 240, rect=(10, 20, 300, 200), seed=1234)
    n = (p.x1 - p.x0) * (p.y1 - p.y0)
    d = ctx.malloc(n * S.RAY_DTYPE.itemsize)
    for sample in (0, 3):
        ctx.primary_rays_device(sid, p, sample, d)
        ctx.synchronize()
        got = ctx.to_host(d, n * S.RAY_DTYPE.itemsize, S.RAY_DTYPE)
        want = ref.primary_rays(p, sample)
        assert np.array_equal(got["origin"].view(np.uint32), want["origin"].view(np.uint32))
        assert np.array_equal(got["dir"].view(np.uint32), want["dir"].view(np.uint32)), f"sample {sample}"
        assert np.array_equal(got["time"], want["time"])
    ctx.free(d)
    ctx.release(sid)


def _random_segments(rng, n, lo, hi, length):
    o = rng.uniform(lo, hi, (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    d = (d * length).astype(np.float32)
    return S.make_rays(o, d)


def test_random_segments_spheres(ctx):
    """secondary-like rays: random origins inside the sphere cloud, short and long segments"""
    hs = _scene(0)
    sid = ctx.upload(hs)
    ref = oracle.OracleScene(hs.desc_ptr(), 32)
    rng = np.random.default_rng(7)
    for length in (3.0, 30.0, 1e4):
        rays = _random_segments(rng, 50000, -12, 12, length)
        hits = ctx.trace(sid, rays)
        ohits, ext = ref.trace(rays)
        _compare(hits, ohits, ext, f"spheres random len {length}")
    ctx.release(sid)


def test_any_hit_shadow_rays(ctx):
    """shadow test == Scene::cast_ray(ray) != NULL (any-hit vs closest-hit found)"""
    hs = _scene(0)
    sid = ctx.upload(hs)
    ref = oracle.OracleScene(hs.desc_ptr(), 32)
    p = S.render_params(320, 240)
    rays = ref.primary_rays(p)
    ohits, _ = ref.trace(rays)
    m = ohits["obj"] >= 0
    o = rays["origin"][m] + rays["dir"][m] * ohits["t"][m, None]
    light = np.array([0.0, 20.0, 0.0], np.float32)
    srays = S.make_rays(o.astype(np.float32), (light - o).astype(np.float32))
    got = ctx.trace(sid, srays, any_hit=True)["obj"] >= 0
    want = ref.trace(srays, any_hit=True)[0]["obj"] >= 0
    closest = ref.trace(srays)[0]["obj"] >= 0
    assert np.array_equal(want, closest)
    assert np.array_equal(got, want), f"{(got != want).sum()} shadow decisions differ"
    print(f"shadow rays {len(srays)}, occluded {want.sum()}")
    ctx.release(sid)


def test_heightfield_small_primary_and_random(ctx):
    hs = _scene(1, variant=4)  # 250x125 quads (62,500 triangles), mesh octree 8/32
    sid = ctx.upload(hs)
    ref = oracle.OracleScene(hs.desc_ptr(), 32)
    p = S.render_params(480, 270, iter=6)
    rays = ref.primary_rays(p)
    hits = ctx.trace(sid, rays)
    ohits, ext = ref.trace(rays)
    _compare(hits, ohits, ext, "heightfield/4 primary")
    rng = np.random.default_rng(11)
    rays = _random_segments(rng, 40000, -40, 40, 60.0)
    hits = ctx.trace(sid, rays)
    ohits, ext = ref.trace(rays)
    _compare(hits, ohits, ext, "heightfield/4 random")
    ctx.release(sid)


def test_heightfield_full_sampled(ctx):
    """the full 1M-triangle workload: every 7th primary ray of 1920x1080"""
    hs = _scene(1)
    sid = ctx.upload(hs)
    ref = oracle.OracleScene(hs.desc_ptr(), 32)
    p = S.render_params(1920, 1080, iter=6)
    rays = ref.primary_rays(p)[::7]
    hits, cnt = ctx.trace(sid, rays, counts=True)
    ohits, ext = ref.trace(rays)
    _compare(hits, ohits, ext, "heightfield full primary (1/7)")
    print(f"visits per ray: nodes {cnt.nodes / len(rays):.1f} tris {cnt.triangles / len(rays):.1f} "
          f"(exhaustive checker: nodes {ext['nodes'].mean():.1f} prims {ext['prims'].mean():.1f})")
    ctx.release(sid)


def test_empty_and_degenerate(ctx):
    """empty ray batch, zero-length and axis-aligned rays, rays with zero
    direction components (IEEE 1/0 in the slab test, Q8)"""
    hs = _scene(0)
    sid = ctx.upload(hs)
    ref = oracle.OracleScene(hs.desc_ptr(), 32)
    assert len(ctx.trace(sid, S.make_rays(np.zeros((0, 3), np.float32), np.zeros((0, 3), np.float32)))) == 0
    o = np.array([[0, 0, -30], [0, 0, -30], [0.5, -0.25, -30], [0, 0, 0], [3, 4, 5], [0, 0, -30]], np.float32)
    d = np.array([[0, 0, 1e4], [0, 0, -1e4], [0, 0, 100], [0, 0, 0], [1e4, 0, 0], [-0.0, 0, 1e4]], np.float32)
    rays = S.make_rays(o, d)
    hits = ctx.trace(sid, rays)
    ohits, ext = ref.trace(rays)
    _compare(hits, ohits, ext, "degenerate")
    ctx.release(sid)
