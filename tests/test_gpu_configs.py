"""Parity at the workloads' own sizes, and for batched and served requests.

Every image is compared with the CPU checker's render (the strict float
restatement of render_block / trace_ray / shade_phong, which spawns the
same secondary rays): RMSE <= 1e-3 in linear [0,1] colour and identical
ray counts per kind and sample counts.  Photon-mapped renders give the
checker the device's own photon arrays (downloaded), so the gathers work
on identical photons.  Batched and served requests are checked against
the checker, not against the engine's own solo renders.
"""
import ctypes as C
import os
import time

import numpy as np
import pytest

import oracle
import sray_amd as S
from scenes import build, desc_copy, look_at, phong

pytestmark = pytest.mark.gpu
RMSE_TOL = 1e-3
THREADS = min(16, len(os.sched_getaffinity(0)))


def rmse(a, b):
    return float(np.sqrt(np.mean((a.astype(np.float64) - b.astype(np.float64)) ** 2)))


def _checker(desc_ptr, photons=None, prec=32):
    ref = oracle.OracleScene(desc_ptr, prec)
    for kind, ph in enumerate(photons or ()):
        if ph is not None and len(ph):
            ref.set_photons(kind, ph)
    return ref


def _compare(img, st, ref, p, label):
    """st: the render's sray_stats (rays per kind and samples), or a
    request's total ray count (sray_result.rays)"""
    t0 = time.perf_counter()
    oimg, ost = ref.render(p, nthreads=THREADS)
    e = rmse(img, oimg)
    if isinstance(st, int):
        g, c = st, ost.rays()
    else:
        g = (st.rays_primary, st.rays_shadow, st.rays_reflect, st.rays_refract, st.rays_diffuse, st.samples)
        c = (ost.rays_primary, ost.rays_shadow, ost.rays_reflect, ost.rays_refract, ost.rays_diffuse, ost.samples)
    msg = f"{label}: rmse {e:.3e}, rays/samples gpu {g} cpu {c} (checker {time.perf_counter() - t0:.1f} s)"
    print(msg, flush=True)
    assert e <= RMSE_TOL, msg
    assert g == c, msg


def _photons_of(ctx, ids):
    return [ctx.pmap_download(i) if i else None for i in ids]


def test_heightfield_full_frame(ctx):
    """the traversal-bound workload as configured: 1M triangles + 40
    spheres, mesh octree depth 8 / 32 items, 1920x1080, depth 6"""
    hs = S.HostScene.generate(1)
    assert hs.desc.num_triangles == 1_000_000 and hs.desc.mesh_oct_items == 32 and hs.desc.mesh_oct_depth == 8
    sid = ctx.upload(hs)
    p = S.render_params(1920, 1080, iter=6)
    img, st = ctx.render(sid, p)
    _compare(img, st, _checker(hs.desc_ptr()), p, "heightfield 1920x1080 depth 6")
    ctx.release(sid)


def test_room_full_photon_maps(ctx):
    """the photon-mapped room as configured: 50k-triangle icosphere, 2M
    global + 500k caustic photons shot on the device, radius 0.1, 1280x720,
    adaptive 1-16 spp (threshold 0.05), 4 area-light shadow samples and one
    diffuse ray per hit"""
    hs = S.HostScene.generate(2)
    sid = ctx.upload(hs)
    ids, pst = ctx.shoot_photons(sid, 500_000, 2_000_000, 300.0, 5, 0, 0, 0x5EED)
    photons = _photons_of(ctx, ids)
    print(f"photons stored: caustic {len(photons[0])}, global {len(photons[1])}", flush=True)
    diag = S.scene_diag(hs.desc_ptr())
    p = S.render_params(1280, 720, samples=(1, 16), max_var=0.05, iter=5, shadow_samples=4, diffuse_samples=1,
                        caust_photons=500_000, gi_photons=2_000_000, gather_dist=0.1 / diag, seed=100)
    img, st = ctx.render(sid, p)
    _compare(img, st, _checker(hs.desc_ptr(), photons), p, "room 1280x720 1-16 spp, 2M+500k photons")
    ctx.release(sid)


def test_batched_requests_match_checker(ctx):
    """the batched workload's shape, scaled down: scenes of the batch
    generator (displaced icospheres + spheres, 10% moving), resolutions
    64..192, half with device-shot photon maps, motion blur (shutter 40),
    all in flight together; each image against the checker"""
    rng = np.random.default_rng(0x5EED0004)
    jobs, refs, keep = [], [], []
    for k in range(10):
        ntri = int(np.exp(rng.uniform(np.log(1000), np.log(20_000))))
        hs = S.HostScene.generate(3, seed=0x5EED0400 + k, variant=ntri)
        d = desc_copy(hs.desc_ptr(), shutter=40)
        keep += [hs, d]
        sid = ctx.upload(C.pointer(d))
        res = 64 * int(rng.integers(1, 4))
        kw, photons = {}, None
        if k % 2 == 0:
            nph = int(rng.integers(20_000, 100_001))
            ids, _ = ctx.shoot_photons(sid, nph // 5, nph - nph // 5, 300.0, 5, 0, 40, 0x5EED + k)
            photons = _photons_of(ctx, ids)
            kw = dict(caust_photons=1, gi_photons=1, gather_dist=0.1 / S.scene_diag(C.pointer(d)))
        p = S.render_params(res, res, samples=(2, 6), max_var=0.01, iter=4, mblur=1, frame_time=20, seed=k, **kw)
        jobs.append((sid, p, None, None))
        refs.append((_checker(C.pointer(d), photons), p, sid))
    tickets = ctx.submit_many(jobs)
    for k, t in enumerate(tickets):
        img, res = ctx.poll(t)
        ref, p, sid = refs[k]
        _compare(img, int(res.rays), ref, p, f"batched request {k} ({p.x1}x{p.y1})")
    for _, _, sid in refs:
        ctx.release(sid)


def test_served_requests_match_checker(ctx):
    """the served workload's shape, scaled down: resident photon-mapped
    rooms of the room generator, requests with random camera poses on a
    shell around the room centre and tiles of 64x64, 128x128 and 320x180,
    arriving while earlier ones are on the GPU; each image against the
    checker rendering the same request alone"""
    rng = np.random.default_rng(0x5EED0005)
    scenes = []
    for k in range(3):
        hs = S.HostScene.generate(2, seed=k)
        sid = ctx.upload(hs)
        ids, _ = ctx.shoot_photons(sid, 25_000, 100_000, 300.0, 5, 0, 0, 0x5EED + k)
        scenes.append((hs, sid, _photons_of(ctx, ids), S.scene_diag(hs.desc_ptr())))
    sizes = [(64, 64), (128, 128), (320, 180)]
    reqs = []
    for k in range(12):
        sc = int(rng.integers(0, len(scenes)))
        w, h = sizes[int(rng.choice(3, p=[0.5, 0.3, 0.2]))]
        v = rng.normal(size=3)
        v /= np.linalg.norm(v)
        cam = look_at(np.array([0.0, 4.0, 0.0]) + 3.5 * v, (0.0, 4.0, 0.0), 60.0)
        hs, sid, photons, diag = scenes[sc]
        p = S.render_params(w, h, samples=(1, 16), max_var=0.05, iter=5, shadow_samples=4, caust_photons=1,
                            gi_photons=1, gather_dist=0.1 / diag, seed=k)
        reqs.append((sid, p, cam, sc))
    tickets = []
    for k, (sid, p, cam, sc) in enumerate(reqs):
        tickets.append(ctx.submit(sid, p, cam))
        time.sleep(float(rng.exponential(0.002)))
    for k, t in enumerate(tickets):
        img, res = ctx.poll(t)
        sid, p, cam, sc = reqs[k]
        hs, _, photons, _ = scenes[sc]
        d = desc_copy(hs.desc_ptr(), camera=cam)
        _compare(img, int(res.rays), _checker(C.pointer(d), photons), p, f"served request {k} ({p.x1}x{p.y1})")
    for _, sid, _, _ in scenes:
        ctx.release(sid)


def _cylinder_scene():
    cyl = [((-3.0, 0.0, 4.0), 1.0, -2.0, 2.0, 1), ((0.5, -1.0, 6.0), 0.6, 0.0, 3.5, 2),
           ((3.0, 0.5, 3.0), 1.4, -1.0, 1.0, 0), ((-1.0, 2.0, 9.0), 2.0, -3.0, -0.5, 1)]
    sph = [((1.5, -1.5, 4.0), 0.8, 2), ((-2.0, 2.5, 6.5), 0.7, 0)]
    mats = [phong(kd=(0.8, 0.3, 0.2), ks=0.5, shin=30.0), phong(kd=(0.2, 0.6, 0.9), ks=1.0, shin=60.0, refl=0.5),
            phong(kd=(0.1, 0.1, 0.1), ks=1.0, shin=80.0, refl=0.1, refr=0.9, ior=1.5)]
    return build(spheres=sph, cylinders=cyl, materials=mats, lights=[((0.0, 8.0, -4.0), (1.0, 1.0, 1.0))],
                 camera=(0.0, 0.5, -6.0), ambient=(0.05, 0.05, 0.05, 1.0), bg=(0.1, 0.1, 0.2, 1.0),
                 scene_oct=(3, 1))


def test_cylinders(ctx):
    """Cylinder::intersect (src/cylinder.cc:125-185) on the device: hit IDs
    and t of random segments bit-exact on every traversal instance, then a
    render with mirror and glass materials against the checker"""
    b = _cylinder_scene()
    sid = ctx.upload(b.ptr())
    ref = oracle.OracleScene(b.ptr(), 32)
    rng = np.random.default_rng(41)
    n = 40000
    o = rng.uniform(-5, 5, (n, 3)) + np.array([0.0, 0.0, 5.0])
    d = rng.normal(size=(n, 3)) * rng.choice([3.0, 30.0, 1e4], (n, 1))
    from scenes import rays
    r = rays(o, d)
    oh, _ = ref.trace(r)
    cyl_hits = int(((oh["obj"] >= 2) & (oh["obj"] < 6)).sum())
    assert cyl_hits > 1000
    for kernel in ("plain", "wave", "sparse"):
        h = ctx.trace(sid, r, kernel=kernel)
        assert np.array_equal(h["obj"], oh["obj"]), kernel
        m = oh["obj"] >= 0
        assert np.array_equal(h["t"][m].view(np.uint32), oh["t"][m].view(np.uint32)), kernel
    print(f"cylinders: {cyl_hits} of {n} segments end on a cylinder", flush=True)
    p = S.render_params(320, 240, samples=(1, 4), max_var=0.01, iter=5, seed=3)
    img, st = ctx.render(sid, p)
    _compare(img, st, ref, p, "cylinders 320x240 1-4 spp")
    ctx.release(sid)


@pytest.mark.parametrize("config,variant,size", [(0, 0, (640, 480)), (2, 10, (320, 180))])
def test_xml_scene_renders_like_the_generated_one(ctx, config, variant, size):
    """a scene written as reference-format XML and loaded back through
    sray_scene_load_xml renders on the device like the checker's render of
    the scene it was written from"""
    hs = S.HostScene.generate(config, variant=variant)
    loaded = S.HostScene.load_xml(text=hs.to_xml())
    assert loaded.desc.num_objects == hs.desc.num_objects
    assert loaded.desc.num_triangles == hs.desc.num_triangles
    sid = ctx.upload(loaded)
    W, H = size
    p = S.render_params(W, H, samples=(1, 4), max_var=0.05, iter=5, shadow_samples=2, seed=9)
    img, st = ctx.render(sid, p)
    _compare(img, st, _checker(hs.desc_ptr()), p, f"XML-loaded config {config} {W}x{H}")
    ctx.release(sid)
