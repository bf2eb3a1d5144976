"""Photon shooting on the device (sray_gpu_shoot_photons) against the CPU
checker's photon tracer, and renders from device-shot maps.

Bar: the stored photon arrays (storage order = emission order) equal the
strict float restatement's bit for bit — the emission draws, traversal,
surface points and power bookkeeping are the same IEEE float operations;
renders from the maps within image RMSE 1e-3 of the checker's render from
its own maps.
"""
import numpy as np
import pytest

import oracle
import sray_amd as S
from scenes import build, material, plane

pytestmark = pytest.mark.gpu
SEED = 0xC0FFEE


def _compare(got, want, label):
    assert got.shape == want.shape, f"{label}: gpu {got.shape[0]} photons, checker {want.shape[0]}"
    diff = np.flatnonzero(np.any(got.view(np.uint32) != want.view(np.uint32), axis=1))
    print(f"{label}: {len(got)} photons, {len(diff)} differ")
    if len(diff):
        i = diff[0]
        raise AssertionError(f"{label}: {len(diff)} photons differ; first {i}: gpu {got[i]} cpu {want[i]}")


def _shoot_both(ctx, desc_ptr, sid, caust, gi, iter=5, t=(0, 0)):
    ids, st = ctx.shoot_photons(sid, caust, gi, 300.0, iter, t[0], t[1], SEED)
    ref = oracle.OracleScene(desc_ptr, 32)
    c, g, seg = ref.shoot_photons(S.photon_params(caust, gi, 300.0, iter, t[0], t[1], SEED))
    gc = ctx.pmap_download(ids[0]) if ids[0] else np.zeros((0, 12), np.float32)
    gg = ctx.pmap_download(ids[1]) if ids[1] else np.zeros((0, 12), np.float32)
    assert st.stored[0] == len(gc) and st.stored[1] == len(gg)
    assert st.segments == seg, (st.segments, seg)
    return ids, st, (gc, gg), (c, g), ref


def test_plane_scenes_match_checker(ctx):
    """the closed-form scenes of the CPU suite: mirror, glass, coloured floor"""
    W = 1000.0
    cases = {
        "mirror": ([(plane(0.0, W), (0, 0, 0), 0), (plane(2.0, W, up=False), (0, 0, 0), 1)],
                   [material(ks=1.0, refl=1.0), material()]),
        "glass": ([(plane(0.0, W), (0, 0, 0), 0), (plane(-1.5, W), (0, 0, 0), 1)],
                  [material(ks=1.0, refr=1.0, ior=1.5), material()]),
        "diffuse": ([(plane(0.0, W), (0, 0, 0), 0), (plane(2.0, W, up=False), (0, 0, 0), 1)],
                    [material(kd=(0.9, 0.6, 0.3)), material()]),
    }
    for name, (meshes, mats) in cases.items():
        b = build(meshes=meshes, materials=mats, lights=[((0, 1.0, 0), (1, 0.5, 0.25))])
        sid = ctx.upload(b.ptr())
        ids, st, (gc, gg), (c, g), _ = _shoot_both(ctx, b.ptr(), sid, 3000, 3000)
        _compare(gc, c, f"{name} caustic")
        _compare(gg, g, f"{name} global")
        for pid in ids:
            if pid:
                ctx.pmap_release(pid)
        ctx.release(sid)


def test_room_photons_match_checker(ctx):
    """photon room: box area light, glass spheres, diffuse walls; both maps"""
    hs = S.HostScene.generate(2, variant=10)
    sid = ctx.upload(hs)
    ids, st, (gc, gg), (c, g), _ = _shoot_both(ctx, hs.desc_ptr(), sid, 20000, 20000)
    _compare(gc, c, "room caustic")
    _compare(gg, g, "room global")
    assert st.shot[0] >= 20000 and st.shot[1] >= 20000 and len(gc) > 0 and len(gg) > 0
    ctx.release(sid)


def test_sphere_and_box_lights_with_motion(ctx):
    """sphere + box lights, a moving object and a frame interval: emission
    time draws, projection maps over both interval ends"""
    hs = S.HostScene.generate(3, variant=300)
    sid = ctx.upload(hs)
    _, _, (gc, gg), (c, g), _ = _shoot_both(ctx, hs.desc_ptr(), sid, 5000, 5000, t=(0, 1000))
    _compare(gc, c, "batch caustic")
    _compare(gg, g, "batch global")
    ctx.release(sid)


def test_shoot_is_deterministic_and_rebinds(ctx):
    hs = S.HostScene.generate(2, variant=10)
    sid = ctx.upload(hs)
    ids1, _ = ctx.shoot_photons(sid, 4000, 4000, 300.0, 5, 0, 0, 7)
    ids2, _ = ctx.shoot_photons(sid, 4000, 4000, 300.0, 5, 0, 0, 7)
    for a, b in zip(ids1, ids2):
        assert a != b
        assert np.array_equal(ctx.pmap_download(a), ctx.pmap_download(b))
    ids0, st = ctx.shoot_photons(sid, 0, 0)
    assert ids0 == (0, 0) and st.shot[0] == 0
    ctx.release(sid)


def test_render_from_device_photons(ctx):
    """shoot on the device, render with both maps bound; the checker shoots
    and renders from its own maps"""
    hs = S.HostScene.generate(2, variant=10)
    sid = ctx.upload(hs)
    ctx.shoot_photons(sid, 20000, 20000, 300.0, 5, 0, 0, SEED)
    p = S.render_params(96, 54, iter=3, caust_photons=20000, gi_photons=20000, diffuse_samples=2,
                        gather_dist=0.03, seed=17)
    img, st = ctx.render(sid, p)
    ref = oracle.OracleScene(hs.desc_ptr(), 32)
    ref.shoot_photons(S.photon_params(20000, 20000, 300.0, 5, 0, 0, SEED))
    oimg, ost = ref.render(p)
    e = float(np.sqrt(np.mean((img.astype(np.float64) - oimg) ** 2)))
    print(f"render from device photons: rmse {e:.3e}, mean {img.mean():.4f}, diffuse rays {st.rays_diffuse}")
    assert e <= 1e-3
    assert st.rays_diffuse == ost.rays_diffuse
    ctx.release(sid)


def test_render_with_k_nearest(ctx):
    """gather_k: the k nearest photons within the radius in the render's
    gathers (k = 200 as the photon room workload), against the checker's
    render with the same k; a k no gather reaches renders the range
    image bit for bit"""
    hs = S.HostScene.generate(2, variant=10)
    sid = ctx.upload(hs)
    ctx.shoot_photons(sid, 20000, 20000, 300.0, 5, 0, 0, SEED)
    ref = oracle.OracleScene(hs.desc_ptr(), 32)
    ref.shoot_photons(S.photon_params(20000, 20000, 300.0, 5, 0, 0, SEED))
    imgs = {}
    for k in (0, 10, 200, 1 << 30):
        p = S.render_params(96, 54, iter=3, caust_photons=20000, gi_photons=20000, diffuse_samples=2,
                            gather_dist=0.08, seed=17, gather_k=k)
        img, st = ctx.render(sid, p)
        imgs[k] = img
        if k in (10, 200):
            oimg, ost = ref.render(p)
            e = float(np.sqrt(np.mean((img.astype(np.float64) - oimg) ** 2)))
            print(f"render with k = {k}: rmse {e:.3e}, mean {img.mean():.4f}", flush=True)
            assert e <= 1e-3 and st.rays_diffuse == ost.rays_diffuse
    assert np.array_equal(imgs[1 << 30], imgs[0])
    assert not np.array_equal(imgs[10], imgs[0])  # the cap binds
    ctx.release(sid)
