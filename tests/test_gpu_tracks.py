"""Keyframed transforms on the device: objects rotating, scaling and turning
about a pivot, a moving light and a moving target camera, read from the
per-msec rows at each ray's time (Object::intersect at ray.time,
src/sphere.cc:88; Camera::get_primary_ray, src/camera.cc:141;
Light::get_point(tm), src/shader.cc:258).

Bar: hit IDs and t bit-exact against the checker at random ray times on
every traversal instance; renders with motion blur within image RMSE 1e-3
with identical ray counts; device-shot photon arrays equal the checker's
bit for bit; ray times outside the rows' interval are refused.
"""
import numpy as np
import pytest

import oracle
import sray_amd as S
from scenes import animated_scene_xml, rays
from test_gpu_photons import SEED, _compare

pytestmark = pytest.mark.gpu


def _scene(msec=20):
    hs = S.HostScene.load_xml(text=animated_scene_xml(40))
    hs.set_frame(msec)
    return hs


def rmse(a, b):
    return float(np.sqrt(np.mean((a.astype(np.float64) - b.astype(np.float64)) ** 2)))


def test_trace_at_ray_times(ctx):
    hs = _scene()
    sid = ctx.upload(hs)
    ref = oracle.OracleScene(hs.desc_ptr(), 32)
    rng = np.random.default_rng(5)
    n = 60000
    o = rng.uniform(-5, 5, (n, 3)) + np.array([0.0, 1.0, 0.0])
    r = rays(o, rng.normal(size=(n, 3)) * rng.choice([4.0, 40.0], (n, 1)))
    r["time"] = rng.integers(-5, 46, n)  # also outside the rows (nearest row)
    oh, ext = ref.trace(r)
    for kernel in ("plain", "wave", "sparse"):
        h = ctx.trace(sid, r, kernel=kernel)
        bad = np.flatnonzero((h["obj"] != oh["obj"]) | (h["prim"] != oh["prim"]) |
                             ((oh["obj"] >= 0) & (h["t"].view(np.uint32) != oh["t"].view(np.uint32))))
        print(f"animated trace [{kernel}]: {n} rays, {(oh['obj'] >= 0).sum()} hits, {len(bad)} mismatches", flush=True)
        assert len(bad) == 0, f"ray {bad[0]}: gpu {h[bad[0]]} cpu {oh[bad[0]]}"
    ctx.release(sid)


@pytest.mark.parametrize("mblur", [0, 1])
def test_render_with_keyframes(ctx, mblur):
    hs = _scene()
    sid = ctx.upload(hs)
    p = S.render_params(320, 240, samples=(2, 6), max_var=0.005, iter=5, mblur=mblur, frame_time=20, seed=31)
    img, st = ctx.render(sid, p)
    oimg, ost = oracle.OracleScene(hs.desc_ptr(), 32).render(p)
    e = rmse(img, oimg)
    g = (st.rays_primary, st.rays_shadow, st.rays_reflect, st.rays_refract, st.samples)
    c = (ost.rays_primary, ost.rays_shadow, ost.rays_reflect, ost.rays_refract, ost.samples)
    print(f"animated render mblur {mblur}: rmse {e:.3e}, gpu {g} cpu {c}", flush=True)
    assert e <= 1e-3 and g == c
    ctx.release(sid)


def test_frames_differ_and_match(ctx):
    """two frames of the animation: different images, each the checker's"""
    imgs = []
    for msec in (0, 200):
        hs = _scene(msec)
        sid = ctx.upload(hs)
        p = S.render_params(160, 120, samples=(1, 1), iter=4, frame_time=msec, seed=2)
        img, st = ctx.render(sid, p)
        oimg, _ = oracle.OracleScene(hs.desc_ptr(), 32).render(p)
        assert rmse(img, oimg) <= 1e-3
        imgs.append(img)
        ctx.release(sid)
    assert rmse(imgs[0], imgs[1]) > 1e-2


def test_photons_with_keyframes(ctx):
    hs = _scene()
    sid = ctx.upload(hs)
    ids, st = ctx.shoot_photons(sid, 4000, 8000, 300.0, 5, 0, 40, SEED)
    c, g, seg = oracle.OracleScene(hs.desc_ptr(), 32).shoot_photons(
        S.photon_params(4000, 8000, 300.0, 5, 0, 40, SEED))
    gc = ctx.pmap_download(ids[0]) if ids[0] else np.zeros((0, 12), np.float32)
    gg = ctx.pmap_download(ids[1]) if ids[1] else np.zeros((0, 12), np.float32)
    assert st.segments == seg
    _compare(gc, c, "animated caustic")
    _compare(gg, g, "animated global")
    ctx.release(sid)


def test_times_outside_the_rows_are_refused(ctx):
    hs = _scene()
    sid = ctx.upload(hs)
    with pytest.raises(S.SrayError):
        ctx.render(sid, S.render_params(32, 32, frame_time=100))
    with pytest.raises(S.SrayError):
        ctx.render(sid, S.render_params(32, 32, frame_time=5, mblur=1))  # 5 - 20 < 0
    with pytest.raises(S.SrayError):
        ctx.shoot_photons(sid, 100, 100, 300.0, 5, 0, 80, SEED)
    ctx.render(sid, S.render_params(32, 32, frame_time=20, mblur=1))
    ctx.release(sid)
