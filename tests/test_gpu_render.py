"""Device render parity: sray_gpu_render (wavefront) against the CPU checker's
recursive render_block restatement.

Bar: image RMSE <= 1e-3 in linear [0,1] colour against both the strict
float restatement (identical secondary rays) and the double restatement
(the reference's own precision).  Ray counts per kind must match the float
checker exactly (same spawning decisions).
"""
import numpy as np
import pytest

import oracle
import sray_amd as S

pytestmark = pytest.mark.gpu

RMSE_TOL = 1e-3


def rmse(a, b):
    return float(np.sqrt(np.mean((a.astype(np.float64) - b.astype(np.float64)) ** 2)))


def _check(ctx, hs, p, label, check64=True):
    sid = ctx.upload(hs)
    img, st = ctx.render(sid, p)
    ref = oracle.OracleScene(hs.desc_ptr(), 32)
    oimg, ost = ref.render(p)
    e = rmse(img, oimg)
    msg = (f"{label}: rmse vs float checker {e:.3e}; rays gpu p{st.rays_primary} s{st.rays_shadow} "
           f"r{st.rays_reflect} t{st.rays_refract} / cpu p{ost.rays_primary} s{ost.rays_shadow} "
           f"r{ost.rays_reflect} t{ost.rays_refract}")
    e64 = None
    if check64:
        r64 = oracle.OracleScene(hs.desc_ptr(), 64)
        oimg64, _ = r64.render(p)
        e64 = rmse(img, oimg64)
        msg += f"; rmse vs double checker {e64:.3e}"
    print(msg)
    assert e <= RMSE_TOL, msg
    assert e64 is None or e64 <= RMSE_TOL, msg
    assert st.rays_primary == ost.rays_primary, msg
    assert st.samples == ost.samples, msg
    assert (st.rays_shadow, st.rays_reflect, st.rays_refract) == (ost.rays_shadow, ost.rays_reflect,
                                                                   ost.rays_refract), msg
    ctx.release(sid)
    return img


def test_render_spheres(ctx):
    _check(ctx, S.HostScene.generate(0), S.render_params(640, 480, iter=5), "spheres 640x480 depth 5")


def test_render_heightfield_small(ctx):
    _check(ctx, S.HostScene.generate(1, variant=4), S.render_params(480, 270, iter=6),
           "heightfield/4 480x270 depth 6")


def test_render_room_adaptive_area_light(ctx):
    """adaptive supersampling (1-8 spp, variance 0.05) + box area light with
    4 shadow samples, no photon maps"""
    p = S.render_params(160, 90, samples=(1, 8), max_var=0.05, iter=5, shadow_samples=4, seed=99)
    _check(ctx, S.HostScene.generate(2, variant=10), p, "room/10 160x90 1-8spp")


def test_render_zero_shadow_samples(ctx):
    """shadow_samples 0: area lights take no samples, point lights still one
    (calc_lighting, src/shader.cc:252); shade_undef samples every light once"""
    p = S.render_params(120, 70, iter=3, shadow_samples=0, seed=7)
    _check(ctx, S.HostScene.generate(2, variant=10), p, "room/10 shadow_samples 0")
    _check(ctx, S.HostScene.generate(0), S.render_params(96, 72, iter=3, shadow_samples=0), "spheres shadow_samples 0")


def test_render_q1_as_written(ctx):
    """SRAY_RENDER_REF_Q1 reproduces img[x] *= rcp_lut[i] (src/render.cc:240)"""
    p = S.render_params(200, 100, samples=(2, 4), max_var=0.001, iter=4, flags=S.abi.SRAY_RENDER_REF_Q1)
    _check(ctx, S.HostScene.generate(0), p, "spheres Q1 2-4spp", check64=False)


def test_render_tile_equals_crop(ctx):
    """a tile request gives the same pixels as the crop of the full frame"""
    hs = S.HostScene.generate(0)
    sid = ctx.upload(hs)
    full, _ = ctx.render(sid, S.render_params(320, 240, iter=5))
    tile, _ = ctx.render(sid, S.render_params(320, 240, rect=(64, 32, 200, 150), iter=5))
    assert np.array_equal(tile, full[32:150, 64:200])
    ctx.release(sid)


def test_render_deterministic(ctx):
    hs = S.HostScene.generate(0)
    sid = ctx.upload(hs)
    p = S.render_params(256, 192, samples=(1, 4), max_var=0.01, iter=5, seed=5)
    a, _ = ctx.render(sid, p)
    b, _ = ctx.render(sid, p)
    assert np.array_equal(a, b)
    ctx.release(sid)


def test_render_motion_blur(ctx):
    """moving objects and a camera shutter: sample times drawn per pixel and
    sample (Camera::get_primary_ray, src/camera.cc:137-139), object matrices
    at the ray time"""
    import ctypes as C
    from sray_amd import abi
    hs = S.HostScene.generate(3, variant=300)
    d = abi.SceneDesc()
    C.memmove(C.byref(d), hs.desc_ptr(), C.sizeof(d))
    d.camera.shutter = 40
    p = S.render_params(128, 96, samples=(2, 6), max_var=0.005, iter=3, mblur=1, seed=21)
    sid = ctx.upload(C.pointer(d))
    img, st = ctx.render(sid, p)
    ref = oracle.OracleScene(C.pointer(d), 32)
    oimg, ost = ref.render(p)
    e = rmse(img, oimg)
    print(f"motion blur: rmse {e:.3e}, samples {st.samples}/{ost.samples}")
    assert e <= RMSE_TOL and st.samples == ost.samples and st.rays_primary == ost.rays_primary
    ctx.release(sid)
