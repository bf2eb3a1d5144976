"""Texture maps and bump normals on the device: mapped diffuse, specular,
shininess, reflect, refract and ior attributes (MatAttrib::get_color,
src/material.inl:18-28, read in shade_phong, src/shader.cc:108-133, and
the photon tracers, src/scene.cc:583-605, 677-711) and the "normal"
attribute (get_bump_normal, src/shader.cc:137-138, 293-316).

Bar: renders within image RMSE 1e-3 of the checker's float restatement
with identical ray counts, over the textured scene's PNG / PPM / PGM /
PFM maps (repeat wrap, 5-texel cross), spheres, a cylinder and meshes
with and without tangents; device-shot photon arrays equal the checker's
bit for bit; the served path renders the same image.
"""
import numpy as np
import pytest

import oracle
import sray_amd as S
from scenes import textured_scene
from test_gpu_photons import SEED, _compare

pytestmark = pytest.mark.gpu


def rmse(a, b):
    return float(np.sqrt(np.mean((a.astype(np.float64) - b.astype(np.float64)) ** 2)))


def _counts(st):
    return (st.rays_primary, st.rays_shadow, st.rays_reflect, st.rays_refract, st.samples)


@pytest.mark.parametrize("bump_scale", [1.5, 0.4])
def test_textured_render_matches_checker(ctx, tmp_path, bump_scale):
    hs = textured_scene(str(tmp_path), bump_scale)
    sid = ctx.upload(hs)
    p = S.render_params(320, 240, samples=(1, 4), max_var=0.005, iter=5, seed=13)
    img, st = ctx.render(sid, p)
    oimg, ost = oracle.OracleScene(hs.desc_ptr(), 32).render(p)
    e = rmse(img, oimg)
    print(f"textured render (bump {bump_scale}): rmse {e:.3e}, gpu {_counts(st)} cpu {_counts(ost)}", flush=True)
    assert e <= 1e-3 and _counts(st) == _counts(ost)
    ctx.release(sid)


def test_textured_photons_match_checker(ctx, tmp_path):
    hs = textured_scene(str(tmp_path))
    sid = ctx.upload(hs)
    ids, st = ctx.shoot_photons(sid, 6000, 6000, 300.0, 5, 0, 0, SEED)
    c, g, seg = oracle.OracleScene(hs.desc_ptr(), 32).shoot_photons(S.photon_params(6000, 6000, 300.0, 5, 0, 0, SEED))
    gc = ctx.pmap_download(ids[0]) if ids[0] else np.zeros((0, 12), np.float32)
    gg = ctx.pmap_download(ids[1]) if ids[1] else np.zeros((0, 12), np.float32)
    assert st.segments == seg
    assert len(gc) > 0 and len(gg) > 0
    _compare(gc, c, "textured caustic")
    _compare(gg, g, "textured global")
    ctx.release(sid)


def test_textured_render_with_photon_maps(ctx, tmp_path):
    hs = textured_scene(str(tmp_path))
    sid = ctx.upload(hs)
    ctx.shoot_photons(sid, 8000, 8000, 300.0, 5, 0, 0, SEED)
    p = S.render_params(128, 96, iter=3, caust_photons=8000, gi_photons=8000, diffuse_samples=2,
                        gather_dist=0.05, seed=19)
    img, st = ctx.render(sid, p)
    ref = oracle.OracleScene(hs.desc_ptr(), 32)
    ref.shoot_photons(S.photon_params(8000, 8000, 300.0, 5, 0, 0, SEED))
    oimg, ost = ref.render(p)
    e = rmse(img, oimg)
    print(f"textured render with photon maps: rmse {e:.3e}, diffuse rays {st.rays_diffuse}", flush=True)
    assert e <= 1e-3 and st.rays_diffuse == ost.rays_diffuse
    ctx.release(sid)


def test_textured_served_request(ctx, tmp_path):
    hs = textured_scene(str(tmp_path))
    sid = ctx.upload(hs)
    p = S.render_params(160, 120, samples=(1, 2), iter=4, seed=23)
    t = ctx.submit(sid, p)
    img, res = ctx.poll(t)
    oimg, _ = oracle.OracleScene(hs.desc_ptr(), 32).render(p)
    e = rmse(img, oimg)
    print(f"textured served request: rmse {e:.3e}", flush=True)
    assert e <= 1e-3
    ctx.release(sid)
