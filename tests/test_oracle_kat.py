"""Pins the CPU checker (oracle/) with closed forms, known answers, brute
force and the reference's own invariants, in both precisions.

The reference ships no tests, fixtures or golden images and cannot be built
here (it needs libvmath, kdtree and libimago), so these are the pins:
each one is computed independently of the checker's code (analytically or by
brute force) and chosen so a plausible slip (sign, index, dropped term,
swapped operand) fails it.  tests/golden/kat_closed_form.json holds the
closed-form vectors (written by tests/golden/make_golden.py).
"""
import json
import math
import os

import numpy as np
import pytest

import oracle
import sray_amd as S
from scenes import build, phong, ray, rays

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "kat_closed_form.json")
PRECS = (32, 64)


def trace(b, r, prec, flags=0, any_hit=False):
    o = oracle.OracleScene(b.ptr(), prec, flags)
    h, e = o.trace(r, any_hit=any_hit)
    return h, e


def close(a, b, tol):
    return np.allclose(np.asarray(a, np.float64), np.asarray(b, np.float64), rtol=0, atol=tol)


def golden():
    with open(GOLDEN) as f:
        return json.load(f)


# ---------------------------------------------------------------- RNG
def test_splitmix64_published_values():
    # SplitMix64 seeded with 0: first two outputs (Steele, Lea, Flood 2014)
    assert oracle.splitmix64(0) == 0xE220A8397B1DCDAF
    assert oracle.splitmix64(0x9E3779B97F4A7C15) == 0x6E789E6AA1B965F4


def test_rng_streams_distinct_and_reproducible():
    a = [oracle.rng_u32(1, p, s, 1, 0) for p in range(64) for s in range(4)]
    assert len(set(a)) == len(a)
    assert a == [oracle.rng_u32(1, p, s, 1, 0) for p in range(64) for s in range(4)]


# ---------------------------------------------------------------- primitives
@pytest.mark.parametrize("prec", PRECS)
def test_sphere_closed_forms(prec):
    g = golden()["sphere"]
    b = build(spheres=[(g["center"], g["radius"], 0)])
    for case in g["cases"]:
        h, _ = trace(b, ray(case["o"], case["d"]), prec, oracle.BRUTE_FORCE)
        if case["hit"]:
            assert h["obj"][0] == 0, case["name"]
            assert h["t"][0] == np.float32(case["t"]), (case["name"], h["t"][0], case["t"])
        else:
            assert h["obj"][0] == -1, case["name"]


@pytest.mark.parametrize("prec", PRECS)
def test_cylinder_closed_forms(prec):
    g = golden()["cylinder"]
    b = build(cylinders=[(g["center"], g["radius"], g["start"], g["end"], 0)])
    for case in g["cases"]:
        h, _ = trace(b, ray(case["o"], case["d"]), prec, oracle.BRUTE_FORCE)
        if case["hit"]:
            assert h["obj"][0] == 0, case["name"]
            assert h["t"][0] == np.float32(case["t"]), (case["name"], h["t"][0], case["t"])
        else:
            assert h["obj"][0] == -1, case["name"]


@pytest.mark.parametrize("prec", PRECS)
def test_triangle_closed_forms(prec):
    g = golden()["triangle"]
    b = build(meshes=[([g["tri"]], (0, 0, 0), 0)])
    for case in g["cases"]:
        h, _ = trace(b, ray(case["o"], case["d"]), prec, oracle.BRUTE_FORCE)
        if case["hit"]:
            assert h["obj"][0] == 0 and h["prim"][0] == 0, case["name"]
            assert h["t"][0] == np.float32(case["t"]), (case["name"], h["t"][0], case["t"])
        else:
            assert h["obj"][0] == -1, case["name"]


@pytest.mark.parametrize("prec", PRECS)
def test_shared_edge_is_inclusive_and_tie_goes_to_first(prec):
    """both triangles of a shared edge report the hit (src/mesh.cc:314); the
    first in depth-first order wins with strict '<' (src/octree.inl:211-234)"""
    v0, v1, v2, v3 = (-1, -1, 5), (1, -1, 5), (0, 1, 5), (0, -3, 5)
    for order in ([[v0, v1, v2], [v1, v0, v3]], [[v1, v0, v3], [v0, v1, v2]]):
        b = build(meshes=[(order, (0, 0, 0), 0)])
        h, e = trace(b, ray((0, -1, 0), (0, 0, 10)), prec)
        assert h["obj"][0] == 0 and h["prim"][0] == 0 and h["t"][0] == np.float32(0.5)
        assert e["tie"][0] == 1
        for single in order:
            bs = build(meshes=[([single], (0, 0, 0), 0)])
            assert trace(bs, ray((0, -1, 0), (0, 0, 10)), prec)[0]["obj"][0] == 0


@pytest.mark.parametrize("prec", PRECS)
def test_zero_direction_sign_in_slab_test(prec):
    """the slab test selects bounds with dir < 0 and divides 1/dir (Q8): a
    +0.0 component passes, a -0.0 one gives -inf and fails the box test when
    the origin is strictly inside the slab, so the octree path misses what
    the brute-force path hits"""
    b = build(spheres=[((0, 0, 10), 2, 0)])
    hp, _ = trace(b, ray((0, 0, 0), (0.0, 0, 100)), prec)
    hn, _ = trace(b, ray((0, 0, 0), (-0.0, 0, 100)), prec)
    hb, _ = trace(b, ray((0, 0, 0), (-0.0, 0, 100)), prec, oracle.BRUTE_FORCE)
    assert hp["obj"][0] == 0 and hb["obj"][0] == 0
    assert hn["obj"][0] == -1


@pytest.mark.parametrize("prec", PRECS)
def test_ray_on_box_face_is_nan_miss(prec):
    """a ray lying in a box face plane with a zero direction component makes
    (bound - origin) * (1/0) = 0 * inf = NaN, every comparison with it is
    false and the slab test fails: the tangent ray x = r misses through the
    octree path although the sphere test alone reports the tangent hit"""
    b = build(spheres=[((0, 0, 10), 2, 0)])
    r = ray((2, 0, 0), (0, 0, 100))
    assert trace(b, r, prec)[0]["obj"][0] == -1
    assert trace(b, r, prec, oracle.BRUTE_FORCE)[0]["obj"][0] == 0


# ---------------------------------------------------------------- octree vs brute force
def _random_scene(seed, nsph=120, ntri=1500):
    rng = np.random.default_rng(seed)
    sph = [(tuple(rng.uniform(-10, 10, 3)), float(rng.uniform(0.2, 1.5)), 0) for _ in range(nsph)]
    c = rng.uniform(-8, 8, (ntri, 1, 3))
    tris = (c + rng.normal(scale=0.6, size=(ntri, 3, 3))).astype(np.float32)
    return build(spheres=sph, meshes=[(tris, (0.5, -0.25, 0.125), 0)], scene_oct=(5, 3), mesh_oct=(6, 8))


@pytest.mark.parametrize("prec", PRECS)
def test_octree_equals_brute_force(prec):
    """the reference's own invariant: Scene::cast_ray / Mesh::intersect give
    the same closest hit with or without the octrees (src/scene.cc:771-793,
    src/mesh.cc:271-287)"""
    b = _random_scene(3)
    rng = np.random.default_rng(4)
    n = 20000
    o = rng.uniform(-12, 12, (n, 3))
    d = rng.normal(size=(n, 3))
    d = d / np.linalg.norm(d, axis=1, keepdims=True) * rng.choice([2.0, 20.0, 1e4], (n, 1))
    r = rays(o, d)
    ht, et = trace(b, r, prec)
    hb, eb = trace(b, r, prec, oracle.BRUTE_FORCE)
    assert np.array_equal(ht["t"].view(np.uint32), hb["t"].view(np.uint32))
    nt = (et["tie"] == 0) & (eb["tie"] == 0)
    assert np.array_equal(ht["obj"][nt], hb["obj"][nt])
    assert np.array_equal(ht["prim"][nt], hb["prim"][nt])
    assert (ht["obj"] >= 0).sum() > n // 4
    # any-hit (shadow) == closest-hit found
    ha, _ = trace(b, r, prec, any_hit=True)
    assert np.array_equal(ha["obj"] >= 0, ht["obj"] >= 0)


def test_q9_retraversal_gives_identical_results():
    b = _random_scene(5)
    rng = np.random.default_rng(6)
    r = rays(rng.uniform(-12, 12, (5000, 3)), rng.normal(size=(5000, 3)) * 50)
    a, _ = trace(b, r, 32)
    q, _ = trace(b, r, 32, oracle.Q9_RETRAVERSE)
    assert np.array_equal(a, q)


# ---------------------------------------------------------------- camera
@pytest.mark.parametrize("prec", PRECS)
def test_primary_ray_pixel_corners(prec):
    """px = (x/W)*xsz - xsz/2, py = 1 - (y/H)*2 at the pixel corner, z = 1/tan(fov/2),
    scaled by RAY_MAGNITUDE (src/camera.cc:119-142)"""
    b = build(spheres=[((0, 0, 10), 1, 0)])
    o = oracle.OracleScene(b.ptr(), prec)
    p = S.render_params(640, 480)
    r = o.primary_rays(p)
    W, H = 640, 480
    a = W / H
    fz = 1.0 / math.tan(math.pi / 8)
    for x, y in ((0, 0), (320, 240), (639, 479), (100, 333)):
        want = np.array([((x / W) * 2 * a - a), 1 - (y / H) * 2, fz]) * 1e4
        got = r["dir"][y * W + x].astype(np.float64)
        assert np.allclose(got, want, rtol=1e-6, atol=1e-3), (x, y, got, want)
    c = r["dir"][240 * W + 320]
    assert c[0] == 0.0 and c[1] == 0.0
    assert np.all(r["origin"] == 0)
    j = o.primary_rays(p, sample=3)["dir"].astype(np.float64)
    off = j - r["dir"].astype(np.float64)
    assert np.all(np.abs(off[:, 0]) <= 0.5 * 2 * a / W * 1e4 * (1 + 1e-5))
    assert np.all(np.abs(off[:, 1]) <= 0.5 * 2 / H * 1e4 * (1 + 1e-5))


# ---------------------------------------------------------------- shading
def _center_pixel(b, prec, **kw):
    o = oracle.OracleScene(b.ptr(), prec)
    p = S.render_params(64, 48, rect=(32, 24, 33, 25), **kw)
    img, st = o.render(p, nthreads=1)
    return img[0, 0].astype(np.float64), st


@pytest.mark.parametrize("prec", PRECS)
def test_phong_diffuse_specular_reflect_closed_form(prec):
    """centre ray hits the sphere at normal incidence; light at the eye:
    (amb*kd + ndotl)*kd + (rdotl^shin + refl*fres*bg)*ks with fres = 1 for an
    opaque material (src/shader.cc:142-154, 191-201, 238-246, 266-275)"""
    g = golden()["phong"]
    kd, ks, refl, amb, bg = g["kd"], g["ks"], g["refl"], g["ambient"], g["bg"]
    b = build(spheres=[((0, 0, 10), 2, 0)], materials=[phong(kd=kd, ks=ks, shin=10.0, refl=refl)],
              lights=[((0, 0, 0), (1, 1, 1))], ambient=(*amb, 1), bg=(*bg, 1))
    c, st = _center_pixel(b, prec, iter=3)
    # float: the hit point comes from a cancellation (-10 + 8), its rounding
    # reaches the shading normal and pow(rdotl, 10)
    tol = 1e-7 if prec == 64 else 1e-4
    assert np.allclose(c[:3], g["expected"], rtol=tol, atol=tol), (c, g["expected"])
    assert c[3] == 1.0
    assert st.rays_shadow == 1 and st.rays_reflect == 1


@pytest.mark.parametrize("prec", PRECS)
def test_glass_normal_incidence_closed_form(prec):
    """refraction through a glass sphere along its axis: Schlick r0 = 0.04 at
    both interfaces, weight (1-r0)^2 on the background; the ior stack enters
    and leaves (src/shader.cc:204-229, 285-291)"""
    g = golden()["glass"]
    b = build(spheres=[((0, 0, 10), 2, 0)], materials=[phong(kd=(0, 0, 0), ks=1.0, shin=50, refr=1.0, ior=1.5)],
              lights=[((0, 100, 10), (1, 1, 1))], bg=(*g["bg"], 1))
    c, st = _center_pixel(b, prec, iter=4)
    tol = 1e-7 if prec == 64 else 1e-5
    assert np.allclose(c[:3], g["expected"], rtol=tol, atol=tol), (c, g["expected"])
    assert st.rays_refract == 2 and st.rays_reflect == 0


@pytest.mark.parametrize("prec", PRECS)
def test_energy_cutoff_and_iter(prec):
    """no reflect ray when refl*energy <= min_energy or iter == 0 (src/shader.cc:191-192)"""
    b = build(spheres=[((0, 0, 10), 2, 0)], materials=[phong(refl=0.5)], lights=[((0, 0, 0), (1, 1, 1))])
    assert _center_pixel(b, prec, iter=0)[1].rays_reflect == 0
    assert _center_pixel(b, prec, iter=1, min_energy=0.5)[1].rays_reflect == 0
    assert _center_pixel(b, prec, iter=1, min_energy=0.49)[1].rays_reflect == 1


# ---------------------------------------------------------------- render_block
@pytest.mark.parametrize("prec", PRECS)
def test_adaptive_stop_needs_two_samples(prec):
    """constant samples have zero variance, but the test needs i > 1, so
    every pixel takes max(min_samples, 2) samples (src/render.cc:232-237)"""
    b = build(bg=(0.2, 0.4, 0.6, 1.0))
    o = oracle.OracleScene(b.ptr(), prec)
    for mins, want in ((1, 2), (2, 2), (3, 3)):
        img, st = o.render(S.render_params(20, 10, samples=(mins, 8), max_var=0.01))
        assert st.samples == want * 200
        assert np.allclose(img[..., :3], (0.2, 0.4, 0.6)) and np.all(img[..., 3] == 1.0)


@pytest.mark.parametrize("prec", PRECS)
def test_q1_scaling_pattern(prec):
    """img[x] *= rcp_lut[i] scales float x of each block row (src/render.cc:240):
    with 2 samples everywhere, floats [4*org, 4*org+bw) of each 48-px block
    row are halved, all other floats keep the 2-sample sum"""
    bg = np.array([0.2, 0.4, 0.6, 1.0], np.float32)
    b = build(bg=tuple(bg))
    o = oracle.OracleScene(b.ptr(), prec)
    img, _ = o.render(S.render_params(100, 3, samples=(2, 2), flags=S.abi.SRAY_RENDER_REF_Q1))
    row = img.reshape(3, -1)
    want = np.tile(2 * bg, 100)
    for org, bw in ((0, 48), (48, 48), (96, 4)):
        want[4 * org:4 * org + bw] *= 0.5
    for y in range(3):
        assert np.array_equal(row[y], want.astype(np.float32))


# ---------------------------------------------------------------- gather
def test_gather_range_brute_force_and_threshold():
    """irradiance: sum of photon colours within r whose dir faces the normal,
    / (pi r^2), zero if fewer than 8 photons are in range (src/pmap.cc:109-139);
    radiance: col * (-dir.n) when -dir.n > 0 and norm.n > cos(pi/2.5)"""
    rng = np.random.default_rng(2)
    n = 400
    ph = np.zeros((n, 12), np.float32)
    ph[:, 0:3] = rng.uniform(-1, 1, (n, 3))
    d = rng.normal(size=(n, 3))
    ph[:, 3:6] = d / np.linalg.norm(d, axis=1, keepdims=True)
    nn = rng.normal(size=(n, 3))
    ph[:, 6:9] = nn / np.linalg.norm(nn, axis=1, keepdims=True)
    ph[:, 9:12] = rng.uniform(0, 1, (n, 3))
    b = build()
    o = oracle.OracleScene(b.ptr(), 64)
    o.set_photons(0, ph)
    o.set_photons(1, ph)
    q = np.zeros((50, 10), np.float32)
    q[:, 0:3] = rng.uniform(-1, 1, (50, 3))
    qn = rng.normal(size=(50, 3))
    q[:, 3:6] = qn / np.linalg.norm(qn, axis=1, keepdims=True)
    for r in (0.05, 0.3, 0.6):
        irr = o.gather(0, q, r, mode=0)
        rad = o.gather(1, q, r, mode=1)
        for i in range(50):
            d2 = ((ph[:, :3].astype(np.float64) - q[i, :3]) ** 2).sum(1)
            inr = d2 <= np.float64(np.float32(r)) ** 2
            front = (ph[:, 3:6].astype(np.float64) @ q[i, 3:6]) < 0
            if inr.sum() < 8:
                want = np.zeros(3)
            else:
                want = ph[inr & front, 9:12].sum(0) / (math.pi * r * r)
            assert np.allclose(irr[i], want, rtol=1e-5, atol=1e-6)
            ndl = -(ph[:, 3:6].astype(np.float64) @ q[i, 3:6])
            ok = inr & (ndl > 0) & ((ph[:, 6:9].astype(np.float64) @ q[i, 3:6]) > math.cos(math.pi / 2.5))
            want = (ph[ok, 9:12] * ndl[ok, None]).sum(0) / (math.pi * r * r)
            assert np.allclose(rad[i], want, rtol=1e-5, atol=1e-6)


def test_gather_knn_reduces_to_range():
    """k-NN mode equals range mode whenever at most k photons are in range,
    and keeps the k nearest (ties by photon index) otherwise"""
    rng = np.random.default_rng(9)
    ph = np.zeros((300, 12), np.float32)
    ph[:, 0:3] = rng.uniform(-1, 1, (300, 3))
    ph[:, 3:6] = (0, 0, -1)
    ph[:, 9:12] = 1.0
    b = build()
    o = oracle.OracleScene(b.ptr(), 64)
    o.set_photons(0, ph)
    q = np.zeros((1, 10), np.float32)
    q[0, 5] = 1.0
    full = o.gather(0, q, 0.5, k=0)
    assert np.allclose(o.gather(0, q, 0.5, k=10000), full)
    k20 = o.gather(0, q, 0.5, k=20)
    assert np.allclose(k20[0], 20 / (math.pi * 0.25))


# ---------------------------------------------------------------- surface points
@pytest.mark.parametrize("prec", PRECS)
def test_surface_points_closed_forms(prec):
    """SurfPoint of the winning item: sphere normal (p-c)/r with the
    spherical texture coordinates, cylinder normal (x,0,z)/r, and a
    triangle's interpolated normal and texcoords = barycentric interpolation
    of its vertex attributes (interp_face's inverse bilinear map of the
    degenerate quad v0 v1 v2 v2 reduces to it)"""
    tol = 1e-4 if prec == 32 else 1e-9
    rng = np.random.default_rng(17)
    # sphere
    c, r = np.array([1.0, 2.0, 3.0]), 1.5
    b = build(spheres=[(tuple(c), r, 0)])
    o = oracle.OracleScene(b.ptr(), prec)
    dirs = rng.normal(size=(200, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    targets = c + r * dirs
    origins = targets + 5.0 * dirs  # outside, looking back at the target point
    sp = o.surface(rays(origins, (targets - origins) * 1.5))
    assert close(sp[:, 0:3], targets, 1e-5 if prec == 32 else 1e-6)
    assert close(sp[:, 3:6], dirs, 1e-5 if prec == 32 else 1e-6)
    lp = targets - c
    u = (np.arctan2(lp[:, 2], lp[:, 0]) + math.pi) / (2 * math.pi)
    v = np.arccos(np.clip(lp[:, 1] / r, -1, 1)) / math.pi
    du = np.abs(sp[:, 6] - u)
    du = np.minimum(du, 1 - du)  # the seam at theta = +-pi
    assert du.max() < 1e-5 and np.abs(sp[:, 7] - v).max() < 1e-4
    # cylinder along y, radius 0.8, y in [-1, 2)
    b = build(cylinders=[((0.5, 0.0, -0.5), 0.8, -1.0, 2.0, 0)])
    o = oracle.OracleScene(b.ptr(), prec)
    ang = rng.uniform(0, 2 * math.pi, 100)
    ys = rng.uniform(-0.9, 1.9, 100)
    nrm = np.stack([np.cos(ang), np.zeros(100), np.sin(ang)], 1)
    targets = np.array([0.5, 0.0, -0.5]) + 0.8 * nrm + np.stack([np.zeros(100), ys, np.zeros(100)], 1)
    origins = targets + 4.0 * nrm
    sp = o.surface(rays(origins, (targets - origins) * 1.5))
    assert close(sp[:, 0:3], targets, 1e-5 if prec == 32 else 1e-6)
    assert close(sp[:, 3:6], nrm, 1e-5 if prec == 32 else 1e-6)
    # triangles with vertex normals and texcoords
    tris, nrms, uvs, pts, want_n, want_uv = [], [], [], [], [], []
    for k in range(60):
        v0 = rng.uniform(-1, 1, 3) + np.array([3.0 * (k % 8), 3.0 * (k // 8), 0.0])
        e1, e2 = rng.uniform(-1, 1, 3), rng.uniform(-1, 1, 3)
        tri = np.array([v0, v0 + e1, v0 + e2])
        n = rng.normal(size=(3, 3))
        n /= np.linalg.norm(n, axis=1, keepdims=True)
        t = rng.uniform(0, 1, (3, 2))
        a, bb = rng.uniform(0.1, 0.45, 2)
        w = np.array([1 - a - bb, a, bb])
        p = w @ tri
        tris.append(tri), nrms.append(n), uvs.append(t), pts.append(p)
        ni = w @ n
        want_n.append(ni / np.linalg.norm(ni))
        want_uv.append(w @ t)
    b = build(meshes=[(np.array(tris), (0, 0, 0), 0)], normals=np.array(nrms), texcoords=np.array(uvs))
    o = oracle.OracleScene(b.ptr(), prec, oracle.BRUTE_FORCE)
    pts = np.array(pts)
    gn = np.cross(np.array(tris)[:, 1] - np.array(tris)[:, 0], np.array(tris)[:, 2] - np.array(tris)[:, 0])
    gn /= np.linalg.norm(gn, axis=1, keepdims=True)
    origins = pts + 0.7 * gn
    sp = o.surface(rays(origins, (pts - origins) * 1.5))
    rel = 5e-6  # the rays and the stored positions are float
    assert np.all(np.abs(sp[:, 0:3] - pts) <= rel * (1.0 + np.abs(pts)))
    # expected attributes at the returned hit point, from the float-stored
    # triangle and attributes: barycentric interpolation
    T32 = np.array(tris, np.float32).astype(np.float64)
    N32, U32 = np.array(nrms, np.float32).astype(np.float64), np.array(uvs, np.float32).astype(np.float64)
    for k in range(len(tris)):
        e = np.stack([T32[k, 1] - T32[k, 0], T32[k, 2] - T32[k, 0]], 1)
        ab = np.linalg.lstsq(e, sp[k, 0:3].astype(np.float64) - T32[k, 0], rcond=None)[0]
        w = np.array([1 - ab[0] - ab[1], ab[0], ab[1]])
        n = w @ N32[k]
        assert close(sp[k, 3:6], n / np.linalg.norm(n), 2e-5), k
        assert close(sp[k, 6:8], w @ U32[k], 2e-5), k


@pytest.mark.parametrize("prec", PRECS)
def test_shade_undef_checker_closed_form(prec):
    """an object without a material uses shade_undef (src/shader.cc:36-71):
    checker colour from the spherical texcoords (20 x 10 cells), diffuse 0.1
    + ndotl, Blinn highlight ndoth^50"""
    c0 = np.array([0.7, 0.4, 10.0])
    b = build(spheres=[(tuple(c0), 2.0, -1)], materials=[], lights=[((0, 0, 0), (1, 1, 1))])
    col, st = _center_pixel(b, prec, iter=2)
    lp = np.array([-0.7, -0.4, -math.sqrt(4 - 0.49 - 0.16)])  # local hit point of the +z ray
    u = (math.atan2(lp[2], lp[0]) + math.pi) / (2 * math.pi)
    v = math.acos(lp[1] / 2.0) / math.pi
    x, y = int(u * 20), int(v * 10)
    tex = np.array([1.0, 0.3, 0.2]) if x % 2 == y % 2 else np.array([0.2, 0.3, 1.0])
    n = lp / 2.0
    ldir = np.array([0.0, 0.0, -1.0])  # light at the eye, hit on the axis
    ndotl = max(float(n @ ldir), 0.0)
    want = (0.1 + ndotl) * tex + ndotl ** 50
    tol = 1e-6 if prec == 64 else 1e-5  # the framebuffer is float
    assert np.allclose(col[:3], want, rtol=tol, atol=tol), (col, want)
    assert col[3] == 1.0 and st.rays_shadow == 1


@pytest.mark.parametrize("prec", PRECS)
def test_motion_translation_closed_form(prec):
    """linear translation keys (src/anim.inl:188-261): a sphere moving by
    (3,0,0) over [0,40] ms is met at the position of the ray's time, clamped
    outside the key interval"""
    b = build(spheres=[((0, 0, 10), 2.0, 0)])
    b.desc.objects[0].motion[:] = (3.0, 0.0, 0.0)
    b.desc.objects[0].motion_t0, b.desc.objects[0].motion_t1 = 0, 40
    o = oracle.OracleScene(b.ptr(), prec, oracle.BRUTE_FORCE)
    r = rays([(1.5, 0, 0)] * 5, [(0, 0, 100)] * 5)
    r["time"] = [-10, 0, 20, 40, 70]
    h, _ = o.trace(r)
    off = np.abs(1.5 - np.array([0.0, 0.0, 1.5, 3.0, 3.0]))
    want = (10.0 - np.sqrt(4.0 - off ** 2)) / 100.0
    assert np.allclose(h["t"], want, rtol=1e-6), (h["t"], want)
    assert np.all(h["obj"] == 0)
