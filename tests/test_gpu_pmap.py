"""Photon maps on the device: GPU-built kd-tree + gathers, against the CPU
checker's brute-force gather on identical photon arrays.

Bar: irradiance / radiance estimates within 1e-4 relative (plus 1e-6
absolute) of the double restatement (src/pmap.cc:107-178 semantics:
dist^2 <= r^2 on float positions with a float radius, >= 8 photons for
irradiance, the cos(pi/2.5) normal filter for radiance); k-nearest mode
with ties broken by photon index against the checker's sorted selection.
"""
import numpy as np
import pytest

import oracle
import sray_amd as S
from scenes import build

pytestmark = pytest.mark.gpu
RTOL, ATOL = 1e-4, 1e-6


def _photons(rng, n, planar=False):
    ph = np.zeros((n, 12), np.float32)
    if planar:
        ph[:, 0] = rng.uniform(-2, 2, n)
        ph[:, 2] = rng.uniform(-2, 2, n)
        ph[:, 1] = 0.0
        d = rng.normal(size=(n, 3)) * 0.3 + (0, -1, 0)
        ph[:, 6:9] = (0, 1, 0)
    else:
        ph[:, 0:3] = rng.uniform(-1, 1, (n, 3))
        d = rng.normal(size=(n, 3))
        nn = rng.normal(size=(n, 3))
        ph[:, 6:9] = nn / np.linalg.norm(nn, axis=1, keepdims=True)
    ph[:, 3:6] = d / np.linalg.norm(d, axis=1, keepdims=True)
    ph[:, 9:12] = rng.uniform(0, 1, (n, 3))
    return ph


def _queries(rng, n, planar=False):
    q = np.zeros((n, 10), np.float32)
    if planar:
        q[:, 0] = rng.uniform(-2, 2, n)
        q[:, 2] = rng.uniform(-2, 2, n)
        q[:, 4] = 1.0
    else:
        q[:, 0:3] = rng.uniform(-1, 1, (n, 3))
        qn = rng.normal(size=(n, 3))
        q[:, 3:6] = qn / np.linalg.norm(qn, axis=1, keepdims=True)
    return q


def _gpu_gather(ctx, pid, q, radius, k, mode):
    qs = np.ascontiguousarray(q, np.float32)
    out = np.zeros((len(qs), 3), np.float32)
    import ctypes as C
    rc = S.lib().sray_gpu_pmap_gather(ctx._h, pid, qs.ctypes.data_as(C.POINTER(S.abi.GatherQuery)), len(qs),
                                      radius, k, mode, out.ctypes.data_as(C.POINTER(C.c_float)))
    ctx._check(rc)
    return out


def _upload(ctx, ph, kind=0, sid=0):
    import ctypes as C
    pid = C.c_uint32()
    arr = np.ascontiguousarray(ph, np.float32)
    ctx._check(S.lib().sray_gpu_pmap_upload(ctx._h, sid, arr.ctypes.data_as(C.POINTER(S.abi.Photon)), len(arr),
                                            kind, C.byref(pid)))
    return pid.value


@pytest.mark.parametrize("planar", [False, True])
@pytest.mark.parametrize("n", [1, 7, 1000, 200_000])
def test_range_gather_matches_checker(ctx, n, planar):
    rng = np.random.default_rng(n + planar)
    ph = _photons(rng, n, planar)
    q = _queries(rng, 2000, planar)
    pid = _upload(ctx, ph)
    b = build()
    ref = oracle.OracleScene(b.ptr(), 64)
    ref.set_photons(0, ph)
    for radius in (0.02, 0.1, 0.4):
        for mode in (0, 1):
            got = _gpu_gather(ctx, pid, q, radius, 0, mode)
            want = ref.gather(0, q, radius, 0, mode)
            assert np.allclose(got, want, rtol=RTOL, atol=ATOL), (n, planar, radius, mode,
                                                                  np.abs(got - want).max())
    S.lib().sray_gpu_pmap_release(ctx._h, pid)


@pytest.mark.parametrize("planar", [False, True])
@pytest.mark.parametrize("n", [1, 1000, 200_000])
def test_grid_gather_matches_checker(ctx, n, planar):
    """the uniform-grid gather (the renderer's shading path) finds the same
    photons as the reference's range query"""
    rng = np.random.default_rng(100 + n + planar)
    ph = _photons(rng, n, planar)
    q = _queries(rng, 2000, planar)
    q[:50, 0:3] = ph[np.arange(50) % n, 0:3]  # queries on photons
    q[50:60, 0:3] = 50.0  # far outside the photons' box
    pid = _upload(ctx, ph)
    b = build()
    ref = oracle.OracleScene(b.ptr(), 64)
    ref.set_photons(0, ph)
    for radius in (0.02, 0.1, 0.4, 3.0):
        for mode in (0, 1):
            got = _gpu_gather(ctx, pid, q, radius, 0, mode | S.abi.SRAY_GATHER_GRID)
            want = ref.gather(0, q, radius, 0, mode)
            assert np.allclose(got, want, rtol=RTOL, atol=ATOL), (n, planar, radius, mode,
                                                                  np.abs(got - want).max())
    S.lib().sray_gpu_pmap_release(ctx._h, pid)


@pytest.mark.parametrize("grid", [0, S.abi.SRAY_GATHER_GRID], ids=["kdtree", "grid"])
@pytest.mark.parametrize("k", [1, 8, 20, 200])
def test_knn_gather_tie_break_by_index(ctx, k, grid):
    """the k nearest by (dist^2, photon index): the kd-tree's quickselect
    and the grid's radix select of the renderer's gathers"""
    rng = np.random.default_rng(k)
    ph = _photons(rng, 50_000)
    # duplicated positions force exact distance ties
    ph[25_000:30_000, 0:3] = ph[0:5_000, 0:3]
    # 600 photons on one point: more than k tie at every distance
    ph[40_000:40_600, 0:3] = ph[7, 0:3]
    q = _queries(rng, 3000)
    q[:200, 0:3] = ph[:200, 0:3]
    q[200:260, 0:3] = ph[7, 0:3] + rng.normal(0, 0.01, (60, 3))
    pid = _upload(ctx, ph)
    b = build()
    ref = oracle.OracleScene(b.ptr(), 64)
    ref.set_photons(0, ph)
    for radius in (0.05, 0.2):
        for mode in (0, 1):
            got = _gpu_gather(ctx, pid, q, radius, k, mode | grid)
            want = ref.gather(0, q, radius, k, mode)
            assert np.allclose(got, want, rtol=RTOL, atol=ATOL), (k, radius, mode, np.abs(got - want).max())
    S.lib().sray_gpu_pmap_release(ctx._h, pid)


def test_empty_map_and_bad_arguments(ctx):
    pid = _upload(ctx, np.zeros((0, 12), np.float32))
    q = _queries(np.random.default_rng(0), 10)
    assert np.all(_gpu_gather(ctx, pid, q, 0.5, 0, 0) == 0)
    with pytest.raises(S.SrayError):
        _gpu_gather(ctx, pid, q, 0.0, 0, 0)
    with pytest.raises(S.SrayError):
        _gpu_gather(ctx, 999999, q, 0.5, 0, 0)


def test_render_with_photon_maps(ctx):
    """room scene with caustic + global maps bound: caustic irradiance at
    Phong hits and diffuse GI rays ending in radiance estimates"""
    hs = S.HostScene.generate(2, variant=10)
    sid = ctx.upload(hs)
    rng = np.random.default_rng(5)
    n = 30_000
    ph = np.zeros((n, 12), np.float32)
    ph[:, 0] = rng.uniform(-5, 5, n)
    ph[:, 2] = rng.uniform(-5, 5, n)
    ph[:, 1] = rng.choice([0.0, 8.0], n)
    d = rng.normal(size=(n, 3)) * 0.2
    d[:, 1] += np.where(ph[:, 1] == 0.0, -1.0, 1.0)
    ph[:, 3:6] = d / np.linalg.norm(d, axis=1, keepdims=True)
    ph[:, 6:9] = 0
    ph[:, 7] = np.where(ph[:, 1] == 0.0, 1.0, -1.0)
    ph[:, 9:12] = rng.uniform(0, 0.05, (n, 3))
    _upload(ctx, ph, S.abi.SRAY_PMAP_CAUSTIC, sid)
    _upload(ctx, ph, S.abi.SRAY_PMAP_GLOBAL, sid)
    p = S.render_params(96, 54, iter=3, caust_photons=1, gi_photons=1, diffuse_samples=2, gather_dist=0.03,
                        seed=17)
    img, st = ctx.render(sid, p)
    for prec in (32, 64):
        ref = oracle.OracleScene(hs.desc_ptr(), prec)
        ref.set_photons(0, ph)
        ref.set_photons(1, ph)
        oimg, ost = ref.render(p)
        e = float(np.sqrt(np.mean((img.astype(np.float64) - oimg) ** 2)))
        print(f"photon render prec {prec}: rmse {e:.3e}; diffuse rays gpu {st.rays_diffuse} cpu {ost.rays_diffuse}")
        assert e <= 1e-3
        if prec == 32:
            assert st.rays_diffuse == ost.rays_diffuse and st.rays_shadow == ost.rays_shadow
    ctx.release(sid)
