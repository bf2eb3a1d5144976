"""The photon gather that allocates nothing (sray_gpu_pmap_gather_device_noalloc)
against the gather it is a variant of (sray_gpu_pmap_gather, k = 0), through
the C ABI: bit-identical results on two calls in a row, the second with
fewer queries (the per-GPU scratch keeps the first call's contents), for
both estimates; sizes beyond the scratch's maximum are refused."""
import ctypes as C

import numpy as np
import pytest

import sray_amd as S
from sray_amd import abi
from test_gpu_pmap import _gpu_gather, _photons, _queries, _upload

pytestmark = pytest.mark.gpu


def _noalloc(ctx, pid, q, radius, mode):
    qs = np.ascontiguousarray(q, np.float32)
    dq = ctx.to_device(qs)
    dout = ctx.malloc(len(qs) * 12)
    rc = S.lib().sray_gpu_pmap_gather_device_noalloc(ctx._h, pid, C.c_void_p(dq), len(qs), radius, mode,
                                                     C.c_void_p(dout))
    ctx._check(rc)
    ctx.synchronize()
    out = ctx.to_host(dout, len(qs) * 12, np.float32).reshape(-1, 3)
    ctx.free(dq)
    ctx.free(dout)
    return out


@pytest.mark.parametrize("mode", [abi.SRAY_GATHER_IRRADIANCE, abi.SRAY_GATHER_RADIANCE])
def test_noalloc_gather_equals_gather(ctx, mode):
    rng = np.random.default_rng(31 + mode)
    ph = _photons(rng, 60_000)
    pid = _upload(ctx, ph)
    radius = 0.08
    for n in (20_000, 3_000):  # the second call is smaller
        q = _queries(rng, n)
        want = _gpu_gather(ctx, pid, q, radius, 0, mode)
        got = _noalloc(ctx, pid, q, radius, mode)
        print(f"mode {mode}: {n} queries, {int((want != 0).any(axis=1).sum())} non-zero estimates", flush=True)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), f"mode {mode}, n {n}"
    ctx.pmap_release(pid)


def test_noalloc_gather_limits(ctx):
    rng = np.random.default_rng(5)
    pid = _upload(ctx, _photons(rng, 100))
    big = S.abi.SRAY_NOALLOC_MAX_QUERIES + 1
    # refused on the host before any device work (the pointers are never read)
    rc = S.lib().sray_gpu_pmap_gather_device_noalloc(ctx._h, pid, C.c_void_p(16), big, 0.1,
                                                     abi.SRAY_GATHER_IRRADIANCE, C.c_void_p(16))
    assert rc == abi.SRAY_ERR_LIMIT
    assert b"at most" in S.lib().sray_gpu_last_error(ctx._h)
    rc = S.lib().sray_gpu_pmap_gather_device_noalloc(ctx._h, pid, C.c_void_p(16), 4, 0.1,
                                                     abi.SRAY_GATHER_IRRADIANCE | abi.SRAY_GATHER_GRID,
                                                     C.c_void_p(16))
    assert rc == abi.SRAY_ERR_ARG
    assert S.lib().sray_gpu_pmap_gather_device_noalloc(ctx._h, pid, None, 0, 0.1, abi.SRAY_GATHER_IRRADIANCE,
                                                       None) == abi.SRAY_OK
    ctx.pmap_release(pid)
