"""Tuning knobs and the autotuner: they change speed, not images.

Bar: renders under every tuning equal the default tuning's render to 1e-6
(sub-chunking of a bounce level may split one sample's rays over two
sub-chunks, which changes the order of two float additions); the
autotuner returns a supported tuning and keeps it in the context.
"""
import numpy as np
import pytest

import oracle
import sray_amd as S

pytestmark = pytest.mark.gpu
TOL = 1e-6


def _variants():
    return [dict(trace_block=128), dict(trace_block=256), dict(tile_w=8), dict(tile_w=16),
            dict(chunk_log2=16), dict(gather_lanes=8), dict(gather_lanes=32), dict(gather_cell=0.25),
            dict(gather_cell=0.5), dict(gather_cell=2.0), dict(graphs=0), dict(persistent=1), dict(persistent=2), dict(sparse_from=64), dict(sparse_from=2)]


def test_tunings_give_the_same_image(ctx):
    hs = S.HostScene.generate(2, variant=10)
    sid = ctx.upload(hs)
    ctx.shoot_photons(sid, 20000, 20000, 300.0, 5, 0, 0, 3)
    p = S.render_params(160, 90, samples=(1, 4), max_var=0.05, iter=4, shadow_samples=2, caust_photons=1,
                        gi_photons=1, gather_dist=0.03, seed=8)
    default = ctx.tuning()
    base, bst = ctx.render(sid, p)
    for kw in _variants():
        ctx.set_tuning(**kw)
        img, st = ctx.render(sid, p)
        ctx.set_tuning(**{k: getattr(default, k) for k in kw})
        err = float(np.abs(img - base).max())
        print(f"{kw}: max diff {err:.2e}")
        assert err <= TOL, kw
        assert st.rays() == bst.rays(), kw
    ctx.release(sid)


def test_sub_chunks_match_checker(ctx):
    """chunks of 2^16 rays split the heightfield's bounce levels into
    sub-chunks (depth first); the image still matches the checker"""
    hs = S.HostScene.generate(1, variant=4)
    sid = ctx.upload(hs)
    p = S.render_params(480, 270, iter=6)
    ctx.set_tuning(chunk_log2=16)
    img, st = ctx.render(sid, p)
    ctx.set_tuning(chunk_log2=21)
    oimg, ost = oracle.OracleScene(hs.desc_ptr(), 32).render(p)
    e = float(np.sqrt(np.mean((img.astype(np.float64) - oimg) ** 2)))
    assert e <= 1e-3 and st.rays() == ost.rays(), e
    ctx.release(sid)


@pytest.mark.parametrize("sparse_from", [0, 1])
def test_sparse_levels_match_checker(ctx, sparse_from):
    """the sparse traversal instances (two triangles in flight, the world
    ray in registers; from bounce level sparse_from on) give the heightfield
    image of the dense instances bit for bit, and match the checker"""
    hs = S.HostScene.generate(1, variant=4)
    sid = ctx.upload(hs)
    p = S.render_params(480, 270, iter=6)
    default = ctx.tuning().sparse_from
    ctx.set_tuning(sparse_from=64)
    try:
        base, bst = ctx.render(sid, p)
        ctx.set_tuning(sparse_from=sparse_from)
        img, st = ctx.render(sid, p)
    finally:
        ctx.set_tuning(sparse_from=default)
    assert np.array_equal(img, base) and st.rays() == bst.rays()
    oimg, ost = oracle.OracleScene(hs.desc_ptr(), 32).render(p)
    e = float(np.sqrt(np.mean((img.astype(np.float64) - oimg) ** 2)))
    assert e <= 1e-3 and st.rays() == ost.rays(), e
    ctx.release(sid)


@pytest.mark.parametrize("persistent", [1, 2])
@pytest.mark.parametrize("config,variant,size", [(1, 4, (480, 270)), (2, 10, (320, 240)), (3, 2000, (256, 256))])
def test_persistent_runs_same_image(ctx, persistent, config, variant, size):
    """persistent grids walk each level in runs of 16, 32 or 64 rays per
    wave (the fewest that cover the level), the other lanes idle or helping
    in the wave-wide leaf tests: every ray keeps its own walk, so the image
    and the ray counts equal those of one run of 64 per wave bit for bit,
    and the heightfield matches the checker"""
    hs = S.HostScene.generate(config, variant=variant)
    sid = ctx.upload(hs)
    p = S.render_params(*size, samples=(1, 2), max_var=0.05, iter=6, shadow_samples=2, seed=7)
    default = ctx.tuning()
    ctx.set_tuning(persistent=0, chain=0)
    try:
        base, bst = ctx.render(sid, p)
        ctx.set_tuning(persistent=persistent)
        img, st = ctx.render(sid, p)
    finally:
        ctx.set_tuning(persistent=default.persistent, chain=default.chain)
    assert np.array_equal(img, base)
    assert st.rays() == bst.rays() and st.rays_shadow == bst.rays_shadow
    if config == 1:
        oimg, ost = oracle.OracleScene(hs.desc_ptr(), 32).render(p)
        e = float(np.sqrt(np.mean((img.astype(np.float64) - oimg) ** 2)))
        assert e <= 1e-3 and st.rays() == ost.rays(), e
    ctx.release(sid)


@pytest.mark.parametrize("config,variant,size", [(2, 10, (320, 240)), (3, 2000, (256, 256)), (0, 0, (320, 240))])
def test_sparse_instances_same_image(ctx, config, variant, size):
    """the sparse instances (two triangles in flight, box tests in min/max
    form where exact) render the image and ray counts of the dense
    instances on the room (axis-aligned walls), a batch scene (moving
    objects) and the spheres"""
    hs = S.HostScene.generate(config, variant=variant)
    sid = ctx.upload(hs)
    p = S.render_params(*size, samples=(1, 2), max_var=0.05, iter=5, shadow_samples=2, seed=5)
    default = ctx.tuning()
    ctx.set_tuning(sparse_from=64, chain=1)
    try:
        base, bst = ctx.render(sid, p)
        ctx.set_tuning(sparse_from=0, chain=0)
        img, st = ctx.render(sid, p)
    finally:
        ctx.set_tuning(sparse_from=default.sparse_from, chain=default.chain)
    assert np.array_equal(img, base) and st.rays() == bst.rays()
    ctx.release(sid)


@pytest.mark.parametrize("config,variant,size", [(1, 4, (480, 270)), (2, 10, (400, 300)), (0, 0, (400, 300))])
def test_lanes_give_the_same_image(ctx, config, variant, size):
    """a pass spread over 2 or 4 streams renders the image and ray counts of
    one stream.  Chunks of 2^16 pixels make these small frames several
    chunks (the room's levels also outgrow a sub-chunk; the spheres' cannot)"""
    hs = S.HostScene.generate(config, variant=variant)
    sid = ctx.upload(hs)
    p = S.render_params(*size, samples=(1, 2), max_var=0.05, iter=5, shadow_samples=2, seed=3)
    ctx.set_tuning(chunk_log2=16, lanes=1)
    try:
        base, bst = ctx.render(sid, p)
        for n in (2, 4):
            ctx.set_tuning(lanes=n)
            img, st = ctx.render(sid, p)
            err = float(np.abs(img - base).max())
            print(f"lanes={n}: max diff {err:.2e}")
            assert err <= TOL and st.rays() == bst.rays(), n
    finally:
        ctx.set_tuning(lanes=4, chunk_log2=21)
    ctx.release(sid)


@pytest.mark.parametrize("config,variant,size,samples", [(0, 0, (320, 240), (1, 1)), (0, 0, (160, 120), (1, 4)),
                                                          (3, 2000, (128, 96), (8, 8))])
def test_chain_gives_the_same_image(ctx, config, variant, size, samples):
    """one-child chains in one kernel (reflect-only spheres; a motion-blurred
    batch scene) render the image and ray counts of the level queues"""
    hs = S.HostScene.generate(config, variant=variant)
    sid = ctx.upload(hs)
    p = S.render_params(*size, samples=samples, max_var=0.01, iter=5, mblur=1 if config == 3 else 0, seed=11)
    ctx.set_tuning(chain=0)
    try:
        base, bst = ctx.render(sid, p)
        ctx.set_tuning(chain=1)
        img, st = ctx.render(sid, p)
    finally:
        ctx.set_tuning(chain=1)
    assert np.array_equal(img, base)
    assert st.rays() == bst.rays() and st.rays_shadow == bst.rays_shadow and st.rays_reflect == bst.rays_reflect
    ctx.release(sid)


def test_chain_mixed_scenes_serve(ctx):
    """requests on two sphere scenes in flight together (one-child chains of
    several scenes and requests) equal the same requests on the level queues"""
    a = S.HostScene.generate(0, seed=5)
    b = S.HostScene.generate(0, seed=6)
    sa, sb = ctx.upload(a), ctx.upload(b)
    jobs = [(sa, S.render_params(96, 64, samples=(1, 4), max_var=0.01, iter=5, seed=1), None, None),
            (sb, S.render_params(64, 80, rect=(8, 8, 56, 72), iter=4, seed=2), None, None),
            (sa, S.render_params(50, 40, iter=6, seed=3), None, None)]
    out = {}
    for ch in (0, 1):
        ctx.set_tuning(chain=ch)
        res = [ctx.poll(t) for t in ctx.submit_many(jobs)]
        out[ch] = [(img, r.rays) for img, r in res]
    ctx.set_tuning(chain=1)
    for (i0, r0), (i1, r1) in zip(out[0], out[1]):
        assert np.array_equal(i0, i1) and r0 == r1
    ctx.release(sa)
    ctx.release(sb)


def _runs_jobs(ctx):
    hb = S.HostScene.generate(3, variant=2000)
    hr = S.HostScene.generate(2, variant=10)
    sb, sr = ctx.upload(hb), ctx.upload(hr)
    ctx.shoot_photons(sr, 20000, 20000, 300.0, 5, 0, 0, 3)
    room = dict(shadow_samples=2, caust_photons=1, gi_photons=1, gather_dist=0.03)
    jobs = [
        # 231 pixels, adaptive: one sample per pass, and the list length
        # is not a multiple of 16 when the next request joins
        (sr, S.render_params(33, 7, samples=(1, 4), max_var=0.05, iter=4, seed=1, **room)),
        (sb, S.render_params(64, 48, samples=(8, 8), iter=4, mblur=1, seed=2)),   # 8 per pass
        (sr, S.render_params(40, 30, samples=(4, 4), iter=4, seed=3, **room)),    # 4 per pass
        (sb, S.render_params(50, 30, samples=(16, 16), iter=3, seed=4)),          # 16 per pass
        (sb, S.render_params(31, 17, samples=(8, 8), iter=3, seed=5)),            # 527 x 8 entries: one per pass
        (sr, S.render_params(24, 16, samples=(2, 2), iter=4, seed=6, **room)),    # 2 per pass
        # adaptive: the first pass takes the two samples every pixel takes
        # (the variance test starts at two), then one per pass
        (sr, S.render_params(32, 16, samples=(1, 8), max_var=0.05, iter=4, seed=7, **room)),
        (sb, S.render_params(48, 20, samples=(3, 16), max_var=0.02, iter=3, seed=8)),
    ]
    return (sb, sr), jobs


def test_sample_runs_same_images(ctx):
    """a request's first pass takes the samples every pixel takes (all of
    max_samples when min_samples >= max_samples; else max(min_samples, 2)),
    one queue entry per pixel and sample, summed in sample order with the
    variance test after each: each image and ray count equals one sample
    per pass bit for bit, rendered alone and submitted together with
    requests of one sample per pass (whose entries leave the list's length
    off a multiple of 16, so the run request's entries start at the last
    multiple and the entries past it move behind them)"""
    sids, jobs = _runs_jobs(ctx)
    default = ctx.tuning().sample_runs
    try:
        ctx.set_tuning(sample_runs=0)
        base = [ctx.render(sid, p) for sid, p in jobs]
        ctx.set_tuning(sample_runs=1)
        for k, (sid, p) in enumerate(jobs):
            img, st = ctx.render(sid, p)
            assert np.array_equal(img, base[k][0]), k
            assert st.rays() == base[k][1].rays() and st.rays_shadow == base[k][1].rays_shadow, k
        for rnd in range(2):
            tickets = ctx.submit_many([(sid, p, None, None) for sid, p in jobs])
            for k, t in enumerate(tickets):
                img, res = ctx.poll(t)
                print(f"round {rnd} request {k}: rays {res.rays} (one per pass {base[k][1].rays()})")
                assert np.array_equal(img, base[k][0]), (rnd, k)
                assert res.rays == base[k][1].rays(), (rnd, k)
    finally:
        ctx.set_tuning(sample_runs=default)
    for sid in sids:
        ctx.release(sid)


def test_bad_tuning_rejected(ctx):
    for kw in (dict(trace_block=96), dict(tile_w=12), dict(gather_lanes=4), dict(chunk_log2=22),
               dict(gather_cell=3.0), dict(graphs=2), dict(persistent=3), dict(lanes=3), dict(lanes=0),
               dict(sparse_from=-1), dict(sparse_from=65), dict(chain=2), dict(sample_runs=2)):
        with pytest.raises(S.SrayError):
            ctx.set_tuning(**kw)


def test_autotune(ctx):
    hs = S.HostScene.generate(0)
    sid = ctx.upload(hs)
    p = S.render_params(320, 240, iter=5)
    before, _ = ctx.render(sid, p)
    t, ms = ctx.autotune(sid, p, rounds=2)
    assert ms > 0 and t.trace_block in (64, 128, 256) and t.tile_w in (8, 16, 32, 64)
    kept = ctx.tuning()
    assert (kept.trace_block, kept.tile_w, kept.chunk_log2) == (t.trace_block, t.tile_w, t.chunk_log2)
    after, _ = ctx.render(sid, p)
    assert float(np.abs(after - before).max()) <= TOL
    ctx.set_tuning(trace_block=64, tile_w=64, gather_lanes=16, chunk_log2=21, gather_cell=1.0, graphs=1, persistent=0,
                   lanes=4, sparse_from=64, chain=1, sample_runs=1)
    ctx.release(sid)


def test_autotune_octree(ctx):
    hs = S.HostScene.generate(1, variant=8)
    p = S.render_params(240, 135, iter=3)
    # the last candidate (one item per leaf, depth up to 30) builds a tree
    # deeper than the traversal stack holds: its upload fails and it is skipped
    cands = [(14, 7), (32, 8), (8, 6), (1, 30)]
    k, ms = ctx.autotune_octree(hs.desc_ptr(), p, cands, rounds=1)
    assert 0 <= k < 3 and ms > 0
