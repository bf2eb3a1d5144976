"""Exact ties on the device: hit IDs must equal the checker's on tie rays.

The reference keeps the first depth-first occurrence of the smallest t
(src/octree.inl:200-241).  The scenes of tests/scenes.py (tie_*) put the
tied items so that for rays heading to -x that occurrence lies in a leaf
beyond the hit, which a near-to-far walk that skips such leaves never
visits.  Every kernel instance that walks the octrees is checked: the
one-lane walk (API traces, one-child chains, photon paths) and the level
kernels' instances (dense and sparse), with all lanes of a wave busy and
with one lane in sixteen (the waves then test leaves together).
"""
import numpy as np
import pytest

import oracle
import sray_amd as S
from scenes import rays, sparse_lanes, tie_mesh_scene, tie_object_scene, tie_rays

pytestmark = pytest.mark.gpu

KERNELS = ["plain", "wave", "sparse"]


def _check(ctx, sid, ref, r, label, min_ties=0):
    ohits, ext = ref.trace(r)
    ties = int((ext["tie"] != 0).sum())
    for kernel in KERNELS:
        h = ctx.trace(sid, r, kernel=kernel)
        bad = np.flatnonzero((h["obj"] != ohits["obj"]) | (h["prim"] != ohits["prim"]) |
                             ((ohits["obj"] >= 0) & (h["t"].view(np.uint32) != ohits["t"].view(np.uint32))))
        print(f"{label} [{kernel}]: {len(r)} rays, {ties} ties, {len(bad)} mismatches", flush=True)
        assert len(bad) == 0, f"{label} [{kernel}] ray {bad[0]}: gpu {h[bad[0]]} cpu {ohits[bad[0]]}"
    assert ties >= min_ties, f"{label}: only {ties} ties"
    return ties


@pytest.mark.parametrize("spread", [1, 16])
def test_mesh_level_ties(ctx, spread):
    b, pairs = tie_mesh_scene()
    sid = ctx.upload(b.ptr())
    ref = oracle.OracleScene(b.ptr(), 32)
    r, want, neg = tie_rays(pairs, per_pair=64, seed=11, edge=True)
    if spread > 1:
        r, _ = sparse_lanes(r, spread)
    _check(ctx, sid, ref, r, f"mesh-level ties (1 lane in {spread})", min_ties=100)
    ctx.release(sid)


@pytest.mark.parametrize("spread", [1, 16])
def test_object_level_ties(ctx, spread):
    b, pairs = tie_object_scene()
    sid = ctx.upload(b.ptr())
    ref = oracle.OracleScene(b.ptr(), 32)
    r, want, neg = tie_rays(pairs, per_pair=64, seed=12, edge=False)
    if spread > 1:
        r, _ = sparse_lanes(r, spread)
    _check(ctx, sid, ref, r, f"object-level ties (1 lane in {spread})", min_ties=100)
    # coincident spheres: every hit is a tie between objects 0 and 1
    n = 512
    rng = np.random.default_rng(13)
    o = np.tile([-4.0, 4.0, -6.0], (n, 1)) + np.c_[rng.uniform(-0.6, 0.6, (n, 2)), np.zeros(n)]
    _check(ctx, sid, ref, rays(o, np.tile([0.0, 0.0, 12.0], (n, 1))), "coincident spheres", min_ties=100)
    ctx.release(sid)


def test_random_rays_in_tie_scenes(ctx):
    """the tie scenes under random segments (mostly no ties): nothing else changes"""
    rng = np.random.default_rng(14)
    for build in (tie_mesh_scene, tie_object_scene):
        b, _ = build()
        sid = ctx.upload(b.ptr())
        ref = oracle.OracleScene(b.ptr(), 32)
        n = 20000
        o = rng.uniform(-6, 6, (n, 3))
        d = rng.normal(size=(n, 3)) * rng.choice([2.0, 12.0], (n, 1))
        _check(ctx, sid, ref, rays(o, d), f"{build.__name__} random")
        ctx.release(sid)


def test_room_split_quads(ctx):
    """the photon room's walls are quads split into two triangles: rays
    aimed at the split diagonals, plus random rays; every ray is checked"""
    hs = S.HostScene.generate(2, variant=10)
    sid = ctx.upload(hs)
    ref = oracle.OracleScene(hs.desc_ptr(), 32)
    rng = np.random.default_rng(15)
    n = 40000
    o = rng.uniform(-0.9, 0.9, (n, 3)) * np.array([4.0, 2.0, 4.0]) + np.array([0.0, 2.0, 0.0])
    d = rng.normal(size=(n, 3)) * 30.0
    ties = _check(ctx, sid, ref, rays(o, d), "room random")
    print(f"room: {ties} ties among random rays")
    # exact hits on the floor's split diagonals (a -> c of each quad, the
    # edge the two triangles share) and on the seam z = 0 between the two
    # floor quads (two objects): t = 0.5, every value dyadic
    m = 4096
    s_ = np.floor(rng.uniform(0, 1, m) * 64) / 64
    k = rng.integers(0, 2, m)
    a = np.where(k[:, None] == 0, [-5.0, 0.0, -5.0], [-5.0, 0.0, 0.0])
    c = np.where(k[:, None] == 0, [5.0, 0.0, 0.0], [5.0, 0.0, 5.0])
    diag = a + s_[:, None] * (c - a)
    seam = np.c_[np.floor(rng.uniform(-4, 4, m) * 16) / 16, np.zeros(m), np.zeros(m)]
    p = np.concatenate([diag, seam])
    dd = np.c_[np.floor(rng.uniform(-4, 4, 2 * m) * 4) / 4, np.full(2 * m, -8.0),
               np.floor(rng.uniform(-4, 4, 2 * m) * 4) / 4]
    ties = _check(ctx, sid, ref, rays(p - 0.5 * dd, dd), "room floor diagonals and seam", min_ties=100)
    print(f"room: {ties} ties on the floor's diagonals and seam")
    ctx.release(sid)
