"""Item boxes in the wave-wide walks: a mesh leaf's triangles are tested
only when the ray's segment meets their (8-bit coded, padded) bounding boxes.
The filter must never drop a hit the triangle test reports.

Bar: hit object / triangle IDs and t bit-exact against the strict float
checker, and any-hit occlusion equal, for the level-kernel instances
("wave", "sparse") that use the boxes, on rays chosen to sit where a coded
box is tight: aimed at triangle interiors, vertices and edge midpoints,
grazing triangle planes, with near-zero direction components (huge inverse
directions), from origins far from the mesh, and short segments ending at a
hit.  The mesh sits far from the origin (coordinates near 3000) and mixes
large, small, tiny and sliver triangles, so items reach beyond their cells
(open codes) and the padding must scale with the coordinates.
"""
import numpy as np
import pytest

import oracle
import sray_amd as S
from scenes import build, phong

pytestmark = pytest.mark.gpu

OFFSET = np.array([3000.0, -2000.0, 1500.0])


def _triangles(rng):
    tris = []
    # ordinary triangles, 0.2 - 3 units, over a 40-unit cube
    c = rng.uniform(-20, 20, (1500, 1, 3))
    tris.append(c + rng.normal(0, 1, (1500, 3, 3)) * rng.uniform(0.1, 1.5, (1500, 1, 1)))
    # tiny ones
    c = rng.uniform(-20, 20, (300, 1, 3))
    tris.append(c + rng.normal(0, 1e-3, (300, 3, 3)))
    # slivers: a long edge and a vertex almost on it
    a = rng.uniform(-20, 20, (100, 3))
    b = a + rng.normal(0, 4, (100, 3))
    m = 0.5 * (a + b) + rng.normal(0, 1e-5, (100, 3))
    tris.append(np.stack([a, b, m], 1))
    # large ones, spanning many cells
    c = rng.uniform(-15, 15, (40, 1, 3))
    tris.append(c + rng.normal(0, 12, (40, 3, 3)))
    # axis-aligned quads on round coordinates (edges on octree split planes)
    for k in range(-16, 17, 4):
        q = np.array([[k, -8, -8], [k, 8, -8], [k, 8, 8]], float)
        tris.append(q[None])
        tris.append(np.array([[k, -8, -8], [k, 8, 8], [k, -8, 8]], float)[None])
    return np.concatenate(tris).astype(np.float32)


def _rays(rng, tris):
    """world-space rays at the mesh (translated by OFFSET)"""
    n = len(tris)
    world = tris.astype(np.float64) + OFFSET
    os_, ds_ = [], []

    def add(o, target, length=None):
        d = target - o
        if length is not None:
            d = d / np.linalg.norm(d, axis=-1, keepdims=True) * length[:, None]
        os_.append(o)
        ds_.append(d)

    k = rng.integers(0, n, 4000)
    # interior points, from 5-60 units away, segments of length 1e4 (camera-like)
    w = rng.dirichlet((1, 1, 1), 4000)
    p = np.einsum("ki,kij->kj", w, world[k])
    o = p + rng.normal(0, 1, (4000, 3)) * rng.uniform(5, 60, (4000, 1))
    add(o, p, np.full(4000, 1e4))
    # vertices and edge midpoints: hit points on the boxes' faces
    k = rng.integers(0, n, 3000)
    v = world[k, rng.integers(0, 3, 3000)]
    e = 0.5 * (world[k, 0] + world[k, 1])
    for pts in (v, e):
        o = pts + rng.normal(0, 1, (3000, 3)) * rng.uniform(2, 40, (3000, 1))
        add(o, pts, np.full(3000, 1e4))
    # segments ending just past the hit (shadow-like, |d| small)
    k = rng.integers(0, n, 3000)
    w = rng.dirichlet((1, 1, 1), 3000)
    p = np.einsum("ki,kij->kj", w, world[k])
    o = p + rng.normal(0, 1, (3000, 3)) * 3.0
    add(o, p + (p - o) * 1e-3)
    # grazing: in the triangle's plane, nudged off it
    k = rng.integers(0, n, 3000)
    a, b, c = world[k, 0], world[k, 1], world[k, 2]
    nrm = np.cross(b - a, c - a)
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True) + 1e-30
    w = rng.dirichlet((1, 1, 1), 3000)
    p = np.einsum("ki,kij->kj", w, world[k])
    tang = (b - a) / (np.linalg.norm(b - a, axis=1, keepdims=True) + 1e-30)
    o = p - tang * rng.uniform(3, 30, (3000, 1)) + nrm * rng.normal(0, 1e-3, (3000, 1))
    add(o, p, np.full(3000, 1e4))
    # near-zero direction components: inverse directions near 1e20-1e30
    k = rng.integers(0, n, 2000)
    p = world[k].mean(1)
    axis = rng.integers(0, 3, 2000)
    d = np.zeros((2000, 3))
    d[np.arange(2000), axis] = rng.choice([-1.0, 1.0], 2000)
    d += rng.choice([-1.0, 1.0], (2000, 3)) * 10.0 ** rng.uniform(-30, -20, (2000, 3))
    o = p - d * rng.uniform(5, 30, (2000, 1))
    os_.append(o)
    ds_.append(d * 1e4)
    # far origins
    k = rng.integers(0, n, 2000)
    p = world[k].mean(1)
    o = p + rng.normal(0, 1, (2000, 3)) * 1e5
    add(o, p, np.full(2000, 3e5))
    o = np.concatenate(os_).astype(np.float32)
    d = np.concatenate(ds_).astype(np.float32)
    return S.make_rays(o, d)


@pytest.fixture(scope="module")
def far_mesh():
    rng = np.random.default_rng(21)
    tris = _triangles(rng)
    b = build(meshes=[(tris, tuple(OFFSET), 0)], spheres=[(tuple(OFFSET + [0, 30, 0]), 2.0, 0)],
              materials=[phong()], lights=[(tuple(OFFSET + [0, 100, 0]), (1, 1, 1))], mesh_oct=(7, 8))
    return b, _rays(rng, tris)


@pytest.mark.parametrize("kernel", ["wave", "sparse"])
def test_item_boxes_closest(ctx, far_mesh, kernel):
    b, rays = far_mesh
    sid = ctx.upload(b.ptr())
    ref = oracle.OracleScene(b.ptr(), 32)
    hits = ctx.trace(sid, rays, kernel=kernel)
    ohits, ext = ref.trace(rays)
    ctx.release(sid)
    bad = np.flatnonzero((hits["obj"] != ohits["obj"]) | (hits["prim"] != ohits["prim"]) |
                         ((ohits["obj"] >= 0) & (hits["t"].view(np.uint32) != ohits["t"].view(np.uint32))))
    print(f"{kernel}: {len(rays)} rays, {(ohits['obj'] >= 0).sum()} hits, {(ext['tie'] != 0).sum()} ties, "
          f"mismatches {len(bad)}")
    assert len(bad) == 0, f"first ray {bad[0]}: gpu {hits[bad[0]]} cpu {ohits[bad[0]]}"


@pytest.mark.parametrize("kernel", ["wave", "sparse"])
def test_item_boxes_any_hit(ctx, far_mesh, kernel):
    b, rays = far_mesh
    sid = ctx.upload(b.ptr())
    ref = oracle.OracleScene(b.ptr(), 32)
    hits = ctx.trace(sid, rays, any_hit=True, kernel=kernel)
    ohits, _ = ref.trace(rays, any_hit=True)
    ctx.release(sid)
    occ, oocc = hits["obj"] >= 0, ohits["obj"] >= 0
    print(f"{kernel} any hit: {len(rays)} rays, {oocc.sum()} occluded, mismatches {(occ != oocc).sum()}")
    assert np.array_equal(occ, oocc)
