"""The box tests of the traversal in min/max form (slab and expand_children
in sray_amd/csrc/traverse.h) accept exactly the children the reference's
comparison sequence accepts (AABox::intersect, src/aabb.cc:81-125), under
the preconditions the kernels check: a finite origin, every 1/d component
finite and non-zero, every box finite with lo <= hi, and a bound of at least
2^-149 (the bound is FLT_MAX or a hit distance, and hits are at t >= 1e-6;
at bound 0 the folded compare would differ, as the last test shows).

Both forms are restated here in numpy float32 with the kernels' operations:
the sequence with its NaN-faithful comparisons, the min/max form with the
folded compare max(entry, 2^-149) <= min(exit, min(bound, 1 - 2^-24)).
Pins: equality of the accepted child sets (and of the node's own entry
distance) on random rays and boxes, and on adversarial ones: origins on box
faces and centres, entries and exits exactly 0 and 1, bounds equal to an
entry, denormal distances, huge and tiny direction components.
"""
import numpy as np

F = np.float32


def _seq(txmin, txmax, tymin, tymax, tzmin, tzmax, bound):
    """slab()'s sequence (float comparisons, updates by '>' / '<')"""
    tmin, tmax = txmin, txmax
    if tmin > tymax or tymin > tmax:
        return False
    if tymin > tmin:
        tmin = tymin
    if tymax < tmax:
        tmax = tymax
    if tmin > tzmax or tzmin > tmax:
        return False
    if tzmin > tmin:
        tmin = tzmin
    if tzmax < tmax:
        tmax = tzmax
    return bool(tmin < F(1.0) and tmax > F(0.0) and not (tmin > bound))


def _axis(lo, hi, o, inv, neg):
    """near, centre, far distances of one axis (the kernels' operations)"""
    c = F(F(lo + hi) * F(0.5))
    near = F(F((hi if neg else lo) - o) * inv)
    far = F(F((lo if neg else hi) - o) * inv)
    mid = F(F(c - o) * inv)
    return near, mid, far


def _children(lo, hi, o, d, bound):
    inv = [F(F(1.0) / d[a]) for a in range(3)]
    ax = [_axis(lo[a], hi[a], o[a], inv[a], d[a] < 0) for a in range(3)]
    seq = 0
    for k in range(8):
        iv = []
        for a in range(3):
            n, c, f = ax[a]
            iv += [c, f] if (k >> a) & 1 else [n, c]
        if _seq(*iv, bound):
            seq |= 1 << k
    lo_t = F(2.0 ** -149)
    cb = min(F(bound), F(1.0) - F(2.0 ** -24))
    (nx, xc, fx), (ny, yc, fy), (nz, zc, fz) = ax
    nx2, xc2, fx2, xc3 = max(nx, lo_t), max(xc, lo_t), min(fx, cb), min(xc, cb)
    mm = 0
    for k in range(8):
        tx = (xc2, fx2) if k & 1 else (nx2, xc3)
        ty = (yc, fy) if k & 2 else (ny, yc)
        tz = (zc, fz) if k & 4 else (nz, zc)
        a = max(max(tx[0], ty[0]), tz[0])
        b = min(min(tx[1], ty[1]), tz[1])
        if a <= b:
            mm |= 1 << k
    # the node's own test in both forms, and the claim that a failing node
    # leaves no child
    node_seq = _seq(nx, fx, ny, fy, nz, fz, bound)
    node_mm = max(max(nx, ny), nz) <= min(min(fx, fy), fz) and max(max(nx, ny), nz) < 1 and \
        min(min(fx, fy), fz) > 0 and max(max(nx, ny), nz) <= bound
    return seq, mm, node_seq, node_mm


def _check(lo, hi, o, d, bound):
    with np.errstate(all="ignore"):
        inv = [F(1.0) / F(v) for v in d]
        if not all(np.isfinite(v) and v != 0 for v in inv):
            return None
        seq, mm, ns, nm = _children(lo, hi, o, d, F(bound))
    assert seq == mm, (lo, hi, o, d, bound, seq, mm)
    assert ns == nm, (lo, hi, o, d, bound)
    if not ns:
        assert mm == 0, (lo, hi, o, d, bound)
    return mm


def test_random_boxes_and_rays():
    rng = np.random.default_rng(3)
    hits = 0
    for _ in range(6000):
        lo = rng.uniform(-10, 10, 3).astype(F)
        hi = (lo + rng.uniform(0, 8, 3)).astype(F)
        o = rng.uniform(-20, 20, 3).astype(F)
        # towards a point near the box, reaching it at t in (0, 2)
        target = lo + (hi - lo) * rng.uniform(-0.3, 1.3, 3)
        d = ((target - o) / rng.uniform(0.05, 2.0) + rng.normal(size=3) * 0.01).astype(F)
        bound = F(rng.choice([3.4e38, rng.uniform(1e-6, 1.2)]))
        m = _check(lo, hi, o, d, bound)
        hits += bool(m)
    assert hits > 1000


def test_adversarial_cases():
    rng = np.random.default_rng(8)
    n = 0
    for _ in range(6000):
        lo = rng.integers(-4, 4, 3).astype(F)
        hi = (lo + rng.integers(0, 3, 3)).astype(F)      # flat boxes too
        c = (lo + hi) * F(0.5)
        pick = rng.integers(0, 4, 3)
        o = np.where(pick == 0, lo, np.where(pick == 1, hi, np.where(pick == 2, c, lo - 1))).astype(F)
        d = rng.choice([-4.0, -1.0, -0.5, 0.25, 1.0, 2.0, 3.0, 1e-30, -1e-30, 1e30], 3).astype(F)
        # bounds: 1, the smallest positive float, hit-like distances and
        # entries of the box (ties at 0.25, 0.5)
        bound = F(rng.choice([3.4e38, 1.0, 2.0 ** -149, 1e-6, 0.25, 0.5]))
        if _check(lo, hi, o, d, bound) is not None:
            n += 1
    assert n > 3000


def test_fold_needs_a_positive_bound():
    """at bound 0 the sequence accepts an entry <= 0 < exit and the fold
    does not: why the kernels' bound (never below 1e-6) matters"""
    lo, hi = np.array([0, -2, -4], F), np.array([2, 0, -3], F)
    o, d = np.array([2, -1, -4], F), np.array([-0.5, 0.25, 1.0], F)
    seq, mm, _, _ = _children(lo, hi, o, d, F(0.0))
    assert seq != mm
    seq, mm, _, _ = _children(lo, hi, o, d, F(1e-6))
    assert seq == mm
