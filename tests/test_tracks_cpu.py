"""Keyframe tracks on the host (XFormNode, src/anim.inl:188-275,
src/anim.cc:68-121, 391-427): the per-msec rows the loader evaluates,
pinned to closed forms: rotation angle linear in time under slerp about a
fixed axis, rotation about a pivot leaving the pivot fixed, linear scale,
step and clamped keys, repeat extrapolation, Catmull-Rom through the keys
and exact on linear data, the target camera's look-at, light positions,
and every row's inverse."""
import math

import numpy as np
import pytest

import oracle
import sray_amd as S
from scenes import animated_scene_xml, rays, rows_of


def _load(msec=20, shutter=40):
    hs = S.HostScene.load_xml(text=animated_scene_xml(shutter))
    hs.set_frame(msec)
    return hs


def _m(row):
    m = np.eye(4)
    m[:3] = np.asarray(row[:12], np.float64).reshape(3, 4)
    return m


def test_frame_interval_and_rows():
    hs = _load(20)
    d = hs.desc
    assert (d.frame_t0, d.frame_t1) == (0, 40)
    assert rows_of(d, "object", 0) is None           # the static floor
    for i in (1, 2, 3, 4, 5):
        r = rows_of(d, "object", i)
        assert r is not None and r.shape == (41, 24)
        for row in r[::5]:
            assert np.allclose(_m(row) @ _m(row[12:]), np.eye(4), atol=2e-5)
    hs.set_frame(100)
    assert (hs.desc.frame_t0, hs.desc.frame_t1) == (80, 120)


def test_rotation_is_linear_in_angle_under_slerp():
    hs = _load(20)
    r = rows_of(hs.desc, "object", 1)
    for t in (0, 7, 20, 33, 40):
        a = math.radians(90.0 * t / 40)
        s = np.array([1.0 + 0.5 * t / 40, 1.0 - 0.4 * t / 40, 1.0])
        R = np.array([[math.cos(a), 0, math.sin(a)], [0, 1, 0], [-math.sin(a), 0, math.cos(a)]])
        want = R @ np.diag(s)
        got = _m(r[t])[:3, :3]
        assert np.allclose(got, want, atol=2e-6), (t, got, want)
        assert np.allclose(_m(r[t])[:3, 3], [-2.5, -1.0, 1.0], atol=1e-6)


def test_rotation_about_a_pivot_keeps_the_pivot():
    hs = _load(20)
    r = rows_of(hs.desc, "object", 3)
    for t in (0, 13, 40):
        m = _m(r[t])
        # pivot * T * R * S * pivot^-1: the pivot point maps to pivot + translation
        p = m @ np.array([0.5, 0.0, 0.0, 1.0])
        assert np.allclose(p[:3], [0.5, -1.0, 3.0], atol=2e-6), (t, p)
    # 45 degrees about z at t = 40 (quaternion cos 22.5, sin 22.5 about z)
    a = math.radians(45.0)
    assert np.allclose(_m(r[40])[:3, :3], [[math.cos(a), -math.sin(a), 0], [math.sin(a), math.cos(a), 0], [0, 0, 1]],
                       atol=2e-6)


def test_repeat_extrapolation_and_linear_keys():
    hs = _load(20)
    r = rows_of(hs.desc, "object", 2)
    for t in (0, 12, 30):
        u = t / 30.0
        assert np.allclose(_m(r[t])[:3, 3], [1.5, 0.2 + u, 0.5 * u], atol=2e-6)
        assert np.allclose(np.diag(_m(r[t])[:3, :3]), [1.0 + 0.6 * u, 1.0 - 0.3 * u, 1.0], atol=2e-6)
    # past the last key the track repeats: t = 35 -> 5, t = 40 -> 10
    for t, tr in ((35, 5), (40, 10)):
        assert np.allclose(r[t], r[tr], atol=0), t
    # before the first key: frame at msec -10 (rows -30..10), t = -10 -> 20
    hs.set_frame(0)
    r0 = rows_of(hs.desc, "object", 2)
    assert (hs.desc.frame_t0, hs.desc.frame_t1) == (-20, 20)
    assert np.allclose(r0[0], r[10], atol=0)   # msec -20 -> 10
    assert np.allclose(r0[10], r[20], atol=0)  # msec -10 -> 20


def test_step_and_clamp():
    hs = _load(20)
    r = rows_of(hs.desc, "object", 4)
    for t in (0, 10, 19):
        assert np.allclose(_m(r[t])[:3, 3], [-0.5, 0.0, -1.5])
    for t in (20, 30, 40):
        assert np.allclose(_m(r[t])[:3, 3], [0.0, 0.3, -1.5])


def test_cubic_passes_through_keys():
    hs = _load(20)
    r = rows_of(hs.desc, "object", 5)
    for t, p in ((0, (3.0, 2.0, 2.0)), (10, (3.5, 2.5, 2.0)), (25, (3.0, 3.0, 2.5)), (40, (2.5, 2.0, 2.0))):
        assert np.allclose(_m(r[t])[:3, 3], p, atol=1e-6), t
    # between keys it departs from the straight line (the spline is not linear here)
    mid = _m(r[5])[:3, 3]
    assert not np.allclose(mid, [3.25, 2.25, 2.0], atol=1e-4)


def test_cubic_reproduces_linear_motion():
    x = animated_scene_xml().replace(
        '<xform interpolate="cubic" time="0" pos="3.0 2.0 2.0"/>\n    <xform time="10" pos="3.5 2.5 2.0"/>\n'
        '    <xform time="25" pos="3.0 3.0 2.5"/>\n    <xform time="40" pos="2.5 2.0 2.0"/>',
        '<xform interpolate="cubic" time="0" pos="0.0 0.0 0.0"/>\n    <xform time="10" pos="1.0 2.0 3.0"/>\n'
        '    <xform time="20" pos="2.0 4.0 6.0"/>\n    <xform time="30" pos="3.0 6.0 9.0"/>')
    hs = S.HostScene.load_xml(text=x)
    hs.set_frame(20)
    r = rows_of(hs.desc, "object", 5)
    # exact on the inner segment; the end segments repeat their end key as
    # the outer control point (src/anim.inl:244-252), so they bend
    for t in (10, 11, 15, 17, 20):
        assert np.allclose(_m(r[t])[:3, 3], np.array([1.0, 2.0, 3.0]) * t / 10, atol=2e-6), t
    assert not np.allclose(_m(r[3])[:3, 3], np.array([1.0, 2.0, 3.0]) * 0.3, atol=1e-3)


def test_light_and_target_camera_rows():
    hs = _load(20)
    d = hs.desc
    lr = rows_of(d, "light", 0)
    for t in (0, 10, 40):
        u = t / 40
        assert np.allclose(lr[t], [4.0 - 8.0 * u, 8.0 + u, -6.0 + u], atol=1e-6)
    cr = rows_of(d, "camera")
    for t in (0, 20, 40):
        u = t / 40
        pos = np.array([0.0, 2.5, -9.0]) + u * np.array([1.0, 0.5, 0.5])
        tg = u * np.array([0.3, 0.2, 0.5])
        f = (tg - pos) / np.linalg.norm(tg - pos)
        rt = np.cross([0.0, 1.0, 0.0], f)
        rt /= np.linalg.norm(rt)
        up = np.cross(f, rt)
        m = _m(cr[t])
        assert np.allclose(m[:3, 0], rt, atol=2e-6) and np.allclose(m[:3, 1], up, atol=2e-6)
        assert np.allclose(m[:3, 2], f, atol=2e-6) and np.allclose(m[:3, 3], pos, atol=2e-6)


def test_xml_round_trip_keeps_the_tracks():
    hs = _load(20)
    back = S.HostScene.load_xml(text=hs.to_xml())
    back.set_frame(20)
    for i in range(hs.desc.num_objects):
        a, b = rows_of(hs.desc, "object", i), rows_of(back.desc, "object", i)
        assert (a is None) == (b is None)
        if a is not None:
            assert np.array_equal(a, b), i
    assert np.array_equal(rows_of(hs.desc, "camera"), rows_of(back.desc, "camera"))
    assert np.array_equal(rows_of(hs.desc, "light", 0), rows_of(back.desc, "light", 0))


@pytest.mark.parametrize("prec", [32, 64])
def test_checker_octree_equals_brute_force_at_ray_times(prec):
    """the checker reads the rows at each ray's time: its octree walk (bounds
    at the interval's ends) and brute force agree on segments that stay
    inside the bounds of both times"""
    hs = _load(20)
    rng = np.random.default_rng(3)
    n = 4000
    o = rng.uniform(-4, 4, (n, 3)) + np.array([0.0, 1.0, 0.0])
    dd = rng.normal(size=(n, 3)) * 6.0
    r = rays(o, dd)
    r["time"] = rng.integers(0, 41, n)
    a, ea = oracle.OracleScene(hs.desc_ptr(), prec).trace(r)
    b, eb = oracle.OracleScene(hs.desc_ptr(), prec, oracle.BRUTE_FORCE).trace(r)
    hit = (a["obj"] >= 0).sum()
    assert hit > n // 4
    # the tree may miss an object moving out of its t0/t1 bounds in between
    # (the reference's build_tree(t0, t1)); where both hit they agree
    both = (a["obj"] >= 0) & (b["obj"] >= 0) & (ea["tie"] == 0) & (eb["tie"] == 0)
    assert both.sum() > hit * 0.9
    assert np.array_equal(a["obj"][both], b["obj"][both])
    assert np.array_equal(a["t"][both].view(np.uint32), b["t"][both].view(np.uint32))
