"""Writes kat_closed_form.json: closed-form known answers for the ray
primitives and the shading recursion, computed here in double precision
from the mathematics (not from either implementation).

    python tests/golden/make_golden.py
"""
import json
import math
import os


def sphere_t(o, d, c, r):
    oc = [o[i] - c[i] for i in range(3)]
    a = sum(x * x for x in d)
    b = 2 * sum(d[i] * oc[i] for i in range(3))
    cc = sum(x * x for x in oc) - r * r
    disc = b * b - 4 * a * cc
    if disc < 0:
        return None
    t1 = (-b + math.sqrt(disc)) / (2 * a)
    t2 = (-b - math.sqrt(disc)) / (2 * a)
    eps = 1e-6
    if (t1 < eps and t2 < eps) or (t1 > 1 and t2 > 1):
        return None
    if t1 < eps:
        t1 = t2
    if t2 < eps:
        t2 = t1
    return min(t1, t2)


def main():
    out = {}
    c, r = (0.0, 0.0, 10.0), 2.0
    cases = []
    for name, o, d in [("axis", (0, 0, 0), (0, 0, 100)), ("offset_miss", (3, 0, 0), (0, 0, 100)),
                       ("tangent", (2, 0, 0), (0, 0, 100)), ("behind", (0, 0, 20), (0, 0, 100)),
                       ("too_short", (0, 0, 0), (0, 0, 5)), ("inside_q4_t_gt_1", (0, 0, 10), (0, 0, 1)),
                       ("inside_long", (0, 0, 10), (0, 0, 100)), ("ends_inside", (0, 0, 0), (0, 0, 9))]:
        t = sphere_t(o, d, c, r)
        cases.append({"name": name, "o": o, "d": d, "hit": t is not None, "t": t})
    out["sphere"] = {"center": c, "radius": r, "cases": cases}
    # cylinder along local y at (0,0,10), r = 1, y in [0, 2): t = (10 - 1)/100
    out["cylinder"] = {"center": (0.0, 0.0, 10.0), "radius": 1.0, "start": 0.0, "end": 2.0, "cases": [
        {"name": "mid", "o": (0, 1, 0), "d": (0, 0, 100), "hit": True, "t": 0.09},
        {"name": "start_inclusive", "o": (0, 0, 0), "d": (0, 0, 100), "hit": True, "t": 0.09},
        {"name": "end_exclusive", "o": (0, 2, 0), "d": (0, 0, 100), "hit": False, "t": None},
        {"name": "above", "o": (0, 2.5, 0), "d": (0, 0, 100), "hit": False, "t": None},
        {"name": "inside_out", "o": (0, 1, 10), "d": (0, 0, 100), "hit": True, "t": 0.01},
        {"name": "offset_miss", "o": (1.5, 1, 0), "d": (0, 0, 100), "hit": False, "t": None}]}
    # triangle in the plane z = 5: t = 5 / d.z for rays from z = 0
    tri = [(-1, -1, 5), (1, -1, 5), (0, 1, 5)]
    out["triangle"] = {"tri": tri, "cases": [
        {"name": "center", "o": (0, 0, 0), "d": (0, 0, 10), "hit": True, "t": 0.5},
        {"name": "edge_inclusive", "o": (0, -1, 0), "d": (0, 0, 10), "hit": True, "t": 0.5},
        {"name": "vertex_inclusive", "o": (0, 1, 0), "d": (0, 0, 10), "hit": True, "t": 0.5},
        {"name": "outside", "o": (0, -1.25, 0), "d": (0, 0, 10), "hit": False, "t": None},
        {"name": "parallel", "o": (-5, 0, 5), "d": (10, 0, 0), "hit": False, "t": None},
        {"name": "t_gt_1", "o": (0, 0, 0), "d": (0, 0, 4), "hit": False, "t": None},
        {"name": "origin_on_plane", "o": (0, 0, 5), "d": (0, 0, 10), "hit": False, "t": None},
        {"name": "back_face", "o": (0, 0, 10), "d": (0, 0, -10), "hit": True, "t": 0.5}]}
    # Phong at normal incidence, light at the eye, opaque reflector:
    # (amb*kd + 1)*kd + (1 + refl*bg)*ks   (fres = 1 for refract = 0)
    kd, ks, refl = (0.5, 0.25, 1.0), 0.5, 0.5
    amb, bg = (0.1, 0.2, 0.3), (0.2, 0.4, 0.6)
    exp = [(amb[i] * kd[i] + 1.0) * kd[i] + (1.0 + refl * bg[i]) * ks for i in range(3)]
    out["phong"] = {"kd": kd, "ks": ks, "refl": refl, "ambient": amb, "bg": bg, "expected": exp}
    # glass sphere along its axis: Schlick r0 = ((n1-n2)/(n1+n2))^2 = 0.04
    r0 = ((1.0 - 1.5) / (1.0 + 1.5)) ** 2
    gbg = (0.2, 0.4, 0.6)
    out["glass"] = {"bg": gbg, "r0": r0, "expected": [x * (1 - r0) ** 2 for x in gbg]}
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "kat_closed_form.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", path)


if __name__ == "__main__":
    main()
