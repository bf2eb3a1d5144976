"""Host-side C++ of the product (no device): the reference-format XML loader,
the generators, and the octree build/flatten that feeds the kernels —
checked against the CPU checker's octree and against the reference's
subdivision rule recomputed from item boxes."""
import ctypes as C

import numpy as np
import pytest

import oracle
import sray_amd as S
from sray_amd import abi

LEAF = 0x80000000


def _objects(d):
    return [d.objects[i] for i in range(d.num_objects)]


def _same_scene(a, b):
    da, db = a.desc, b.desc
    assert da.num_objects == db.num_objects and da.num_meshes == db.num_meshes
    assert da.num_triangles == db.num_triangles and da.num_lights == db.num_lights
    assert da.num_materials == db.num_materials
    for x, y in zip(_objects(da), _objects(db)):
        assert bytes(x) == bytes(y)
    n = da.num_triangles * 9
    if n:
        assert np.array_equal(np.ctypeslib.as_array(da.tri_pos, (n,)), np.ctypeslib.as_array(db.tri_pos, (n,)))
    for name, per in (("tri_normal", 9), ("tri_texcoord", 6), ("tri_tangent", 9)):
        pa, pb = getattr(da, name), getattr(db, name)
        assert bool(pa) == bool(pb), name
        if pa:
            m = per * da.num_triangles
            assert np.array_equal(np.ctypeslib.as_array(pa, (m,)), np.ctypeslib.as_array(pb, (m,))), name
    for i in range(da.num_materials):
        assert bytes(da.materials[i]) == bytes(db.materials[i])
    for i in range(da.num_lights):
        assert bytes(da.lights[i]) == bytes(db.lights[i])
    assert list(da.camera.xform) == list(db.camera.xform) and da.camera.vfov == db.camera.vfov
    assert list(da.env_ambient) == list(db.env_ambient) and list(da.env_color) == list(db.env_color)


@pytest.mark.parametrize("config,variant", [(0, 0), (1, 10), (2, 10), (3, 2000)])
def test_xml_round_trip(config, variant):
    """generator -> reference-format XML (%.9g) -> loader gives the same scene"""
    a = S.HostScene.generate(config, variant=variant)
    b = S.HostScene.load_xml(text=a.to_xml())
    if config == 2 or config == 1:
        # target camera: the loader recomputes the matrix from position/target
        a.desc.camera.xform[:] = b.desc.camera.xform[:]
    _same_scene(a, b)


def test_generators_deterministic_and_sized():
    a = S.HostScene.generate(1)
    assert a.desc.num_triangles == 1_000_000
    assert a.desc.num_objects == 41
    b = S.HostScene.generate(1)
    n = 9 * 1000
    assert np.array_equal(np.ctypeslib.as_array(a.desc.tri_pos, (n,)), np.ctypeslib.as_array(b.desc.tri_pos, (n,)))
    c = S.HostScene.generate(0)
    assert c.desc.num_objects == 64
    refl = [c.desc.materials[i].reflect for i in range(c.desc.num_materials)]
    assert sum(1 for r in refl if r > 0) == 16
    ico = S.HostScene.generate(2)
    assert ico.desc.num_triangles == 12 * 2 + 20 * 50 * 50


def test_attribute_typing_quirks():
    """src/xmltree.c:76-95: numbers must start with a digit or sign+digit
    (".5" is a string, Q16); '.' makes FLT, several values make VEC"""
    xml = """<scene>
      <env ambient="0.5" bgcolor="0.1 0.2 0.3" ior=".5"/>
      <material name="m" shader="phong"><mattr name="diffuse" value="1 0 0"/>
        <mattr name="shininess" value="7"/></material>
      <object type="sphere"><matref name="m"/><xform pos="1 2 3"/><sphere rad="2.5"/></object>
    </scene>"""
    h = S.HostScene.load_xml(text=xml)
    d = h.desc
    assert list(d.env_ambient) == [0.5, 0.5, 0.5, 1.0]
    assert list(d.env_color) == pytest.approx([0.1, 0.2, 0.3, 1.0])
    assert d.env_ior == 1.0          # ".5" is not a number -> ignored
    m = d.materials[0]
    assert list(m.diffuse) == [1, 0, 0, 1] and m.shininess == 7.0 and m.flags & abi.SRAY_MAT_PHONG
    o = d.objects[0]
    assert o.radius == 2.5 and (o.xform[3], o.xform[7], o.xform[11]) == (1, 2, 3)
    assert o.material == 0


def test_material_defaults():
    """Material(): diffuse 1, specular/reflect/refract 0, ior 1.5, shininess
    from def_attr = 1, shade_undef without shader="phong" (src/material.cc:77-89)"""
    h = S.HostScene.load_xml(text='<scene><material name="m"/>'
                                  '<object type="sphere"><matref name="m"/><sphere/></object></scene>')
    m = h.desc.materials[0]
    assert list(m.diffuse) == [1, 1, 1, 1] and m.reflect == 0 and m.refract == 0 and m.ior == 1.5
    assert m.shininess == 1.0 and not (m.flags & abi.SRAY_MAT_PHONG)
    assert h.desc.objects[0].radius == 1.0


@pytest.mark.parametrize("xml,frag", [
    ("<scene><object type='sphere'><matref name='nope'/><sphere/></object></scene>", "does not exist"),
    ("<scene><object type='cube'/></scene>", "unknown type"),
    ("<scene><object type='sphere'></object></scene>", "<sphere>"),
    ("<scene><object", "unterminated"),
    ("<scene><object type='mesh'><mesh><vertex id='0' val='0 0 0'/><face><vref vertex='5'/>"
     "<vref vertex='0'/><vref vertex='0'/></face></mesh></object></scene>", "missing vertex"),
    ("<scene><camera type='target'/></scene>", "no <target>"),
])
def test_load_errors_are_reported(xml, frag):
    with pytest.raises(S.SrayError) as e:
        S.HostScene.load_xml(text=xml)
    assert e.value.code == abi.SRAY_ERR_PARSE and frag in str(e.value)


def test_mesh_load_ids_and_attributes():
    """triangle ids = face order; vertex elements may come unordered (sorted by id)"""
    xml = """<scene><object type="mesh"><mesh>
      <vertex id="1" val="1 0 0"/><vertex id="0" val="0 0 0"/><vertex id="2" val="0 1 0"/>
      <normal id="0" val="0 0 1"/>
      <face><vref vertex="0" normal="0"/><vref vertex="1" normal="0"/><vref vertex="2" normal="0"/></face>
      <face><vref vertex="2"/><vref vertex="1"/><vref vertex="0"/></face>
    </mesh></object></scene>"""
    h = S.HostScene.load_xml(text=xml)
    d = h.desc
    tp = np.ctypeslib.as_array(d.tri_pos, (18,)).reshape(2, 3, 3)
    assert np.array_equal(tp[0], [[0, 0, 0], [1, 0, 0], [0, 1, 0]])
    assert np.array_equal(tp[1], [[0, 1, 0], [1, 0, 0], [0, 0, 0]])
    tn = np.ctypeslib.as_array(d.tri_normal, (18,)).reshape(2, 3, 3)
    assert np.all(tn[:, :, 2] == 1)


def _check_tree(hs, level):
    info, lo, hi, items = S.octree_host(hs.desc_ptr(), level)
    ref = oracle.OracleScene(hs.desc_ptr(), 32)
    oinfo, oitems, ocounts = ref.tree(level)
    # identical leaves, in the reference's depth-first order
    assert info.num_leaves == oinfo.num_leaves and info.height == oinfo.height
    assert np.array_equal(items, oitems)
    return info, lo, hi, items


def _leaf_rule(hs, level, lo, hi, items):
    """every stored leaf holds exactly the items whose box overlaps it
    (closed intervals, src/aabb.cc:53-65), in ascending index order; the
    children of a node are its 8 octants around the box centre"""
    d = hs.desc
    if level < 0:
        return  # object boxes need the object transforms; covered by the item comparison
    m = d.meshes[level]
    tp = np.ctypeslib.as_array(d.tri_pos, (d.num_triangles * 9,)).reshape(-1, 3, 3)[m.tri_start:m.tri_start + m.tri_count]
    bmin, bmax = tp.min(1), tp.max(1)
    meta = hi[:, 3].view(np.uint32)
    first = lo[:, 3].view(np.uint32)
    leaves = np.flatnonzero(meta & LEAF)
    rng = np.random.default_rng(0)
    for k in rng.choice(leaves, min(300, len(leaves)), replace=False):
        cnt = meta[k] & ~np.uint32(LEAF)
        got = items[first[k]:first[k] + cnt]
        inside = np.all((bmin <= hi[k, :3]) & (bmax >= lo[k, :3]), axis=1)
        # depth limit: a leaf at max depth keeps all overlapping items; a leaf
        # above it holds at most the item limit
        assert np.array_equal(got, np.flatnonzero(inside)), k
    inner = np.flatnonzero(~meta & LEAF)
    for k in inner[:200]:
        c = (lo[k, :3] + hi[k, :3]) * np.float32(0.5)
        mask = meta[k]
        for bits in range(8):
            if not (mask >> bits) & 1:
                continue
            slot = first[k] + bits  # eight-slot child block
            want_lo = np.where([(bits >> a) & 1 for a in range(3)], c, lo[k, :3])
            want_hi = np.where([(bits >> a) & 1 for a in range(3)], hi[k, :3], c)
            assert np.array_equal(lo[slot, :3], want_lo) and np.array_equal(hi[slot, :3], want_hi)


@pytest.mark.parametrize("config,variant,levels", [(0, 0, [-1]), (1, 8, [-1, 0]), (2, 10, [-1, 12]),
                                                    (3, 3000, [-1, 0])])
def test_host_octree_matches_checker_and_rule(config, variant, levels):
    hs = S.HostScene.generate(config, variant=variant)
    for level in levels:
        info, lo, hi, items = _check_tree(hs, level)
        _leaf_rule(hs, level, lo, hi, items)


def test_empty_scene_tree():
    h = S.HostScene.load_xml(text="<scene/>")
    info, lo, hi, items = S.octree_host(h.desc_ptr(), -1)
    assert info.num_nodes == 1 and info.num_item_refs == 0


def test_scene_diag_equals_checker():
    """the gather-distance unit (scene octree root diagonal) from the host
    octree build equals the checker's"""
    import oracle
    for cfg, var in ((0, 0), (2, 10), (3, 500)):
        hs = S.HostScene.generate(cfg, variant=var)
        assert S.scene_diag(hs.desc_ptr()) == oracle.OracleScene(hs.desc_ptr(), 32).diag()


@pytest.mark.parametrize("ambient,bgcolor,ior,exp_amb,exp_bg,exp_ior", [
    # two words make a VEC; missing components are 1.0
    ("0.25 0.5", "2,5 1 7 9 11", "1.25 ", [0.25, 0.5, 1.0, 1.0], [2.0, 1.0, 7.0, 1.0], 1.25),
    # one word with a '.' is FLT, a sign and a digit open a number
    ("+0.5", "1 1e-1 0x10", "+1.5", [0.5, 0.5, 0.5, 1.0], [1.0, 0.1, 16.0, 1.0], 1.5),
    # INT: no '.'; ior takes only FLT, a single-word bgcolor only VEC
    ("3", "4", "3", [3.0, 3.0, 3.0, 1.0], None, 1.0),
    # leading space, a bare sign, a leading '.': strings, ignored
    (" 1", "- 1 2", ".5", None, None, 1.0),
])
def test_attribute_words(ambient, bgcolor, ior, exp_amb, exp_bg, exp_ior):
    """the word rules of src/xmltree.c:76-95 on values the loader reads as
    scalars or vectors: numbers open with a digit or sign+digit, two or more
    words make a VEC of the first four, a '.' anywhere makes FLT"""
    base = S.HostScene.load_xml(text="<scene><env/></scene>").desc
    xml = (f'<scene><env ambient="{ambient}" bgcolor="{bgcolor}" ior="{ior}"/></scene>')
    d = S.HostScene.load_xml(text=xml).desc
    assert list(d.env_ambient) == pytest.approx(exp_amb if exp_amb else list(base.env_ambient))
    assert list(d.env_color) == pytest.approx(exp_bg if exp_bg else list(base.env_color))
    assert d.env_ior == pytest.approx(exp_ior)
