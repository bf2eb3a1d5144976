"""Texture maps on the host and in the checker: the image reader that
stands in for Image::load (src/img.cc:64-78), the material "map" loading
and search (src/material.cc:64-71, src/datapath.c:56-77), Texture::lookup
(src/texture.cc:60-101), the texture x value attributes
(src/material.inl:18-28) and get_bump_normal (src/shader.cc:293-316).

Pinning: image files are written byte by byte from the format
specifications (tests/imgfiles.py) and must read back to value / maxval;
lookups are held to what the definitions fix — exact texels under
nearest filtering, the centre texel for an affine texture under the
5-texel cross, a whole-image periodic stencil for repeat wrapping,
periodicity in u and v, the texture matrix as a change of coordinates —
and bump normals to the closed forms of flat and axis-aligned normal
maps in a known tangent frame.  libimago (the reference's image library)
is not in this image, so reads are pinned to the formats' specifications,
not to its outputs.
"""
import os

import numpy as np
import pytest

import imgfiles as I
import oracle
import sray_amd as S
from scenes import rays, textured_scene, textured_scene_xml, write_textures


def _same(a, b):
    return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))


# ---------------------------------------------------------------- reader
PNG_CASES = [(0, 1), (0, 2), (0, 4), (0, 8), (0, 16), (2, 8), (2, 16), (3, 1), (3, 2), (3, 4), (3, 8),
             (4, 8), (4, 16), (6, 8), (6, 16)]


@pytest.mark.parametrize("ctype,depth", PNG_CASES)
def test_png_every_type_and_filter(tmp_path, ctype, depth):
    rng = np.random.default_rng(ctype * 100 + depth)
    nch = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}[ctype]
    h, w = 7, 13  # odd width: sub-byte rows end mid-byte
    pal = None
    if ctype == 3:
        pal = rng.integers(0, 256, (1 << depth, 3))
        s = rng.integers(0, 1 << depth, (h, w, 1))
    else:
        s = rng.integers(0, 1 << depth, (h, w, nch))
    p = I.write(tmp_path / "t.png", I.png_bytes(s, ctype, depth, filters=(0, 1, 2, 3, 4), palette=pal))
    got = S.read_image(p)
    assert _same(got, I.png_expected(s, ctype, depth, palette=pal))


def test_png_transparency(tmp_path):
    rng = np.random.default_rng(3)
    pal = rng.integers(0, 256, (4, 3))
    s = rng.integers(0, 4, (5, 6))
    trns = [0, 77]  # index 0 transparent, 1 partly, 2-3 opaque
    p = I.write(tmp_path / "p.png", I.png_bytes(s, 3, 2, palette=pal, trns=trns))
    assert _same(S.read_image(p), I.png_expected(s, 3, 2, palette=pal, trns=trns))
    g = rng.integers(0, 4, (5, 6))
    p = I.write(tmp_path / "g.png", I.png_bytes(g, 0, 8, trns=[0, 2]))
    e = I.png_expected(g, 0, 8, trns=[0, 2])
    assert (e[:, :, 3] == 0).any() and _same(S.read_image(p), e)
    c = rng.integers(0, 2, (5, 6, 3))
    p = I.write(tmp_path / "c.png", I.png_bytes(c, 2, 8, trns=[0, 1, 0, 0, 0, 1]))
    assert _same(S.read_image(p), I.png_expected(c, 2, 8, trns=[0, 1, 0, 0, 0, 1]))


@pytest.mark.parametrize("kind,maxval", [("2", 255), ("3", 255), ("3", 1000), ("5", 255), ("5", 65535),
                                         ("6", 255), ("6", 4095)])
def test_netpbm(tmp_path, kind, maxval):
    rng = np.random.default_rng(int(kind) * 7 + maxval)
    shape = (6, 9, 3) if kind in "36" else (6, 9)
    s = rng.integers(0, maxval + 1, shape)
    p = I.write(tmp_path / "t.pnm", I.pnm_bytes(s, maxval, kind))
    assert _same(S.read_image(p), I.pnm_expected(s, maxval))


@pytest.mark.parametrize("little", [True, False])
@pytest.mark.parametrize("grey", [True, False])
def test_pfm_rows_and_byte_order(tmp_path, little, grey):
    rng = np.random.default_rng(11)
    v = rng.normal(size=(5, 7) if grey else (5, 7, 3)).astype(np.float32)
    p = I.write(tmp_path / "t.pfm", I.pfm_bytes(v, little))
    got = S.read_image(p)
    assert _same(got, I.pfm_expected(v))
    assert got[0, 0, 0] == v[0, 0] if grey else got[0, 0, 0] == v[0, 0, 0]  # top row first


def test_writers_read_back(tmp_path):
    rng = np.random.default_rng(5)
    img = rng.random((9, 11, 4)).astype(np.float32) * 1.2 - 0.1
    S.write_pfm(tmp_path / "o.pfm", img)
    got = S.read_image(tmp_path / "o.pfm")
    assert _same(got[:, :, :3].copy(), img[:, :, :3].copy()) and np.all(got[:, :, 3] == 1.0)
    S.write_png(tmp_path / "o.png", img)
    b = np.clip(np.trunc(img.astype(np.float64) * 255.0), 0, 255)
    assert _same(S.read_image(tmp_path / "o.png"), (b.astype(np.float32) / np.float32(255)))


def test_reader_refuses_bad_files(tmp_path):
    s = np.zeros((4, 4, 3), np.int64)
    bad = {
        "interlaced.png": I.png_bytes(s, 2, 8, interlace=1),
        "crc.png": I.png_bytes(s, 2, 8, crc_ok=False),
        "short.png": I.png_bytes(s, 2, 8)[:40],
        "short.ppm": I.pnm_bytes(s, 255, "6")[:-5],
        "short.pfm": I.pfm_bytes(np.zeros((4, 4, 3), np.float32))[:-3],
        "unknown.img": b"GIF89a....",
    }
    for name, data in bad.items():
        with pytest.raises(S.SrayError):
            S.read_image(I.write(tmp_path / name, data))
    with pytest.raises(S.SrayError):
        S.read_image(tmp_path / "absent.png")


# ---------------------------------------------------------------- lookups
def _tex_scene(tmp_path, texels, name="t.pfm", filt=None, wrap=None, xform=None):
    """one sphere whose diffuse attribute maps an RGB float texture; the
    descriptor's filter / wrap / matrix set as asked before the checker
    copies it"""
    I.write(tmp_path / name, I.pfm_bytes(texels))
    x = ('<scene><material name="m" shader="phong"><mattr name="diffuse" value="1 1 1" map="%s"/></material>'
         '<object type="sphere"><matref name="m"/><sphere rad="1.0"/></object></scene>' % name)
    with open(tmp_path / "s.xml", "w") as f:
        f.write(x)
    hs = S.HostScene.load_xml(path=str(tmp_path / "s.xml"))
    t = hs.desc.textures[0]
    if filt is not None:
        t.filter = filt
    if wrap is not None:
        t.wrap = wrap
    if xform is not None:
        for k in range(12):
            t.xform[k] = xform[k]
    return hs


@pytest.fixture(params=[32, 64], ids=["f32", "f64"])
def prec(request):
    return request.param


def test_nearest_reads_each_texel(tmp_path, prec):
    rng = np.random.default_rng(1)
    h, w = 8, 16
    tex = rng.random((h, w, 3)).astype(np.float32)
    hs = _tex_scene(tmp_path, tex, filt=S.abi.SRAY_TEX_FILTER_NEAREST)
    o = oracle.OracleScene(hs.desc_ptr(), prec)
    yy, xx = np.mgrid[0:h, 0:w]
    uv = np.stack([(xx + 0.5) / w, (yy + 0.5) / h], -1).reshape(-1, 2)
    got = o.tex_lookup(0, uv)
    assert np.array_equal(got[:, :3], tex.reshape(-1, 3)) and np.all(got[:, 3] == 1.0)


def test_cross_filter_keeps_affine_texture(tmp_path, prec):
    """5-texel cross of an affine texture (exact in eighths) is the centre"""
    h, w = 8, 8
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float32)
    tex = np.stack([1 + xx / 8, 2 + yy / 8, 3 + (xx + yy) / 8], -1).astype(np.float32)
    hs = _tex_scene(tmp_path, tex)
    o = oracle.OracleScene(hs.desc_ptr(), prec)
    inner = [(x, y) for y in range(1, h - 1) for x in range(1, w - 1)]
    uv = np.array([((x + 0.5) / w, (y + 0.5) / h) for x, y in inner])
    got = o.tex_lookup(0, uv)
    want = np.array([tex[y, x] for x, y in inner])
    assert np.array_equal(got[:, :3], want)


def test_repeat_is_a_periodic_stencil(tmp_path, prec):
    rng = np.random.default_rng(2)
    h, w = 4, 8
    tex = rng.random((h, w, 3)).astype(np.float32)
    hs = _tex_scene(tmp_path, tex)
    o = oracle.OracleScene(hs.desc_ptr(), prec)
    t = tex.astype(np.float64)
    stencil = (t + np.roll(t, -1, 1) + np.roll(t, 1, 1) + np.roll(t, -1, 0) + np.roll(t, 1, 0)) / 5
    yy, xx = np.mgrid[0:h, 0:w]
    for shift in (0, 1, 3):  # whole periods away: same texels
        uv = np.stack([(xx + 0.5) / w + shift, (yy + 0.5) / h + 2 * shift], -1).reshape(-1, 2)
        got = o.tex_lookup(0, uv)[:, :3].astype(np.float64)
        assert np.allclose(got, stencil.reshape(-1, 3), rtol=0, atol=1e-6)


def test_negative_coordinates_truncate_toward_zero(tmp_path, prec):
    """(int)(c * size) truncates: a coordinate just below 0 reads texel 0,
    not the last one, and -1.25 texels reads -1, wrapped to size - 1"""
    rng = np.random.default_rng(6)
    h, w = 4, 8
    tex = rng.random((h, w, 3)).astype(np.float32)
    hs = _tex_scene(tmp_path, tex, filt=S.abi.SRAY_TEX_FILTER_NEAREST)
    o = oracle.OracleScene(hs.desc_ptr(), prec)
    got = o.tex_lookup(0, [(-0.5 / w, 0.5 / h), (-1.25 / w, -0.5 / h), (-0.5 / w, -1.5 / h)])[:, :3]
    assert np.array_equal(got[0], tex[0, 0])
    assert np.array_equal(got[1], tex[0, w - 1])
    assert np.array_equal(got[2], tex[h - 1, 0])


def test_clamp_reaches_one_past_the_row(tmp_path, prec):
    """clamp holds indices to [0, size]: the texel right of the last column
    is the next row's first (the row-major index of the reference's
    pointer arithmetic); far outside reads the edge"""
    h, w = 4, 4
    tex = np.arange(h * w * 3, dtype=np.float32).reshape(h, w, 3)
    hs = _tex_scene(tmp_path, tex, filt=S.abi.SRAY_TEX_FILTER_NEAREST, wrap=S.abi.SRAY_TEX_WRAP_CLAMP)
    o = oracle.OracleScene(hs.desc_ptr(), prec)
    got = o.tex_lookup(0, [(-5.0, 0.6), (1.0, 0.1), (1.0, 0.4), (30.0, 30.0), (0.3, -2.0)])[:, :3]
    flat = tex.reshape(-1, 3)
    assert np.array_equal(got[0], tex[2, 0])          # u < 0: column 0
    assert np.array_equal(got[1], flat[0 * w + w])    # u = 1: (row 0, column w) = row 1's first
    assert np.array_equal(got[2], flat[1 * w + w])
    assert np.array_equal(got[3], flat[-1])           # beyond the last texel: the last
    assert np.array_equal(got[4], tex[0, 1])


def test_texture_matrix_changes_coordinates(tmp_path, prec):
    rng = np.random.default_rng(4)
    tex = rng.random((8, 8, 3)).astype(np.float32)
    uv = rng.random((200, 2)) * 0.45
    hs1 = _tex_scene(tmp_path, tex)
    hs2 = _tex_scene(tmp_path, tex, xform=[2, 0, 0, 0.125, 0, 2, 0, 0.25, 0, 0, 1, 0])
    a = oracle.OracleScene(hs1.desc_ptr(), prec).tex_lookup(0, uv * 2 + [0.125, 0.25])
    b = oracle.OracleScene(hs2.desc_ptr(), prec).tex_lookup(0, uv)
    assert np.array_equal(a, b)


def test_mapped_attribute_is_texture_times_value(tmp_path, prec):
    """a constant 0.5 texture on diffuse (1.0, 0.6, 0.2) renders as the
    unmapped diffuse (0.5, 0.3, 0.1): halving is exact"""
    I.write(tmp_path / "half.pfm", I.pfm_bytes(np.full((4, 4, 3), 0.5, np.float32)))
    def scene(kd, mp):
        m = f' map="{mp}"' if mp else ""
        return ('<scene><env ambient="0.2 0.2 0.2"/><material name="m" shader="phong">'
                f'<mattr name="diffuse" value="{kd}"{m}/><mattr name="specular" value="0.3 0.3 0.3"/></material>'
                '<object type="sphere"><matref name="m"/><xform pos="0 0 5"/><sphere rad="1.5"/></object>'
                '<light type="point" color="1 1 1"><xform pos="3 4 0"/></light></scene>')
    imgs = []
    for kd, mp in (("1.0 0.6 0.2", "half.pfm"), ("0.5 0.3 0.1", None)):
        with open(tmp_path / "s.xml", "w") as f:
            f.write(scene(kd, mp))
        hs = S.HostScene.load_xml(path=str(tmp_path / "s.xml"))
        assert hs.desc.num_textures == (1 if mp else 0)
        img, _ = oracle.OracleScene(hs.desc_ptr(), prec).render(S.render_params(48, 36, iter=2, seed=9))
        imgs.append(img)
    assert _same(imgs[0], imgs[1]) and imgs[0].std() > 0.01


# ---------------------------------------------------------------- bump
def _floor_scene(tmp_path, colour, scale, mapped=True):
    """a floor quad at y = 0 with normal +y and tangent +x; its normal map a
    constant colour (float, exact)"""
    I.write(tmp_path / "n.pfm", I.pfm_bytes(np.broadcast_to(np.float32(colour), (4, 4, 3)).copy()))
    m = ' map="n.pfm"' if mapped else ""
    x = ('<scene><material name="m" shader="phong"><mattr name="diffuse" value="1 1 1"/>'
         f'<mattr name="normal" value="{scale}"{m}/></material>'
         '<object type="mesh"><matref name="m"/><mesh>'
         '<vertex id="0" val="-5 0 -5"/><vertex id="1" val="5 0 -5"/><vertex id="2" val="5 0 5"/>'
         '<vertex id="3" val="-5 0 5"/><normal id="0" val="0 1 0"/><tangent id="0" val="1 0 0"/>'
         '<texcoord id="0" val="0 0 0"/>'
         '<face><vref vertex="0" normal="0" tangent="0" texcoord="0"/><vref vertex="2" normal="0" tangent="0" '
         'texcoord="0"/><vref vertex="1" normal="0" tangent="0" texcoord="0"/></face>'
         '<face><vref vertex="0" normal="0" tangent="0" texcoord="0"/><vref vertex="3" normal="0" tangent="0" '
         'texcoord="0"/><vref vertex="2" normal="0" tangent="0" texcoord="0"/></face>'
         '</mesh></object></scene>')
    with open(tmp_path / "s.xml", "w") as f:
        f.write(x)
    return S.HostScene.load_xml(path=str(tmp_path / "s.xml"))


def _normals(hs, prec, from_below=False):
    rng = np.random.default_rng(8)
    o = rng.uniform(-3, 3, (64, 3))
    o[:, 1] = -2.0 if from_below else 2.0
    d = np.zeros_like(o)
    d[:, 1] = 4.0 if from_below else -4.0
    d[:, 0] = rng.uniform(-0.5, 0.5, 64)
    out = oracle.OracleScene(hs.desc_ptr(), prec).shade_normal(rays(o, d))
    assert not np.isnan(out).any()
    return out[:, :3].astype(np.float64), out[:, 3:].astype(np.float64)


@pytest.mark.parametrize("colour,scale,want", [
    ((0.5, 0.5, 1.0), 1.7, (0.0, 1.0, 0.0)),      # flat map: the surface normal
    ((1.0, 0.5, 0.5), 0.6, (1.0, 0.0, 0.0)),      # +x of the map: the tangent
    ((0.5, 1.0, 0.5), 2.0, (0.0, 0.0, -1.0)),     # +y: normal x tangent
    ((0.75, 0.5, 1.0), 2.0, (0.5 ** 0.5, 0.5 ** 0.5, 0.0)),  # (1, 0, 1)/sqrt 2: tangent + normal
    ((0.5, 0.25, 1.0), 2.0, (0.0, 0.5 ** 0.5, 0.5 ** 0.5)),  # (0, -1, 1)/sqrt 2
])
def test_bump_frame_closed_forms(tmp_path, prec, colour, scale, want):
    hs = _floor_scene(tmp_path, colour, scale)
    for below in (False, True):  # the frame is the unflipped normal's on both sides
        n, g = _normals(hs, prec, below)
        assert np.allclose(g, [1.0, 0.0, 0.0], atol=1e-6)
        assert np.allclose(n, want, atol=2e-6), (below, n[0])


def test_normal_attribute_without_map_unflips(tmp_path, prec):
    hs = _floor_scene(tmp_path, (0.5, 0.5, 1.0), 1.0, mapped=False)
    n, _ = _normals(hs, prec, from_below=True)
    assert np.allclose(n, [0.0, 1.0, 0.0], atol=1e-7)


def test_no_normal_attribute_faces_the_ray(tmp_path, prec):
    hs = _floor_scene(tmp_path, (0.5, 0.5, 1.0), 1.0)
    d = hs.desc.materials[0]
    d.flags &= ~S.abi.SRAY_MAT_HAS_NORMAL
    n, _ = _normals(hs, prec, from_below=True)
    assert np.allclose(n, [0.0, -1.0, 0.0], atol=1e-7)


# ---------------------------------------------------------------- loading
def test_scene_maps_load_and_share_files(tmp_path):
    hs = textured_scene(str(tmp_path))
    d = hs.desc
    assert d.num_textures == 6  # bumps.png and shin.pgm shared; the missing file loads nothing
    names = {}
    for i in range(d.num_materials):
        m = d.materials[i]
        names[i] = [int(v) for v in m.map]
    assert all(0 <= v <= d.num_textures for ms in names.values() for v in ms)
    assert names[1][S.abi.SRAY_MAP_NORMAL] == names[4][S.abi.SRAY_MAP_NORMAL] > 0
    assert names[6] == [0] * 7  # the missing map: plain colour
    assert d.materials[5].flags & S.abi.SRAY_MAT_HAS_NORMAL and names[5][S.abi.SRAY_MAP_NORMAL] == 0
    t = d.textures[0]
    assert (t.width, t.height, t.filter, t.wrap) == (16, 16, 0, 0)
    assert list(t.xform) == [1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0]


def test_search_path_and_absolute_names(tmp_path, monkeypatch):
    tex_dir = tmp_path / "maps"
    tex_dir.mkdir()
    write_textures(str(tex_dir))
    # scene given as text: "." then SRAY_DATA_PATH
    monkeypatch.setenv("SRAY_DATA_PATH", f"{tmp_path}/nowhere:{tex_dir}")
    hs = S.HostScene.load_xml(text=textured_scene_xml())
    assert hs.desc.num_textures == 6
    monkeypatch.delenv("SRAY_DATA_PATH")
    monkeypatch.chdir(tmp_path)
    assert S.HostScene.load_xml(text=textured_scene_xml()).desc.num_textures == 0
    x = textured_scene_xml().replace('map="checker.png"', f'map="{tex_dir}/checker.png"')
    assert S.HostScene.load_xml(text=x).desc.num_textures == 1


def test_written_scene_keeps_its_maps(tmp_path):
    hs = textured_scene(str(tmp_path))
    text = hs.to_xml()
    assert 'map="checker.png"' in text and 'name="normal"' in text
    with open(tmp_path / "again.xml", "w") as f:
        f.write(text)
    hs2 = S.HostScene.load_xml(path=str(tmp_path / "again.xml"))
    a, b = hs.desc, hs2.desc
    assert a.num_textures == b.num_textures
    for i in range(a.num_materials):
        assert list(a.materials[i].map) == list(b.materials[i].map)
        assert a.materials[i].flags == b.materials[i].flags
    p = S.render_params(64, 48, iter=3, seed=5)
    i1, _ = oracle.OracleScene(hs.desc_ptr(), 32).render(p)
    i2, _ = oracle.OracleScene(hs2.desc_ptr(), 32).render(p)
    assert float(np.abs(i1 - i2).max()) < 1e-3


def test_textured_render_float_against_double(tmp_path):
    """the float restatement against the double one on the textured scene"""
    hs = textured_scene(str(tmp_path))
    p = S.render_params(96, 72, samples=(1, 2), iter=4, seed=21)
    a, _ = oracle.OracleScene(hs.desc_ptr(), 32).render(p)
    b, _ = oracle.OracleScene(hs.desc_ptr(), 64).render(p)
    e = float(np.sqrt(np.mean((a.astype(np.float64) - b) ** 2)))
    assert e < 5e-3, e
