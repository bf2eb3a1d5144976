"""The library's host C++ (XML loader, track evaluation, octree build,
generators, projection maps, image and photon I/O) built with AddressSanitizer
and UndefinedBehaviorSanitizer, and run over valid, malformed and truncated
inputs by tests/san_host_driver.cc.  Any sanitizer report fails the test.

The device code is not part of this build (GPU sanitizers are not used), and
the serving thread, which needs a device, is not covered here.
"""
import os
import shutil
import struct
import subprocess
import sys
import zlib

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, HERE)

import imgfiles as I  # noqa: E402
import scenes  # noqa: E402

CSRC = os.path.join(ROOT, "sray_amd", "csrc")
HOST_SRC = ["accel.cc", "anim_host.cc", "scene_xml.cc", "scenegen.cc", "capi_host.cc", "photon_host.cc", "io_host.cc"]
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer"]


def _build(out_dir):
    exe = os.path.join(out_dir, "san_host_driver")
    srcs = [os.path.join(CSRC, s) for s in HOST_SRC] + [os.path.join(HERE, "san_host_driver.cc")]
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-ffp-contract=off", *SAN, "-o", exe, *srcs, "-lz"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    return exe


def _png_with_size(w, h):
    """a 1x1 RGB PNG whose header claims w x h"""
    good = I.png_bytes(np.zeros((1, 1, 3), np.int64), 2, 8)
    ihdr = struct.pack(">IIBBBBB", w, h, 8, 2, 0, 0, 0)
    body = b"IHDR" + ihdr
    chunk = struct.pack(">I", len(ihdr)) + body + struct.pack(">I", zlib.crc32(body) & 0xFFFFFFFF)
    return good[:8] + chunk + good[8 + len(chunk):]


def _inputs(d):
    rng = np.random.default_rng(3)
    # scenes: animated, textured (with its maps), and malformed variants
    anim = scenes.animated_scene_xml(40)
    with open(f"{d}/anim.xml", "w") as f:
        f.write(anim)
    scenes.write_textures(d)
    with open(f"{d}/textured.xml", "w") as f:
        f.write(scenes.textured_scene_xml())
    for k, cut in enumerate(np.linspace(10, len(anim) - 10, 12).astype(int)):
        with open(f"{d}/cut{k:02d}.xml", "w") as f:
            f.write(anim[:cut])
    bad = {
        "badvref": '<scene><object type="mesh"><mesh><vertex id="0" val="0 0 0"/>'
                   '<face><vref vertex="0"/><vref vertex="7"/><vref vertex="-3"/></face></mesh></object></scene>',
        "twovert": '<scene><object type="mesh"><mesh><vertex id="0" val="0 0 0"/><vertex id="1" val="1 0 0"/>'
                   '<face><vref vertex="0"/><vref vertex="1"/></face></mesh></object></scene>',
        "fivevert": '<scene><object type="mesh"><mesh>' +
                    "".join(f'<vertex id="{i}" val="{i} {i % 2} 0"/>' for i in range(5)) +
                    "<face>" + "".join(f'<vref vertex="{i}"/>' for i in range(5)) + "</face></mesh></object></scene>",
        "nomat": '<scene><object type="sphere"><matref name="missing"/><sphere rad="1"/></object></scene>',
        "unknown": '<scene><object type="torus"/><light type="laser"/><camera type="fisheye"/></scene>',
        "huge": '<scene><object type="sphere"><xform pos="1e39 -1e39 nan" scale="0 0 0"/>'
                '<sphere rad="1e38"/></object><camera fov="0" shutter="-5"/></scene>',
        "keys": '<scene><object type="sphere"><xform time="40" pos="1 0 0"/><xform time="0" pos="0 0 0"/>'
                '<xform time="40" pos="2 0 0"/><sphere rad="1"/></object></scene>',
        "unclosed": '<scene><object type="sphere"><sphere rad="1">',
        "noquote": "<scene><env ambient=0.5/></scene>",
        "entities": '<scene><material name="&lt;&amp;&#65;&#x42;&bogus;" shader="phong"/></scene>',
        "empty": "",
        "deep": "<scene>" + "<a>" * 2000 + "</a>" * 2000 + "</scene>",
        "maps": '<scene><material name="m" shader="phong"><mattr name="diffuse" value="1 1 1" map="nothere.png"/>'
                '<mattr name="normal" value="1" map="img_trunc.png"/></material></scene>',
    }
    for name, text in bad.items():
        with open(f"{d}/{name}.xml", "w") as f:
            f.write(text)
    # images: every PNG colour type and depth, netpbm, PFM, and broken files
    imgs = {}
    for ctype, depth, nch in [(0, 1, 1), (0, 2, 1), (0, 4, 1), (0, 8, 1), (0, 16, 1), (2, 8, 3), (2, 16, 3),
                              (4, 8, 2), (4, 16, 2), (6, 8, 4), (6, 16, 4)]:
        s = rng.integers(0, 1 << depth, (7, 5, nch))
        imgs[f"img_png{ctype}_{depth}.png"] = I.png_bytes(s, ctype, depth, filters=(0, 1, 2, 3, 4))
    pal = rng.integers(0, 256, (16, 3))
    imgs["img_pal4.png"] = I.png_bytes(rng.integers(0, 16, (6, 9)), 3, 4, palette=pal, trns=[0, 99, 255])
    imgs["img_badidx.png"] = I.png_bytes(rng.integers(10, 16, (4, 4)), 3, 4, palette=pal[:4])
    imgs["img_nopal.png"] = I.png_bytes(rng.integers(0, 4, (4, 4)), 3, 2)
    imgs["img_adam7.png"] = I.png_bytes(rng.integers(0, 256, (4, 4, 3)), 2, 8, interlace=1)
    imgs["img_crc.png"] = I.png_bytes(rng.integers(0, 256, (4, 4, 3)), 2, 8, crc_ok=False)
    good = I.png_bytes(rng.integers(0, 256, (9, 11, 4)), 6, 8, filters=(4, 3))
    for k, cut in enumerate((8, 20, 33, len(good) // 2, len(good) - 13)):
        imgs[f"img_trunc{k}.png"] = good[:cut]
    imgs["img_trunc.png"] = good[:40]
    imgs["img_big.png"] = _png_with_size(1 << 20, 1 << 20)
    imgs["img_zero.png"] = _png_with_size(0, 5)
    imgs["img_p2.pgm"] = I.pnm_bytes(rng.integers(0, 1000, (3, 4)), 999, "2")
    imgs["img_p3.ppm"] = I.pnm_bytes(rng.integers(0, 256, (3, 4, 3)), 255, "3")
    imgs["img_p5.pgm"] = I.pnm_bytes(rng.integers(0, 65536, (5, 3)), 65535, "5")
    imgs["img_p6.ppm"] = I.pnm_bytes(rng.integers(0, 256, (5, 3, 3)), 255, "6")
    imgs["img_p6short.ppm"] = I.pnm_bytes(rng.integers(0, 256, (5, 3, 3)), 255, "6")[:-7]
    imgs["img_p3short.ppm"] = b"P3\n4 4\n255\n1 2 3\n"
    imgs["img_pnmbig.ppm"] = b"P6\n99999999 99999999\n255\n"
    imgs["img_pnmneg.pgm"] = b"P5\n-4 3\n0\n"
    imgs["img_pf.pfm"] = I.pfm_bytes(rng.random((3, 5, 3)).astype(np.float32))
    imgs["img_pfbe.pfm"] = I.pfm_bytes(rng.random((4, 2)).astype(np.float32), little=False)
    imgs["img_pfshort.pfm"] = I.pfm_bytes(rng.random((3, 5, 3)).astype(np.float32))[:-5]
    imgs["img_garbage"] = bytes(rng.integers(0, 256, 300).astype(np.uint8))
    imgs["img_empty"] = b""
    for name, data in imgs.items():
        I.write(f"{d}/{name}", data)
    # photon dumps
    lines = ["energy 1.500000"]
    for _ in range(50):
        v = rng.random(12)
        lines.append("<%f %f %f> <%f %f %f> <%f %f %f> <%f %f %f>" % tuple(v))
    text = "\n".join(lines) + "\n"
    for name, t in {"pm_good": text, "pm_short": text[:len(text) // 2 + 7], "pm_empty": "",
                    "pm_head": "energy\n", "pm_zero": "energy 0\n<1 2 3> <1 0 0> <0 1 0> <1 1 1>\n"}.items():
        with open(f"{d}/{name}", "w") as f:
            f.write(t)


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ is not installed")
def test_host_code_under_asan_ubsan(tmp_path):
    exe = _build(str(tmp_path))
    data = tmp_path / "in"
    data.mkdir()
    _inputs(str(data))
    env = dict(os.environ)
    env["ASAN_OPTIONS"] = "detect_leaks=1:abort_on_error=0:halt_on_error=1:detect_stack_use_after_return=1"
    env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1"
    env.pop("LD_PRELOAD", None)
    r = subprocess.run([exe, str(data)], capture_output=True, text=True, env=env, timeout=600)
    out = r.stdout + r.stderr
    assert "ERROR: AddressSanitizer" not in out and "ERROR: LeakSanitizer" not in out, out[-6000:]
    assert "runtime error:" not in out, out[-6000:]
    assert r.returncode == 0, out[-6000:]
    # the cases ran: valid inputs load, broken ones are refused
    assert "xml anim.xml rc 0" in out and "xml textured.xml rc 0" in out
    assert "image img_png6_16.png rc 0" in out and "image img_crc.png rc 0" not in out
    assert "image img_big.png rc 0" not in out and "photons pm_good rc 0 n 50" in out
    assert r.stdout.rstrip().endswith("done")
