"""Batching and serving: requests in flight together (sray_gpu_submit /
sray_gpu_poll, continuous batching) against the same requests rendered
alone (sray_gpu_render).

Bar: a request's image in the mix equals its image alone (every random
draw is keyed by its own seed, pixel, sample and path; every sample is
summed by one thread in queue order), to 1e-6 absolute, with identical
ray counts; admission by memory footprint, per-request cameras and the
scene-release guard behave as documented.
"""
import ctypes as C
import os
import time

import numpy as np
import pytest

import sray_amd as S
from sray_amd import abi

pytestmark = pytest.mark.gpu
TOL = 1e-6


def _desc_with_camera(hs, shutter=None, pos=None):
    d = abi.SceneDesc()
    C.memmove(C.byref(d), hs.desc_ptr(), C.sizeof(d))
    if shutter is not None:
        d.camera.shutter = shutter
    if pos is not None:
        d.camera.xform[3], d.camera.xform[7], d.camera.xform[11] = pos
    return d


def _requests(ctx):
    sph = S.HostScene.generate(0)
    room = S.HostScene.generate(2, variant=10)
    bat = S.HostScene.generate(3, variant=300)
    s_sph, s_room = ctx.upload(sph), ctx.upload(room)
    d_bat = _desc_with_camera(bat, shutter=40)
    s_bat = ctx.upload(C.pointer(d_bat))
    cam_room = _desc_with_camera(room, pos=(1.0, 4.0, -7.0))
    s_room2 = ctx.upload(C.pointer(cam_room))
    reqs = [
        (s_sph, S.render_params(160, 120, iter=5, seed=1), None),
        (s_room, S.render_params(96, 54, samples=(1, 8), max_var=0.05, iter=4, shadow_samples=4, seed=2), None),
        (s_sph, S.render_params(200, 150, rect=(30, 20, 130, 90), samples=(2, 4), max_var=0.001, iter=4,
                                flags=abi.SRAY_RENDER_REF_Q1, seed=3), None),
        (s_bat, S.render_params(128, 128, samples=(1, 4), max_var=0.01, iter=3, mblur=1, seed=4), None),
        # per-request camera: the room through the second scene's camera
        (s_room, S.render_params(80, 60, iter=3, seed=5), cam_room.camera),
        (s_sph, S.render_params(64, 64, iter=6, seed=6), None),
    ]
    alone_sid = {4: s_room2}
    keep = (sph, room, bat, d_bat, cam_room)
    return reqs, alone_sid, keep, (s_sph, s_room, s_bat, s_room2)


def _alone(ctx, reqs, alone_sid):
    out = []
    for k, (sid, p, cam) in enumerate(reqs):
        img, st = ctx.render(alone_sid.get(k, sid), p)
        out.append((img, st.rays()))
    return out


def test_mix_equals_alone(ctx):
    reqs, alone_sid, keep, sids = _requests(ctx)
    alone = _alone(ctx, reqs, alone_sid)
    tickets = [ctx.submit(sid, p, cam) for sid, p, cam in reqs]
    for k, t in enumerate(tickets):
        img, res = ctx.poll(t)
        err = float(np.abs(img - alone[k][0]).max())
        print(f"request {k}: max diff {err:.2e}, rays {res.rays} (alone {alone[k][1]}), "
              f"latency {res.latency_ms:.2f} ms, queued {res.queue_ms:.2f} ms")
        assert err <= TOL, k
        assert res.rays == alone[k][1], k
        assert res.latency_ms >= res.queue_ms >= 0.0
    for s in sids:
        ctx.release(s)


def test_requests_joining_in_flight(ctx):
    """continuous batching: requests submitted while others are on the GPU
    join at pass boundaries and still match their images alone"""
    reqs, alone_sid, keep, sids = _requests(ctx)
    alone = _alone(ctx, reqs, alone_sid)
    order = [1, 3, 0, 5, 2, 4]
    tickets, done = {}, {}
    for n, k in enumerate(order):
        sid, p, cam = reqs[k]
        tickets[k] = ctx.submit(sid, p, cam)
        if n % 2 == 1:
            prev = order[n - 1]
            got = ctx.poll(tickets[prev], wait=False)  # may or may not be done
            if got is not None:
                done[prev] = got
            time.sleep(0.002)
    for k, t in tickets.items():
        img, res = done[k] if k in done else ctx.poll(t)
        assert float(np.abs(img - alone[k][0]).max()) <= TOL, k
    for s in sids:
        ctx.release(s)


def test_many_small_scenes(ctx):
    """a batch of scenes of different sizes and resolutions (the batched
    workload's shape, scaled down): every image equals its plain render"""
    rng = np.random.default_rng(11)
    scenes, sids, reqs = [], [], []
    for k in range(12):
        hs = S.HostScene.generate(3, seed=100 + k, variant=int(rng.integers(200, 3000)))
        scenes.append(hs)
        sids.append(ctx.upload(hs))
        r = int(rng.integers(1, 5)) * 32
        p = S.render_params(r, r, samples=(1, 3), max_var=0.02, iter=3, seed=k)
        reqs.append((sids[-1], p))
    alone = [ctx.render(sid, p)[0] for sid, p in reqs]
    tickets = [ctx.submit(sid, p) for sid, p in reqs]
    for k, t in enumerate(tickets):
        img, res = ctx.poll(t)
        assert float(np.abs(img - alone[k]).max()) <= TOL, k
    for s in sids:
        ctx.release(s)


def test_admission_cap(ctx):
    """a tight cap admits requests one at a time (results unchanged); a
    request larger than the cap on its own fails with NOMEM"""
    hs = S.HostScene.generate(0)
    sid = ctx.upload(hs)
    p = S.render_params(128, 96, iter=4, seed=9)
    alone, _ = ctx.render(sid, p)
    ctx.set_admission_cap(1 << 62)
    t0 = ctx.submit(sid, p)
    ctx.poll(t0)  # starts the serving engine and sizes its queues
    used = ctx.scene_bytes(0)
    footprint = 128 * 96 * 49  # per-pixel arena bytes of one request
    ctx.set_admission_cap(used + int(2.5 * footprint))
    ts = [ctx.submit(sid, p) for _ in range(4)]
    for t in ts:
        img, _ = ctx.poll(t)
        assert np.array_equal(img, alone)
    ctx.set_admission_cap(1)
    t = ctx.submit(sid, p)
    with pytest.raises(S.SrayError) as e:
        ctx.poll(t)
    assert e.value.code == abi.SRAY_ERR_NOMEM
    ctx.set_admission_cap(200 << 30)
    ctx.release(sid)


def test_release_guard_and_bad_requests(ctx):
    hs = S.HostScene.generate(2, variant=10)
    sid = ctx.upload(hs)
    p = S.render_params(320, 180, samples=(1, 16), max_var=0.0, iter=5, shadow_samples=4)
    t = ctx.submit(sid, p)
    try:
        ctx.release(sid)
        released = True  # the request finished before the release call
    except S.SrayError as e:
        assert e.code == abi.SRAY_ERR_PENDING
        released = False
    ctx.poll(t)
    if not released:
        ctx.release(sid)
    with pytest.raises(S.SrayError):
        ctx.submit(sid, p)
    with pytest.raises(S.SrayError):
        ctx.submit(999, S.render_params(16, 16))


def test_device_outputs_one_scene(ctx):
    """requests of one scene (the single-scene traversal variant) writing
    their images to device buffers equal their images alone"""
    hs = S.HostScene.generate(0)
    sid = ctx.upload(hs)
    ps = [S.render_params(96, 72, iter=5, seed=20 + k) for k in range(5)]
    ps.append(S.render_params(120, 90, rect=(10, 5, 70, 60), samples=(1, 4), max_var=0.002, iter=4, seed=30))
    alone = [ctx.render(sid, p)[0] for p in ps]
    outs = [ctx.malloc((p.x1 - p.x0) * (p.y1 - p.y0) * 16) for p in ps]
    tickets = [ctx.submit(sid, p, out=o) for p, o in zip(ps, outs)]
    for k, t in enumerate(tickets):
        img, res = ctx.poll(t)
        assert img is None and res.rays > 0
        p = ps[k]
        got = ctx.to_host(outs[k], (p.x1 - p.x0) * (p.y1 - p.y0) * 16, np.float32).reshape(alone[k].shape)
        assert float(np.abs(got - alone[k]).max()) <= TOL, k
    for o in outs:
        ctx.free(o)
    r = abi.Request()
    r.scene_id, r.params, r.flags = sid, ps[0], 0x100
    t = C.c_uint64()
    buf = np.empty(96 * 72 * 4, np.float32)
    assert S.lib().sray_gpu_submit(ctx._h, C.byref(r), buf.ctypes.data_as(C.POINTER(C.c_float)),
                                   C.byref(t)) == abi.SRAY_ERR_ARG
    ctx.release(sid)


def test_submit_many_equals_alone(ctx):
    """requests queued together (one call) match their images alone; a bad
    request anywhere in the list queues none of them"""
    reqs, alone_sid, keep, sids = _requests(ctx)
    alone = _alone(ctx, reqs, alone_sid)
    tickets = ctx.submit_many([(sid, p, cam, None) for sid, p, cam in reqs])
    assert len(tickets) == len(reqs) and tickets == sorted(tickets)
    for k, t in enumerate(tickets):
        img, res = ctx.poll(t)
        assert float(np.abs(img - alone[k][0]).max()) <= TOL, k
        assert res.rays == alone[k][1], k
    bad = [(sids[0], reqs[0][1], None, None), (999, reqs[0][1], None, None)]
    with pytest.raises(S.SrayError):
        ctx.submit_many(bad)
    for s in sids:
        ctx.release(s)


def test_serving_counters_start_at_zero():
    """a new context's serving engine starts its ray counters from zero,
    also when its counter buffers land on memory freed dirty just before"""
    c = S.Context(0)
    try:
        hs = S.HostScene.generate(0)
        sid = c.upload(hs)
        p = S.render_params(160, 120, iter=5, seed=1)
        alone, ast = c.render(sid, p)
        junk = [c.to_device(np.full(n, 0xFF, np.uint8)) for n in (5, 80, 120, 256) for _ in range(8)]
        for d in junk:
            c.free(d)
        img, res = c.poll(c.submit(sid, p))
        assert float(np.abs(img - alone).max()) <= TOL
        assert res.rays == ast.rays()
        c.release(sid)
    finally:
        c.close()


def _records(fd_r):
    data = b""
    while True:
        try:
            chunk = os.read(fd_r, 65536)
        except BlockingIOError:
            break
        if not chunk:
            break
        data += chunk
    assert len(data) % abi.SRAY_STATUS_RECORD == 0
    recs = []
    for k in range(0, len(data), abi.SRAY_STATUS_RECORD):
        rec = data[k:k + abi.SRAY_STATUS_RECORD]
        text = rec.split(b"\0", 1)[0].decode()
        assert rec == text.encode() + b"\0" * (abi.SRAY_STATUS_RECORD - len(text))
        st, *args = text.split()
        recs.append((st, [int(a) for a in args]))
    return recs


def test_status_stream(ctx):
    """renders and requests report on the status stream in the reference's
    record format (src/render.cc:279-289): a render is one frame of one
    block, a request one block; 'd' carries the time"""
    hs = S.HostScene.generate(0)
    sid = ctx.upload(hs)
    r, w = os.pipe()
    os.set_blocking(r, False)
    ctx.set_status_fd(w)
    try:
        p = S.render_params(160, 120, rect=(10, 20, 110, 80), iter=3, seed=1)
        ctx.render(sid, p)
        recs = _records(r)
        assert [x[0] for x in recs] == ["f", "s", "e", "d"]
        assert recs[0][1][:3] == [1, 1, 100] and recs[1][1] == [10, 20, 100, 60] and recs[2][1] == [10, 20, 100, 60]
        assert recs[3][1][0] >= 0
        tickets = [ctx.submit(sid, S.render_params(64, 48, rect=(k, 0, 32 + k, 16), iter=3, seed=k)) for k in range(3)]
        for t in tickets:
            ctx.poll(t)
        recs = _records(r)
        assert sorted(x[0] for x in recs) == ["d"] * 3 + ["e"] * 3 + ["s"] * 3
        assert sorted(tuple(x[1]) for x in recs if x[0] == "s") == [(k, 0, 32, 16) for k in range(3)]
        for k, (st, args) in enumerate(recs):
            if st == "e":
                assert recs[k + 1][0] == "d" and recs[k + 1][1][0] >= recs[k + 1][1][1] >= 0
    finally:
        ctx.set_status_fd(-1)
        os.close(r)
        os.close(w)
        ctx.release(sid)
