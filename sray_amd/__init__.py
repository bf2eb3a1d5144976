"""sray_amd — Python bindings (ctypes) over the C ABI of libsray_gpu.so.

The product is the native library (gfx950 HIP kernels + host C++); this
module only loads it and converts numpy arrays.  Loading fails loudly if the
library has not been built: there is no Python or CPU fallback.
"""
import ctypes as C
import os

import numpy as np

from . import abi
from .abi import (Hit, OctreeInfo, Ray, RenderParams, SceneDesc, Stats, TraceCounts)

_HERE = os.path.dirname(os.path.abspath(__file__))
BUILD_DIR = os.path.join(_HERE, "_build")
LIB_PATH = os.path.join(BUILD_DIR, "libsray_gpu.so")
HOST_LIB_PATH = os.path.join(BUILD_DIR, "libsray_host.so")

_lib = None


class SrayError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"sray error {code}: {msg}")
        self.code = code


def lib():
    """The loaded libsray_gpu.so (raises if it is missing)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build() "
                              "(or make -C sray_amd/csrc)")
        _lib = abi.declare(C.CDLL(LIB_PATH))
    return _lib


def _host_err():
    return lib().sray_host_last_error().decode(errors="replace")


def abi_sizes():
    n = len(abi.ABI_STRUCTS)
    out = (C.c_uint32 * n)()
    k = lib().sray_abi_sizes(out, n)
    return list(out)[:k]


# ---------------------------------------------------------------------------
# host scenes
# ---------------------------------------------------------------------------
class HostScene:
    """A flat scene description owned by the native library (loaded from
    reference-format XML or generated)."""

    def __init__(self, handle):
        self._h = C.c_void_p(handle)

    @classmethod
    def load_xml(cls, path=None, text=None):
        h = C.c_void_p()
        rc = lib().sray_scene_load_xml(path.encode() if path else None, text.encode() if text else None,
                                       C.byref(h))
        if rc:
            raise SrayError(rc, _host_err())
        return cls(h.value)

    @classmethod
    def generate(cls, config, seed=None, variant=0):
        """config 0 spheres, 1 heightfield, 2 photon room, 3 batch scene."""
        if seed is None:
            seed = 0x5EED0000 + config
        h = C.c_void_p()
        rc = lib().sray_scene_generate(int(config), int(seed), int(variant), C.byref(h))
        if rc:
            raise SrayError(rc, _host_err())
        return cls(h.value)

    @property
    def desc(self):
        return lib().sray_scene_desc_of(self._h).contents

    def desc_ptr(self):
        return lib().sray_scene_desc_of(self._h)

    def set_frame(self, msec):
        """evaluates the scene's keyframe tracks for render(msec): the frame
        interval, the transforms at msec and per-msec rows over the interval"""
        rc = lib().sray_scene_set_frame(self._h, int(msec))
        if rc:
            raise SrayError(rc, _host_err())

    def to_xml(self):
        n = C.c_size_t()
        lib().sray_scene_write_xml(self._h, None, 0, C.byref(n))
        buf = C.create_string_buffer(n.value)
        rc = lib().sray_scene_write_xml(self._h, buf, n.value, C.byref(n))
        if rc:
            raise SrayError(rc, "write_xml")
        return buf.value.decode()

    def close(self):
        if self._h:
            lib().sray_scene_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def octree_host(desc_ptr, level):
    """Flattened octree of one level as built on the host:
    (info, lo[N,4] float32, hi[N,4] float32, items[M] uint32)."""
    info = OctreeInfo()
    rc = lib().sray_octree_build_host(desc_ptr, level, C.byref(info), None, None, None, 0, 0)
    if rc:
        raise SrayError(rc, "octree_build_host")
    lo = np.zeros((info.num_nodes, 4), np.float32)
    hi = np.zeros((info.num_nodes, 4), np.float32)
    items = np.zeros(max(info.num_item_refs, 1), np.uint32)
    fp = C.POINTER(C.c_float)
    rc = lib().sray_octree_build_host(desc_ptr, level, C.byref(info), lo.ctypes.data_as(fp),
                                      hi.ctypes.data_as(fp), items.ctypes.data_as(C.POINTER(C.c_uint32)),
                                      info.num_nodes, info.num_item_refs)
    if rc:
        raise SrayError(rc, "octree_build_host")
    return info, lo, hi, items[:info.num_item_refs]


def scene_diag(desc_ptr):
    """diagonal of the scene octree's root box (Scene::build_tree,
    src/scene.cc:384-387), the unit of render_params.gather_dist"""
    info = OctreeInfo()
    rc = lib().sray_octree_build_host(desc_ptr, -1, C.byref(info), None, None, None, 0, 0)
    if rc:
        raise SrayError(rc, "octree_build_host")
    d = [float(info.root_max[k]) - float(info.root_min[k]) for k in range(3)]
    return float(np.sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]))


# ---------------------------------------------------------------------------
# numpy <-> struct arrays
# ---------------------------------------------------------------------------
RAY_DTYPE = np.dtype([("origin", np.float32, 3), ("dir", np.float32, 3), ("time", np.int32), ("pad", np.int32)])
HIT_DTYPE = np.dtype([("t", np.float32), ("obj", np.int32), ("prim", np.int32), ("rank", np.uint32)])


def make_rays(origins, dirs, time=0):
    r = np.zeros(len(origins), RAY_DTYPE)
    r["origin"] = origins
    r["dir"] = dirs
    r["time"] = time
    return r


def render_params(width, height, rect=None, samples=(1, 1), max_var=0.005, min_energy=1e-4, iter=7,
                  shadow_samples=1, diffuse_samples=1, frame_time=0, mblur=0, caust_photons=0, gi_photons=0,
                  gather_dist=0.001, seed=0x5EED, flags=0, gather_k=0):
    """sray_render_params with the reference's option defaults (src/opt.c:357-388)."""
    p = RenderParams()
    p.width, p.height = width, height
    x0, y0, x1, y1 = rect if rect else (0, 0, width, height)
    p.x0, p.y0, p.x1, p.y1 = x0, y0, x1, y1
    p.min_samples, p.max_samples = samples
    p.max_var, p.min_energy, p.iter = max_var, min_energy, iter
    p.shadow_samples, p.diffuse_samples = shadow_samples, diffuse_samples
    p.frame_time, p.mblur = frame_time, mblur
    p.caust_photons, p.gi_photons, p.gather_dist = caust_photons, gi_photons, gather_dist
    p.seed, p.flags = seed, flags
    p.gather_k = gather_k
    return p


def photon_params(caust=0, gi=0, energy=300.0, iter=7, t0=0, t1=0, seed=0x5EED):
    """sray_photon_params (-c, -g, -photonenergy, -i of src/opt.c)"""
    pp = abi.PhotonParams()
    pp.caust_photons, pp.gi_photons, pp.photon_energy = caust, gi, energy
    pp.iter, pp.t0, pp.t1, pp.seed = iter, t0, t1, seed
    return pp


def projmap(desc_ptr, light, kind, t0=0, t1=0, seed=0x5EED):
    """projection map of one light (host): (cells[32, 64] uint8, coverage)"""
    cells = np.zeros(64 * 32, np.uint8)
    cov = C.c_double()
    rc = lib().sray_projmap_build(desc_ptr, light, kind, t0, t1, seed, cells.ctypes.data, C.byref(cov))
    if rc:
        raise SrayError(rc, _host_err())
    return cells.reshape(32, 64), cov.value


def projmap_cell(x, y, z):
    u, v = C.c_int(), C.c_int()
    lib().sray_projmap_cell(x, y, z, C.byref(u), C.byref(v))
    return u.value, v.value


def _io_check(rc):
    if rc:
        raise SrayError(rc, lib().sray_io_last_error().decode(errors="replace"))


def photon_dump(path, photons, energy):
    """PhotonMap::dump (src/pmap.cc:180-205): photons (n, 12) float32 rows
    pos, dir, norm, col in the reference's text format"""
    arr = np.ascontiguousarray(photons, np.float32).reshape(-1, 12)
    _io_check(lib().sray_photon_dump(os.fsencode(path), arr.ctypes.data_as(C.POINTER(abi.Photon)), len(arr),
                                     float(energy)))


def photon_restore(path, energy):
    """PhotonMap::restore (src/pmap.cc:207-251): (n, 12) float32 photons,
    colours scaled by energy / (the dump's energy)"""
    p, n = C.POINTER(abi.Photon)(), C.c_uint32()
    _io_check(lib().sray_photon_restore(os.fsencode(path), float(energy), C.byref(p), C.byref(n)))
    try:
        if n.value == 0:
            return np.zeros((0, 12), np.float32)
        return np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_float)), shape=(n.value, 12)).copy()
    finally:
        lib().sray_photon_free(p)


def write_pfm(path, rgba):
    """the float framebuffer (h, w, 4), top row first, as a PFM file"""
    img = np.ascontiguousarray(rgba, np.float32)
    h, w = img.shape[:2]
    _io_check(lib().sray_image_write_pfm(os.fsencode(path), img.ctypes.data_as(C.POINTER(C.c_float)), w, h))


def write_png(path, rgba):
    """the float framebuffer (h, w, 4), top row first, as an 8-bit RGBA PNG"""
    img = np.ascontiguousarray(rgba, np.float32)
    h, w = img.shape[:2]
    _io_check(lib().sray_image_write_png(os.fsencode(path), img.ctypes.data_as(C.POINTER(C.c_float)), w, h))


def read_image(path):
    """Image::load (src/img.cc:64-78): a PNG, PPM/PGM or PFM file as an
    (h, w, 4) float32 RGBA array, top row first"""
    p, w, h = C.POINTER(C.c_float)(), C.c_int(), C.c_int()
    _io_check(lib().sray_image_read(os.fsencode(path), C.byref(p), C.byref(w), C.byref(h)))
    try:
        return np.ctypeslib.as_array(p, shape=(h.value, w.value, 4)).copy()
    finally:
        lib().sray_image_free(p)


def status_record(st, a1, a2, a3, a4):
    """one record of the backend status protocol (src/render.cc:279-289)"""
    buf = C.create_string_buffer(abi.SRAY_STATUS_RECORD)
    _io_check(lib().sray_status_record(st.encode() if isinstance(st, str) else st, a1, a2, a3, a4, buf))
    return buf.raw


# ---------------------------------------------------------------------------
# device context
# ---------------------------------------------------------------------------
class Context:
    """One sray_gpu context (one HIP stream) on a device."""

    def __init__(self, device=0):
        h = C.c_void_p()
        rc = lib().sray_gpu_create(int(device), 0, C.byref(h))
        if rc:
            raise SrayError(rc, "sray_gpu_create failed (no HIP device?)")
        self._h = h

    def _check(self, rc):
        if rc:
            raise SrayError(rc, lib().sray_gpu_last_error(self._h).decode(errors="replace"))

    def upload(self, scene):
        """scene: HostScene or SceneDesc pointer -> scene id"""
        sid = C.c_uint32()
        ptr = scene.desc_ptr() if isinstance(scene, HostScene) else scene
        self._check(lib().sray_gpu_scene_upload(self._h, ptr, C.byref(sid)))
        return sid.value

    def release(self, sid):
        self._check(lib().sray_gpu_scene_release(self._h, sid))

    def scene_bytes(self, sid):
        b = C.c_uint64()
        self._check(lib().sray_gpu_scene_bytes(self._h, sid, C.byref(b)))
        return b.value

    def trace(self, sid, rays, any_hit=False, counts=False, kernel="plain"):
        """kernel: "plain" (one walk per lane), "wave" or "sparse" (the
        render path's level-kernel instances)"""
        rays = np.ascontiguousarray(rays, RAY_DTYPE)
        hits = np.zeros(len(rays), HIT_DTYPE)
        cnt = TraceCounts()
        flags = abi.SRAY_TRACE_ANY if any_hit else 0
        flags |= {"plain": 0, "wave": abi.SRAY_TRACE_WAVE,
                  "sparse": abi.SRAY_TRACE_WAVE | abi.SRAY_TRACE_SPARSE}[kernel]
        self._check(lib().sray_gpu_trace(self._h, sid, rays.ctypes.data_as(C.POINTER(Ray)), len(rays),
                                         hits.ctypes.data_as(C.POINTER(Hit)), flags,
                                         C.byref(cnt) if counts else None))
        return (hits, cnt) if counts else hits

    def trace_device(self, sid, d_rays, n, d_hits, any_hit=False):
        self._check(lib().sray_gpu_trace_device(self._h, sid, C.c_void_p(d_rays), n, C.c_void_p(d_hits),
                                                abi.SRAY_TRACE_ANY if any_hit else 0))

    def primary_rays_device(self, sid, params, sample, d_rays):
        self._check(lib().sray_gpu_primary_rays_device(self._h, sid, C.byref(params), int(sample),
                                                       C.c_void_p(d_rays)))

    def render(self, sid, params):
        w, h = params.x1 - params.x0, params.y1 - params.y0
        out = np.zeros((h, w, 4), np.float32)
        st = Stats()
        self._check(lib().sray_gpu_render(self._h, sid, C.byref(params),
                                          out.ctypes.data_as(C.POINTER(C.c_float)), C.byref(st)))
        return out, st

    def shoot_photons(self, sid, caust=0, gi=0, energy=300.0, iter=7, t0=0, t1=0, seed=0x5EED):
        """shoot_photons on the device: ((caustic map id, global map id), PhotonStats);
        the maps are bound to the scene for its renders (id 0 = not shot)"""
        pp = photon_params(caust, gi, energy, iter, t0, t1, seed)
        ids = (C.c_uint32(), C.c_uint32())
        st = abi.PhotonStats()
        self._check(lib().sray_gpu_shoot_photons(self._h, sid, C.byref(pp), C.byref(ids[0]), C.byref(ids[1]),
                                                 C.byref(st)))
        return (ids[0].value, ids[1].value), st

    def pmap_upload(self, sid, photons, kind):
        """photons: (n, 12) float32 rows pos, dir, norm, col -> map id (bound to scene sid if sid)"""
        arr = np.ascontiguousarray(photons, np.float32).reshape(-1, 12)
        pid = C.c_uint32()
        self._check(lib().sray_gpu_pmap_upload(self._h, sid, arr.ctypes.data_as(C.POINTER(abi.Photon)), len(arr),
                                               kind, C.byref(pid)))
        return pid.value

    def pmap_download(self, pid):
        """a map's photons in storage order as (n, 12) float32"""
        n = C.c_uint32()
        self._check(lib().sray_gpu_pmap_download(self._h, pid, None, 0, C.byref(n)))
        out = np.zeros((max(n.value, 1), 12), np.float32)
        self._check(lib().sray_gpu_pmap_download(self._h, pid, C.c_void_p(out.ctypes.data), n.value, C.byref(n)))
        return out[:n.value]

    def pmap_release(self, pid):
        self._check(lib().sray_gpu_pmap_release(self._h, pid))

    def render_device(self, sid, params, d_out):
        st = Stats()
        self._check(lib().sray_gpu_render_device(self._h, sid, C.byref(params), C.c_void_p(d_out), C.byref(st)))
        return st

    # ---- serving: continuous batching behind submit / poll
    def submit(self, sid, params, camera=None, out=None):
        """queue a tile request; returns a ticket.  The image goes to a new
        host array (kept until poll returns it), or to the device buffer out
        (a pointer from malloc of (x1-x0)*(y1-y0)*16 bytes; poll then
        returns None in place of the array)"""
        r = abi.Request()
        r.scene_id = sid
        r.params = params
        if camera is not None:
            r.camera = camera
            r.use_camera = 1
        t = C.c_uint64()
        if out is not None:
            r.flags = abi.SRAY_REQUEST_DEVICE_OUT
            self._check(lib().sray_gpu_submit(self._h, C.byref(r), C.cast(C.c_void_p(out), C.POINTER(C.c_float)),
                                              C.byref(t)))
            return t.value
        w, h = params.x1 - params.x0, params.y1 - params.y0
        host = np.empty((h, w, 4), np.float32)
        self._check(lib().sray_gpu_submit(self._h, C.byref(r), host.ctypes.data_as(C.POINTER(C.c_float)),
                                          C.byref(t)))
        if not hasattr(self, "_outs"):
            self._outs = {}
        self._outs[t.value] = host
        return t.value

    def submit_many(self, jobs):
        """queue (scene, params, camera or None, device output or None)
        requests together; returns their tickets in order"""
        n = len(jobs)
        reqs = (abi.Request * n)()
        outs = (C.POINTER(C.c_float) * n)()
        hosts = {}
        for k, (sid, params, camera, out) in enumerate(jobs):
            r = reqs[k]
            r.scene_id = sid
            r.params = params
            if camera is not None:
                r.camera = camera
                r.use_camera = 1
            if out is not None:
                r.flags = abi.SRAY_REQUEST_DEVICE_OUT
                outs[k] = C.cast(C.c_void_p(out), C.POINTER(C.c_float))
            else:
                w, h = params.x1 - params.x0, params.y1 - params.y0
                hosts[k] = np.empty((h, w, 4), np.float32)
                outs[k] = hosts[k].ctypes.data_as(C.POINTER(C.c_float))
        t = (C.c_uint64 * n)()
        self._check(lib().sray_gpu_submit_many(self._h, reqs, n, outs, t))
        if not hasattr(self, "_outs"):
            self._outs = {}
        for k, arr in hosts.items():
            self._outs[t[k]] = arr
        return [t[k] for k in range(n)]

    def poll(self, ticket, wait=True):
        """(image, sray_result) when done; None if not done and wait is False"""
        res = abi.Result()
        rc = lib().sray_gpu_poll(self._h, ticket, 1 if wait else 0, C.byref(res))
        if rc == abi.SRAY_ERR_PENDING and not wait:
            return None
        out = self._outs.pop(ticket, None) if hasattr(self, "_outs") else None
        self._check(rc)
        return out, res

    # ---- tuning
    def tuning(self):
        t = abi.Tuning()
        self._check(lib().sray_gpu_get_tuning(self._h, C.byref(t)))
        return t

    def set_tuning(self, **kw):
        t = self.tuning()
        for k, v in kw.items():
            setattr(t, k, v)
        self._check(lib().sray_gpu_set_tuning(self._h, C.byref(t)))
        return t

    def autotune(self, sid, params, rounds=3):
        """coordinate-descent search over the runtime knobs; keeps and returns (tuning, median ms)"""
        t, ms = abi.Tuning(), C.c_double()
        self._check(lib().sray_gpu_autotune(self._h, sid, C.byref(params), rounds, C.byref(t), C.byref(ms)))
        return t, ms.value

    def autotune_octree(self, desc_ptr, params, candidates, rounds=3):
        """fastest (mesh items, mesh depth) of candidates: (index, median ms)"""
        n = len(candidates)
        items = (C.c_int32 * n)(*[c[0] for c in candidates])
        depths = (C.c_int32 * n)(*[c[1] for c in candidates])
        best, ms = C.c_int(), C.c_double()
        self._check(lib().sray_gpu_autotune_octree(self._h, desc_ptr, C.byref(params), rounds, items, depths, n,
                                                   C.byref(best), C.byref(ms)))
        return best.value, ms.value

    def set_status_fd(self, fd):
        """status records of renders and requests go to fd (-1: off)"""
        self._check(lib().sray_gpu_set_status_fd(self._h, int(fd)))

    def set_admission_cap(self, nbytes):
        self._check(lib().sray_gpu_set_admission_cap(self._h, int(nbytes)))

    # ---- device memory (the library's own runtime; no torch needed)
    def malloc(self, nbytes):
        p = C.c_void_p()
        self._check(lib().sray_gpu_malloc(self._h, int(nbytes), C.byref(p)))
        return p.value

    def free(self, dptr):
        self._check(lib().sray_gpu_free(self._h, C.c_void_p(dptr)))

    def to_device(self, arr):
        arr = np.ascontiguousarray(arr)
        d = self.malloc(arr.nbytes)
        self._check(lib().sray_gpu_memcpy(self._h, C.c_void_p(d), C.c_void_p(arr.ctypes.data), arr.nbytes, 0))
        return d

    def memcpy_device(self, dst, src, nbytes):
        """device-to-device copy on the context's device"""
        self._check(lib().sray_gpu_memcpy(self._h, C.c_void_p(dst), C.c_void_p(src), int(nbytes), 2))

    def to_host(self, dptr, nbytes, dtype):
        out = np.empty(nbytes // np.dtype(dtype).itemsize, dtype)
        self._check(lib().sray_gpu_memcpy(self._h, C.c_void_p(out.ctypes.data), C.c_void_p(dptr), nbytes, 1))
        return out

    def synchronize(self):
        self._check(lib().sray_gpu_synchronize(self._h))

    def set_profiling(self, on=True):
        self._check(lib().sray_gpu_set_profiling(self._h, 1 if on else 0))

    def kernel_stats(self, reset=False):
        """{kernel name: (launches, total_ms)} from in-library HIP events"""
        n = C.c_int()
        self._check(lib().sray_gpu_kernel_stats(self._h, None, 0, C.byref(n), 0))
        arr = (abi.KernelStat * max(n.value, 1))()
        self._check(lib().sray_gpu_kernel_stats(self._h, arr, n.value, C.byref(n), 1 if reset else 0))
        return {arr[i].name.decode(): (arr[i].launches, arr[i].total_ms) for i in range(n.value)}

    def close(self):
        if self._h:
            lib().sray_gpu_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
