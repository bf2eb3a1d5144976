"""CPU restatement of the S-ray 0.1 hot path — TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this package, and only as the checker (or as the timed CPU baseline).
The product (sray_amd/, libsray_gpu.so) never imports, links or calls it.

Which reference function each oracle function restates is cited in
oracle/oracle.cc.  The reference itself cannot be compiled here (it needs
libvmath, kdtree and libimago, none of which is present), so there is no
oracle/_ref build; the restatement is pinned by closed forms, brute force
and invariants in tests/ (see the parity notes in the test suite).
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
BUILD = os.path.join(_HERE, "_build")


def build():
    subprocess.run(["make", "-C", _HERE, "-s"], check=True)


class _Ext(C.Structure):
    _fields_ = [("tie", C.c_int32), ("pad_", C.c_int32), ("nodes", C.c_uint64), ("prims", C.c_uint64)]


class _TreeInfo(C.Structure):
    _fields_ = [("num_nodes", C.c_uint32), ("num_leaves", C.c_uint32), ("num_item_refs", C.c_uint32),
                ("height", C.c_uint32), ("root_min", C.c_double * 3), ("root_max", C.c_double * 3)]


EXT_DTYPE = np.dtype([("tie", np.int32), ("pad", np.int32), ("nodes", np.uint64), ("prims", np.uint64)])

_libs = {}


def lib(fast=False):
    name = "liboracle_fast.so" if fast else "liboracle.so"
    if name not in _libs:
        path = os.path.join(BUILD, name)
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        vp = C.c_void_p
        L.oracle_scene_create.restype = vp
        L.oracle_scene_create.argtypes = [vp, C.c_int, C.c_uint32]
        L.oracle_scene_destroy.argtypes = [vp]
        L.oracle_trace.argtypes = [vp, vp, C.c_uint32, vp, vp, C.c_int, C.c_int]
        L.oracle_tree.argtypes = [vp, C.c_int, C.POINTER(_TreeInfo), vp, vp, C.c_uint32, C.c_uint32]
        L.oracle_primary_rays.argtypes = [vp, vp, C.c_int, vp]
        L.oracle_render.argtypes = [vp, vp, vp, C.c_int, vp]
        L.oracle_set_photons.argtypes = [vp, C.c_int, vp, C.c_uint32]
        L.oracle_gather.argtypes = [vp, C.c_int, vp, C.c_uint32, C.c_float, C.c_int, C.c_int, vp]
        L.oracle_scene_diag.restype = C.c_double
        L.oracle_scene_diag.argtypes = [vp]
        L.oracle_rng_u32.restype = C.c_uint32
        L.oracle_rng_u32.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32]
        L.oracle_splitmix64.restype = C.c_uint64
        L.oracle_splitmix64.argtypes = [C.c_uint64]
        L.oracle_shoot_photons.argtypes = [vp, vp, vp, C.c_uint32, C.POINTER(C.c_uint32), vp, C.c_uint32,
                                           C.POINTER(C.c_uint32), C.POINTER(C.c_uint64)]
        L.oracle_projmap.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint64, vp,
                                     C.POINTER(C.c_double)]
        L.oracle_surface.argtypes = [vp, vp, C.c_uint32, vp]
        L.oracle_shade_normal.argtypes = [vp, vp, C.c_uint32, vp]
        L.oracle_tex_lookup.argtypes = [vp, C.c_int, vp, C.c_uint32, vp]
        L.oracle_light_points.argtypes = [vp, C.c_int, C.c_uint64, C.c_uint32, C.c_uint32, vp]
        L.oracle_cell_from_dir.argtypes = [C.c_float, C.c_float, C.c_float, C.POINTER(C.c_int),
                                           C.POINTER(C.c_int)]
        _libs[name] = L
    return _libs[name]


Q9_RETRAVERSE = 0x1
BRUTE_FORCE = 0x2
BRUTE_GATHER = 0x4


class OracleScene:
    """The oracle's scene built from an sray_scene_desc (prec 32 or 64)."""

    def __init__(self, desc_ptr, prec=32, flags=0, fast=False):
        self.L = lib(fast)
        self.h = self.L.oracle_scene_create(C.cast(desc_ptr, C.c_void_p), prec, flags)
        if not self.h:
            raise RuntimeError("oracle_scene_create failed")
        self.prec = prec

    def trace(self, rays, any_hit=False, nthreads=8):
        from sray_amd import HIT_DTYPE
        rays = np.ascontiguousarray(rays)
        hits = np.zeros(len(rays), HIT_DTYPE)
        ext = np.zeros(len(rays), EXT_DTYPE)
        rc = self.L.oracle_trace(self.h, rays.ctypes.data, len(rays), hits.ctypes.data, ext.ctypes.data,
                                 1 if any_hit else 0, nthreads)
        if rc:
            raise RuntimeError("oracle_trace failed")
        return hits, ext

    def tree(self, level):
        info = _TreeInfo()
        self.L.oracle_tree(self.h, level, C.byref(info), None, None, 0, 0)
        items = np.zeros(max(info.num_item_refs, 1), np.uint32)
        counts = np.zeros(max(info.num_leaves, 1), np.uint32)
        self.L.oracle_tree(self.h, level, C.byref(info), items.ctypes.data, counts.ctypes.data,
                           info.num_item_refs, info.num_leaves)
        return info, items[:info.num_item_refs], counts[:info.num_leaves]

    def primary_rays(self, params, sample=0):
        from sray_amd import RAY_DTYPE
        n = (params.x1 - params.x0) * (params.y1 - params.y0)
        rays = np.zeros(n, RAY_DTYPE)
        if self.L.oracle_primary_rays(self.h, C.addressof(params), sample, rays.ctypes.data):
            raise RuntimeError("oracle_primary_rays failed")
        return rays

    def render(self, params, nthreads=8):
        from sray_amd.abi import Stats
        w, h = params.x1 - params.x0, params.y1 - params.y0
        out = np.zeros((h, w, 4), np.float32)
        st = Stats()
        if self.L.oracle_render(self.h, C.addressof(params), out.ctypes.data, nthreads, C.addressof(st)):
            raise RuntimeError("oracle_render failed")
        return out, st

    def set_photons(self, kind, photons):
        photons = np.ascontiguousarray(photons, np.float32).reshape(-1, 12)
        self.L.oracle_set_photons(self.h, kind, photons.ctypes.data, len(photons))

    def gather(self, kind, queries, radius, k=0, mode=0):
        q = np.ascontiguousarray(queries, np.float32).reshape(-1, 10)
        out = np.zeros((len(q), 3), np.float32)
        self.L.oracle_gather(self.h, kind, q.ctypes.data, len(q), radius, k, mode, out.ctypes.data)
        return out

    def shoot_photons(self, params):
        """shoot_photons: (caustic photons, global photons, segments traced);
        photons as (n, 12) float32 rows pos, dir, norm, col.  The maps also
        become this scene's photon maps for render/gather."""
        nc, ng, seg = C.c_uint32(), C.c_uint32(), C.c_uint64()
        pp = C.addressof(params)
        if self.L.oracle_shoot_photons(self.h, pp, None, 0, C.byref(nc), None, 0, C.byref(ng), C.byref(seg)):
            raise RuntimeError("oracle_shoot_photons failed")
        c = np.zeros((max(nc.value, 1), 12), np.float32)
        g = np.zeros((max(ng.value, 1), 12), np.float32)
        if self.L.oracle_shoot_photons(self.h, pp, c.ctypes.data, nc.value, C.byref(nc), g.ctypes.data, ng.value,
                                       C.byref(ng), C.byref(seg)):
            raise RuntimeError("oracle_shoot_photons failed")
        return c[:nc.value], g[:ng.value], seg.value

    def projmap(self, light, kind, t0=0, t1=0, seed=0):
        cells = np.zeros(64 * 32, np.uint8)
        cov = C.c_double()
        if self.L.oracle_projmap(self.h, light, kind, t0, t1, seed, cells.ctypes.data, C.byref(cov)):
            raise RuntimeError("oracle_projmap failed")
        return cells.reshape(32, 64), cov.value

    def surface(self, rays):
        """(n, 8) float32: hit position, normal, u, v of each closest hit (NaN on a miss)"""
        rays = np.ascontiguousarray(rays)
        out = np.zeros((len(rays), 8), np.float32)
        if self.L.oracle_surface(self.h, rays.ctypes.data, len(rays), out.ctypes.data):
            raise RuntimeError("oracle_surface failed")
        return out

    def shade_normal(self, rays):
        """(n, 6) float32: the shading normal at each closest hit (facing
        the ray, or the material's "normal" attribute's) and the tangent"""
        rays = np.ascontiguousarray(rays)
        out = np.zeros((len(rays), 6), np.float32)
        if self.L.oracle_shade_normal(self.h, rays.ctypes.data, len(rays), out.ctypes.data):
            raise RuntimeError("oracle_shade_normal failed")
        return out

    def tex_lookup(self, tex, uv):
        """(n, 4) float32: texture tex of the scene looked up at (n, 2) (u, v)"""
        uv = np.ascontiguousarray(uv, np.float32).reshape(-1, 2)
        out = np.zeros((len(uv), 4), np.float32)
        if self.L.oracle_tex_lookup(self.h, tex, uv.ctypes.data, len(uv), out.ctypes.data):
            raise RuntimeError("oracle_tex_lookup failed")
        return out

    def diag(self):
        return self.L.oracle_scene_diag(self.h)

    def light_points(self, light, n, seed=1, pixel=0):
        """(n, 3) float64: sample points of one light, samples 0..n-1 of a pixel"""
        out = np.zeros((n, 3), np.float64)
        if self.L.oracle_light_points(self.h, light, seed, pixel, n, out.ctypes.data):
            raise RuntimeError("oracle_light_points failed")
        return out

    def close(self):
        if self.h:
            self.L.oracle_scene_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def rng_u32(seed, a, b, c, d):
    return lib().oracle_rng_u32(seed, a, b, c, d)


def splitmix64(x):
    return lib().oracle_splitmix64(x)


def cell_from_dir(x, y, z):
    u, v = C.c_int(), C.c_int()
    lib().oracle_cell_from_dir(x, y, z, C.byref(u), C.byref(v))
    return u.value, v.value
