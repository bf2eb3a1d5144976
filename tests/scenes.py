"""Small hand-built scenes for known-answer tests (ctypes SceneDesc with
owned buffers)."""
import ctypes as C
import math

import numpy as np

import sray_amd as S
from sray_amd import abi


def translate(x, y, z):
    m = np.eye(4)[:3].copy()
    m[:, 3] = (x, y, z)
    return m


def inverse(m):
    M = np.vstack([m, [0, 0, 0, 1]])
    return np.linalg.inv(M)[:3]


class Built:
    """keeps the ctypes arrays alive"""

    def __init__(self):
        self.keep = []
        self.desc = abi.SceneDesc()

    def ptr(self):
        return C.pointer(self.desc)


def phong(kd=(1, 1, 1), ks=0.0, shin=1.0, refl=0.0, refr=0.0, ior=1.5):
    m = abi.MaterialDesc()
    m.diffuse[:] = (*kd, 1.0)
    m.specular[:] = (ks, ks, ks, 1.0)
    m.shininess, m.reflect, m.refract, m.ior = shin, refl, refr, ior
    m.flags = abi.SRAY_MAT_HAS_SPECULAR | abi.SRAY_MAT_HAS_REFLECT | abi.SRAY_MAT_HAS_REFRACT | abi.SRAY_MAT_PHONG
    return m


def plane(y, half, up=True):
    """two triangles covering [-half, half]^2 at height y (geometric normal +y if up)"""
    a, b, c, d = (-half, y, -half), (-half, y, half), (half, y, half), (half, y, -half)
    t = [[a, b, c], [a, c, d]] if up else [[a, c, b], [a, d, c]]
    return np.asarray(t, np.float32)


def material(kd=(0, 0, 0), ks=0.0, refl=0.0, refr=0.0, ior=1.5, flags=None):
    """phong material; flags=None sets every attribute (no default-1 fallbacks)"""
    m = phong(kd, ks, 1.0, refl, refr, ior)
    if flags is not None:
        m.flags = flags
    return m


def build(spheres=(), cylinders=(), meshes=(), materials=None, lights=(), camera=None, ambient=(0, 0, 0, 0),
          bg=(0, 0, 0, 0), scene_oct=(5, 5), mesh_oct=(7, 14), normals=None, texcoords=None):
    """spheres: (center, radius, mat); cylinders: (center, radius, start, end, mat);
    meshes: (triangles [N,3,3] local, translation, mat);
    lights: (pos, color[, type, radius, dim])"""
    b = Built()
    objs, mesh_descs, tris = [], [], []
    for c, r, mat in spheres:
        o = abi.ObjectDesc()
        o.type, o.material, o.mesh, o.radius, o.end = abi.SRAY_OBJ_SPHERE, mat, -1, r, 1.0
        m = translate(*c)
        o.xform[:] = m.astype(np.float32).ravel()
        o.inv_xform[:] = inverse(m).astype(np.float32).ravel()
        objs.append(o)
    for c, r, y0, y1, mat in cylinders:
        o = abi.ObjectDesc()
        o.type, o.material, o.mesh, o.radius, o.start, o.end = abi.SRAY_OBJ_CYLINDER, mat, -1, r, y0, y1
        m = translate(*c)
        o.xform[:] = m.astype(np.float32).ravel()
        o.inv_xform[:] = inverse(m).astype(np.float32).ravel()
        objs.append(o)
    ntri = 0
    for t, c, mat in meshes:
        t = np.asarray(t, np.float32).reshape(-1, 3, 3)
        md = abi.MeshDesc()
        md.tri_start, md.tri_count = ntri, len(t)
        mesh_descs.append(md)
        tris.append(t)
        ntri += len(t)
        o = abi.ObjectDesc()
        o.type, o.material, o.mesh, o.radius, o.end = abi.SRAY_OBJ_MESH, mat, len(mesh_descs) - 1, 1.0, 1.0
        m = translate(*c)
        o.xform[:] = m.astype(np.float32).ravel()
        o.inv_xform[:] = inverse(m).astype(np.float32).ravel()
        objs.append(o)
    d = b.desc
    OA = abi.ObjectDesc * max(len(objs), 1)
    oa = OA(*objs)
    b.keep.append(oa)
    d.num_objects, d.objects = len(objs), C.cast(oa, C.POINTER(abi.ObjectDesc))
    MA = abi.MeshDesc * max(len(mesh_descs), 1)
    ma = MA(*mesh_descs)
    b.keep.append(ma)
    d.num_meshes, d.meshes = len(mesh_descs), C.cast(ma, C.POINTER(abi.MeshDesc))
    tp = np.ascontiguousarray(np.concatenate(tris).ravel() if tris else np.zeros(9), np.float32)
    b.keep.append(tp)
    d.num_triangles, d.tri_pos = ntri, tp.ctypes.data_as(C.POINTER(C.c_float))
    if normals is not None:  # per corner, all triangles: [ntri, 3, 3]
        tn = np.ascontiguousarray(np.asarray(normals, np.float32).ravel())
        b.keep.append(tn)
        d.tri_normal = tn.ctypes.data_as(C.POINTER(C.c_float))
    if texcoords is not None:  # [ntri, 3, 2]
        tt = np.ascontiguousarray(np.asarray(texcoords, np.float32).ravel())
        b.keep.append(tt)
        d.tri_texcoord = tt.ctypes.data_as(C.POINTER(C.c_float))
    mats = materials if materials is not None else [phong()]
    MT = abi.MaterialDesc * max(len(mats), 1)
    mt = MT(*mats)
    b.keep.append(mt)
    d.num_materials, d.materials = len(mats), C.cast(mt, C.POINTER(abi.MaterialDesc))
    lts = []
    for light in lights:
        pos, col = light[0], light[1]
        ltype = light[2] if len(light) > 2 else abi.SRAY_LIGHT_POINT
        lt = abi.LightDesc()
        lt.type = ltype
        lt.pos[:] = pos
        lt.color[:] = (*col, 1.0)
        lt.radius = light[3] if len(light) > 3 else 1.0
        lt.dim[:] = light[4] if len(light) > 4 else (1, 1, 1)
        lt.xform[:] = translate(*pos).astype(np.float32).ravel()
        lts.append(lt)
    LA = abi.LightDesc * max(len(lts), 1)
    la = LA(*lts)
    b.keep.append(la)
    d.num_lights, d.lights = len(lts), C.cast(la, C.POINTER(abi.LightDesc))
    cam = abi.CameraDesc()
    cm = translate(*(camera if camera is not None else (0, 0, 0)))
    cam.xform[:] = cm.astype(np.float32).ravel()
    cam.vfov, cam.max_dist, cam.shutter = math.pi / 4, 10000.0, 0
    d.camera = cam
    d.env_ambient[:] = ambient
    d.env_color[:] = bg
    d.env_ior = 1.0
    d.scene_oct_depth, d.scene_oct_items = scene_oct
    d.mesh_oct_depth, d.mesh_oct_items = mesh_oct
    return b


def ray(o, d):
    return S.make_rays(np.asarray([o], np.float32), np.asarray([d], np.float32))


def rays(os_, ds):
    return S.make_rays(np.asarray(os_, np.float32), np.asarray(ds, np.float32))
