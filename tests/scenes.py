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


# ---------------------------------------------------------------- exact ties
# Scenes where distinct items tie at exactly the same distance, built so
# that the reference's first-depth-first-occurrence rule (strict '<' over
# children 0..7 and leaf lists, src/octree.inl:211-234) picks an item whose
# first occurrence lies in a leaf *beyond* the hit.  Every coordinate is a
# small dyadic number, so hit distances and hit points are exact in float:
# the tied items sit in the plane y = 0, their plane normals are (0, ny, 0),
# and every ray is o = p - 0.5 d with p on the tied items, so t = 0.5.
#
# The octree has one split (depth 1, one item per node): extender items at
# the corners make the root box [-8, 8]^3, so the children are the octants
# around the origin (numbered as src/octree.inl:161-169: bit 0 = high x,
# bit 1 = high y, bit 2 = high z, except that children 2/3 and 6/7 run x
# high-then-low: 2 = (+x,+y,-z), 3 = (-x,+y,-z), 6 = (+x,+y,+z),
# 7 = (-x,+y,+z)).
#
# Pair k: item B (lower index) spans x in [0.5, 1.5], item A spans x in
# [-1.5, 0.5]; they share the edge x = 0.5 (mesh level) or are the same
# triangle (scene level, A's box widened by an extra triangle).  A ray
# heading to -x crosses y = 0 at p (x >= 0.5) and then enters the
# low-x low-y octant (child 4 for z > 0, child 0 for z < 0) at t = 0.625:
# that octant holds A only and comes first in depth-first order, so the
# reference returns A; a walk that skips leaves beyond the best hit sees
# B first in every leaf it visits and returns B.  Rays heading to +x
# meet B first in depth-first order (child 5 / child 1) and return B.
TIE_Z = [(0.5, 1.25), (2.0, 2.75), (3.5, 4.25), (5.0, 5.75),
         (-1.25, -0.5), (-2.75, -2.0), (-4.25, -3.5), (-5.75, -5.0)]


def _extenders():
    return [[(-8, -8, -8), (-7.5, -8, -8), (-8, -7.5, -8)], [(8, 8, 8), (7.5, 8, 8), (8, 7.5, 8)]]


def tie_mesh_scene():
    """one mesh: two corner extenders, then for each pair B_k, A_k (B first).
    Returns (built scene, list of (index of B, index of A, z0, z1))."""
    tris = _extenders()
    pairs = []
    for z0, z1 in TIE_Z:
        b = [(0.5, 0, z0), (1.5, 0, z0), (0.5, 0, z1)]
        a = [(0.5, 0, z0), (0.5, 0, z1), (-1.5, 0, z0)]
        pairs.append((len(tris), len(tris) + 1, z0, z1))
        tris += [b, a]
    return build(meshes=[(np.asarray(tris, np.float32), (0, 0, 0), 0)], mesh_oct=(1, 1)), pairs


def tie_object_scene():
    """objects: a pair of coincident spheres (same centre and radius, objects
    0 and 1) away from the pairs, a corner-extender mesh, then per pair mesh
    B_k = [T_k] and mesh A_k = [T_k, X_k] (X_k widens A's box to x = -1.5).
    Returns (built scene, list of (object of B, object of A, z0, z1))."""
    meshes = [(np.asarray(_extenders(), np.float32), (0, 0, 0), 0)]
    pairs = []
    for z0, z1 in TIE_Z:
        t = [(0.5, 0, z0), (1.5, 0, z0), (0.5, 0, z1)]
        x = [(-1.5, 0, z0), (-1.25, 0, z0), (-1.5, 0, z0 + 0.25)]
        pairs.append((len(meshes), len(meshes) + 1, z0, z1))
        meshes += [(np.asarray([t], np.float32), (0, 0, 0), 0), (np.asarray([t, x], np.float32), (0, 0, 0), 0)]
    spheres = [((-4.0, 4.0, 0.0), 1.0, 0), ((-4.0, 4.0, 0.0), 1.0, 0)]
    # build() numbers spheres first, then meshes
    pairs = [(ib + len(spheres), ia + len(spheres), z0, z1) for ib, ia, z0, z1 in pairs]
    return build(spheres=spheres, meshes=meshes, scene_oct=(1, 1), mesh_oct=(1, 14)), pairs


def tie_rays(pairs, per_pair=32, seed=0, edge=True):
    """rays through points p of each pair's tied region at t = 0.5:
    p on the shared edge x = 0.5 (edge=True) or inside the triangle at
    x = 0.75; half head to -x (the reference picks A), half to +x (B).
    Returns (rays, expected item index, mask of -x rays)."""
    rng = np.random.default_rng(seed)
    o, d, want, neg = [], [], [], []
    px = 0.5 if edge else 0.75
    for ib, ia, z0, z1 in pairs:
        lo, hi = (z0 + 0.125, z1 - 0.25) if not edge else (z0, z1)
        for j in range(per_pair):
            # dyadic z in [lo, hi] with 1/64 steps
            pz = lo + np.floor(rng.uniform(0, 1) * (hi - lo) * 64) / 64
            for sx in (-1.0, 1.0):
                dz = 0.25 if rng.uniform() < 0.5 else -0.25
                dd = np.array([4.0 * sx, -2.0, dz])
                p = np.array([px, 0.0, pz])
                o.append(p - 0.5 * dd)
                d.append(dd)
                want.append(ia if sx < 0 else ib)
                neg.append(sx < 0)
    return rays(o, d), np.asarray(want), np.asarray(neg)


def sparse_lanes(r, every=16):
    """spreads rays one per `every` lanes, the rest missing everything (so
    that a wave has few lanes in mesh leaves); returns (rays, index mask)"""
    n = len(r)
    out = np.zeros(n * every, r.dtype)
    out["origin"] = np.array([0.0, 50.0, 0.0], np.float32)
    out["dir"] = np.array([0.0, 1.0, 0.0], np.float32)
    out[::every] = r
    m = np.zeros(n * every, bool)
    m[::every] = True
    return out, m


def look_at(pos, target, vfov_deg, shutter=0):
    """camera -> world matrix of a camera at pos looking at target (+z
    forward, +y up), as an sray_camera_desc"""
    f = np.asarray(target, np.float64) - np.asarray(pos, np.float64)
    f /= np.linalg.norm(f)
    r = np.cross([0.0, 1.0, 0.0], f)
    r /= np.linalg.norm(r)
    u = np.cross(f, r)
    cam = abi.CameraDesc()
    cam.xform[:] = np.array([[r[0], u[0], f[0], pos[0]], [r[1], u[1], f[1], pos[1]],
                             [r[2], u[2], f[2], pos[2]]], np.float32).ravel()
    cam.vfov = math.radians(vfov_deg)
    cam.max_dist = 10000.0
    cam.shutter = shutter
    return cam


def desc_copy(desc_ptr, camera=None, shutter=None):
    """a copy of a scene description (same arrays) with another camera or
    shutter"""
    d = abi.SceneDesc()
    C.memmove(C.byref(d), desc_ptr, C.sizeof(d))
    if camera is not None:
        d.camera = camera
    if shutter is not None:
        d.camera.shutter = shutter
    return d


# ---------------------------------------------------------------- keyframes
def _mat(name, kd, ks=0.5, shin=30.0, refl=0.0, refr=0.0, ior=1.5):
    return (f'  <material name="{name}" shader="phong">\n'
            f'    <mattr name="diffuse" value="{kd[0]} {kd[1]} {kd[2]}"/>\n'
            f'    <mattr name="specular" value="{ks} {ks} {ks}"/>\n'
            f'    <mattr name="shininess" value="{shin}"/>\n'
            f'    <mattr name="reflect" value="{refl}"/>\n'
            f'    <mattr name="refract" value="{refr}"/>\n'
            f'    <mattr name="ior" value="{ior}"/>\n  </material>\n')


def _pyramid():
    v = [(-1.0, 0.0, -1.0), (1.0, 0.0, -1.0), (1.0, 0.0, 1.0), (-1.0, 0.0, 1.0), (0.0, 1.5, 0.0)]
    faces = [(0, 1, 4), (1, 2, 4), (2, 3, 4), (3, 0, 4), (0, 2, 1), (0, 3, 2)]
    s = "    <mesh>\n"
    for i, p in enumerate(v):
        s += f'      <vertex id="{i}" val="{p[0]} {p[1]} {p[2]}"/>\n'
    for f in faces:
        s += "      <face>" + "".join(f'<vref vertex="{k}"/>' for k in f) + "</face>\n"
    return s + "    </mesh>\n"


def animated_scene_xml(shutter=40):
    """a scene whose transforms are keyframed: a pyramid turning 90 degrees
    about y while it scales, a sphere scaled into an ellipsoid with a
    repeating position track, a cylinder turning about a pivot (quaternion
    keys), a step-interpolated sphere, a cubic position track, a moving
    point light and a target camera whose position and target move"""
    x = "<scene>\n"
    x += '  <env ambient="0.05 0.05 0.05" bgcolor="0.1 0.1 0.2" ior="1.0"/>\n'
    x += _mat("red", (0.8, 0.2, 0.2)) + _mat("green", (0.2, 0.8, 0.3), refl=0.3)
    x += _mat("blue", (0.2, 0.3, 0.9), ks=1.0, shin=60.0) + _mat("glass", (0.0, 0.0, 0.0), ks=1.0, shin=80.0,
                                                                 refl=0.1, refr=0.9)
    x += _mat("grey", (0.7, 0.7, 0.7), ks=0.0)
    # a sphere's normal is transformed back without normalising
    # (src/sphere.cc:127-129): under a non-uniform scale a specular term
    # raises a reflected vector longer than 1 to the shininess, so the
    # ellipsoid below is diffuse only
    x += _mat("matte", (0.3, 0.4, 0.9), ks=0.0)
    # floor
    x += ('  <object type="mesh">\n    <matref name="grey"/>\n    <mesh>\n'
          '      <vertex id="0" val="-8.0 -1.0 -8.0"/><vertex id="1" val="8.0 -1.0 -8.0"/>\n'
          '      <vertex id="2" val="8.0 -1.0 8.0"/><vertex id="3" val="-8.0 -1.0 8.0"/>\n'
          '      <face><vref vertex="0"/><vref vertex="2"/><vref vertex="1"/></face>\n'
          '      <face><vref vertex="0"/><vref vertex="3"/><vref vertex="2"/></face>\n'
          '    </mesh>\n  </object>\n')
    x += ('  <object type="mesh">\n    <matref name="red"/>\n'
          '    <xform time="0" pos="-2.5 -1.0 1.0" axisrot="0.0 0.0 1.0 0.0" scale="1.0 1.0 1.0"/>\n'
          '    <xform time="40" axisrot="90.0 0.0 1.0 0.0" scale="1.5 0.6 1.0"/>\n' + _pyramid() + '  </object>\n')
    x += ('  <object type="sphere">\n    <matref name="matte"/>\n'
          '    <xform extrapolate="repeat" time="0" pos="1.5 0.2 0.0" scale="1.0 1.0 1.0"/>\n'
          '    <xform time="30" pos="1.5 1.2 0.5" scale="1.6 0.7 1.0"/>\n'
          '    <sphere rad="0.9"/>\n  </object>\n')
    x += ('  <object type="cylinder">\n    <matref name="green"/>\n'
          '    <xform pivot="0.5 0.0 0.0" time="0" pos="0.0 -1.0 3.0" rot="1.0 0.0 0.0 0.0"/>\n'
          '    <xform time="40" rot="0.9238795 0.0 0.0 0.3826834"/>\n'
          '    <cylinder rad="0.5" start="0.0" end="2.0"/>\n  </object>\n')
    x += ('  <object type="sphere">\n    <matref name="glass"/>\n'
          '    <xform interpolate="step" time="0" pos="-0.5 0.0 -1.5"/>\n'
          '    <xform time="20" pos="0.0 0.3 -1.5"/>\n'
          '    <sphere rad="0.6"/>\n  </object>\n')
    x += ('  <object type="sphere">\n    <matref name="red"/>\n'
          '    <xform interpolate="cubic" time="0" pos="3.0 2.0 2.0"/>\n    <xform time="10" pos="3.5 2.5 2.0"/>\n'
          '    <xform time="25" pos="3.0 3.0 2.5"/>\n    <xform time="40" pos="2.5 2.0 2.0"/>\n'
          '    <sphere rad="0.4"/>\n  </object>\n')
    x += ('  <light type="point" color="1.0 1.0 1.0">\n'
          '    <xform time="0" pos="4.0 8.0 -6.0"/>\n    <xform time="40" pos="-4.0 9.0 -5.0"/>\n  </light>\n')
    x += (f'  <camera type="target" fov="50.0" shutter="{shutter}">\n'
          '    <xform time="0" pos="0.0 2.5 -9.0"/>\n    <xform time="40" pos="1.0 3.0 -8.5"/>\n'
          '    <target><xform time="0" pos="0.0 0.0 0.0"/><xform time="40" pos="0.3 0.2 0.5"/></target>\n'
          '  </camera>\n')
    return x + "</scene>\n"


def rows_of(desc, what, index=0):
    """the per-msec rows of an object (24 floats), light (3) or the camera
    (12) as an array, or None"""
    n = desc.frame_t1 - desc.frame_t0 + 1
    if what == "object":
        p, w = desc.objects[index].time_xforms, 24
    elif what == "light":
        p, w = desc.lights[index].time_pos, 3
    else:
        p, w = desc.camera.time_xforms, 12
    if not p:
        return None
    return np.ctypeslib.as_array(p, shape=(n * w,)).reshape(n, w).copy()


def _normal_map(n, amp=0.35, freq=3):
    """an 8-bit RGB normal map of bumps h = amp sin(2 pi f u) sin(2 pi f v):
    n = normalise(-dh/du, -dh/dv, 1) encoded as (n + 1) / 2"""
    u = (np.arange(n) + 0.5) / n
    U, V = np.meshgrid(u, u)
    k = 2 * np.pi * freq
    du = amp * k * np.cos(k * U) * np.sin(k * V)
    dv = amp * k * np.sin(k * U) * np.cos(k * V)
    nv = np.stack([-du, -dv, np.ones_like(du)], -1)
    nv /= np.linalg.norm(nv, axis=-1, keepdims=True)
    return np.clip(np.round((nv + 1) * 127.5), 0, 255).astype(np.int64)


def write_textures(d, seed=7):
    """the texture files of textured_scene_xml, written into directory d:
    checker.png (RGB, random texel colours, filters 0-4), bumps.png (a
    normal map), spec.ppm (P3 text), mask.pfm (grey, little-endian),
    shin.pgm (P5, 16-bit), alpha.png (palette + tRNS)"""
    import imgfiles as I
    rng = np.random.default_rng(seed)
    I.write(f"{d}/checker.png", I.png_bytes(rng.integers(40, 256, (16, 16, 3)), 2, 8, filters=(0, 1, 2, 3, 4)))
    I.write(f"{d}/bumps.png", I.png_bytes(_normal_map(32), 2, 8, filters=(4, 1)))
    I.write(f"{d}/spec.ppm", I.pnm_bytes(rng.integers(0, 256, (8, 8, 3)), 255, "3"))
    I.write(f"{d}/mask.pfm", I.pfm_bytes((rng.random((4, 8)) > 0.4).astype(np.float32) * 0.9 + 0.05))
    I.write(f"{d}/shin.pgm", I.pnm_bytes(rng.integers(20000, 65536, (8, 8)), 65535, "5"))
    pal = rng.integers(0, 256, (4, 3))
    I.write(f"{d}/alpha.png", I.png_bytes(rng.integers(0, 4, (8, 8)), 3, 2, palette=pal, trns=[255, 128]))


def textured_scene_xml(bump_scale=1.5):
    """a scene whose materials are mapped (the file names of
    write_textures): a textured floor with tangents and texture
    coordinates running to 3 (repeat wrap), a bump-mapped sphere, a
    bump-mapped pyramid without tangents, a sphere whose reflectance is
    masked, a glass sphere with a mapped index of refraction, a bump-mapped
    cylinder that also reflects, and a back-facing wall with a "normal"
    attribute but no map (the unflipped normal)"""
    def mat(name, attrs):
        s = f'  <material name="{name}" shader="phong">\n'
        for a in attrs:
            s += "    <mattr " + " ".join(f'{k}="{v}"' for k, v in a.items()) + "/>\n"
        return s + "  </material>\n"
    x = "<scene>\n"
    x += '  <env ambient="0.08 0.08 0.08" bgcolor="0.1 0.1 0.2" ior="1.0"/>\n'
    x += mat("floor", [dict(name="diffuse", value="1.0 0.9 0.8", map="checker.png"),
                       dict(name="specular", value="0.6 0.6 0.6", map="spec.ppm"),
                       dict(name="shininess", value="60.0", map="shin.pgm")])
    x += mat("bumpy", [dict(name="diffuse", value="0.7 0.5 0.3"), dict(name="specular", value="0.7 0.7 0.7"),
                       dict(name="shininess", value="40.0"), dict(name="normal", value=str(bump_scale), map="bumps.png")])
    x += mat("mirror", [dict(name="diffuse", value="0.3 0.3 0.35"), dict(name="specular", value="0.9 0.9 0.9"),
                        dict(name="shininess", value="80.0"), dict(name="reflect", value="0.9", map="mask.pfm")])
    x += mat("glass", [dict(name="diffuse", value="0.0 0.0 0.0"), dict(name="specular", value="1.0 1.0 1.0"),
                       dict(name="shininess", value="80.0"), dict(name="reflect", value="0.1"),
                       dict(name="refract", value="0.9"), dict(name="ior", value="1.6", map="shin.pgm")])
    x += mat("bumpmirror", [dict(name="diffuse", value="0.4 0.6 0.4", map="alpha.png"),
                            dict(name="specular", value="0.5 0.5 0.5"), dict(name="reflect", value="0.4"),
                            dict(name="normal", value="0.8", map="bumps.png")])
    x += mat("flipwall", [dict(name="diffuse", value="0.6 0.6 0.9"), dict(name="specular", value="0.4 0.4 0.4"),
                          dict(name="normal", value="1.0")])
    x += mat("missing", [dict(name="diffuse", value="0.9 0.3 0.3", map="no_such_file.png")])
    # floor: texture coordinates 0..3 over the quad, tangent +x
    x += ('  <object type="mesh">\n    <matref name="floor"/>\n    <mesh>\n'
          '      <vertex id="0" val="-8.0 -1.0 -8.0"/><vertex id="1" val="8.0 -1.0 -8.0"/>\n'
          '      <vertex id="2" val="8.0 -1.0 8.0"/><vertex id="3" val="-8.0 -1.0 8.0"/>\n'
          '      <normal id="0" val="0.0 1.0 0.0"/>\n      <tangent id="0" val="1.0 0.0 0.0"/>\n'
          '      <texcoord id="0" val="0.0 0.0 0.0"/><texcoord id="1" val="3.0 0.0 0.0"/>\n'
          '      <texcoord id="2" val="3.0 3.0 0.0"/><texcoord id="3" val="0.0 3.0 0.0"/>\n'
          '      <face><vref vertex="0" normal="0" tangent="0" texcoord="0"/>'
          '<vref vertex="2" normal="0" tangent="0" texcoord="2"/>'
          '<vref vertex="1" normal="0" tangent="0" texcoord="1"/></face>\n'
          '      <face><vref vertex="0" normal="0" tangent="0" texcoord="0"/>'
          '<vref vertex="3" normal="0" tangent="0" texcoord="3"/>'
          '<vref vertex="2" normal="0" tangent="0" texcoord="2"/></face>\n'
          '    </mesh>\n  </object>\n')
    x += ('  <object type="sphere">\n    <matref name="bumpy"/>\n    <xform pos="-1.8 0.2 0.5"/>\n'
          '    <sphere rad="1.1"/>\n  </object>\n')
    x += ('  <object type="mesh">\n    <matref name="bumpy"/>\n    <xform pos="-4.0 -1.0 2.5"/>\n'
          + _pyramid() + '  </object>\n')
    x += ('  <object type="sphere">\n    <matref name="mirror"/>\n    <xform pos="1.2 0.0 1.5"/>\n'
          '    <sphere rad="1.0"/>\n  </object>\n')
    x += ('  <object type="sphere">\n    <matref name="glass"/>\n    <xform pos="0.2 -0.3 -1.6"/>\n'
          '    <sphere rad="0.7"/>\n  </object>\n')
    x += ('  <object type="cylinder">\n    <matref name="bumpmirror"/>\n    <xform pos="3.3 -1.0 0.5"/>\n'
          '    <cylinder rad="0.6" start="0.0" end="2.2"/>\n  </object>\n')
    x += ('  <object type="sphere">\n    <matref name="missing"/>\n    <xform pos="-0.6 -0.6 -3.2"/>\n'
          '    <sphere rad="0.4"/>\n  </object>\n')
    # a wall whose geometric normal faces away from the camera
    x += ('  <object type="mesh">\n    <matref name="flipwall"/>\n    <mesh>\n'
          '      <vertex id="0" val="-8.0 -1.0 5.0"/><vertex id="1" val="8.0 -1.0 5.0"/>\n'
          '      <vertex id="2" val="8.0 5.0 5.0"/><vertex id="3" val="-8.0 5.0 5.0"/>\n'
          '      <face><vref vertex="0"/><vref vertex="1"/><vref vertex="2"/></face>\n'
          '      <face><vref vertex="0"/><vref vertex="2"/><vref vertex="3"/></face>\n'
          '    </mesh>\n  </object>\n')
    x += ('  <light type="point" color="1.0 1.0 1.0">\n    <xform pos="3.0 6.0 -5.0"/>\n  </light>\n'
          '  <light type="point" color="0.4 0.4 0.5">\n    <xform pos="-5.0 3.0 -2.0"/>\n  </light>\n')
    x += ('  <camera type="target" fov="55.0">\n    <xform pos="0.5 2.5 -8.0"/>\n'
          '    <target><xform pos="0.0 0.0 0.5"/></target>\n  </camera>\n')
    return x + "</scene>\n"


def textured_scene(d, bump_scale=1.5):
    """writes the textures and textured_scene_xml into directory d and
    loads the scene from there (maps are found next to the scene file)"""
    write_textures(d)
    path = f"{d}/scene.xml"
    with open(path, "w") as f:
        f.write(textured_scene_xml(bump_scale))
    return S.HostScene.load_xml(path=path)
