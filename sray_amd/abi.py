"""ctypes mirror of include/sray_gpu.h (layout checked against the library's
sray_abi_sizes() by the test suite)."""
import ctypes as C

SRAY_OK = 0
SRAY_ERR_ARG = -1
SRAY_ERR_HIP = -2
SRAY_ERR_NOMEM = -3
SRAY_ERR_NOTFOUND = -4
SRAY_ERR_LIMIT = -5
SRAY_ERR_PARSE = -6
SRAY_ERR_PENDING = -7
SRAY_ERR_NODEVICE = -8

SRAY_OBJ_SPHERE, SRAY_OBJ_CYLINDER, SRAY_OBJ_MESH = 0, 1, 2
SRAY_LIGHT_POINT, SRAY_LIGHT_SPHERE, SRAY_LIGHT_BOX = 0, 1, 2
SRAY_MAT_HAS_SPECULAR, SRAY_MAT_HAS_REFLECT, SRAY_MAT_HAS_REFRACT, SRAY_MAT_PHONG, SRAY_MAT_HAS_NORMAL = 1, 2, 4, 8, 16
(SRAY_MAP_DIFFUSE, SRAY_MAP_SPECULAR, SRAY_MAP_SHININESS, SRAY_MAP_REFLECT, SRAY_MAP_REFRACT, SRAY_MAP_IOR,
 SRAY_MAP_NORMAL) = range(7)
SRAY_TEX_FILTER_LINEAR, SRAY_TEX_FILTER_NEAREST, SRAY_TEX_WRAP_REPEAT, SRAY_TEX_WRAP_CLAMP = 0, 1, 0, 1
SRAY_TRACE_CLOSEST, SRAY_TRACE_ANY, SRAY_TRACE_COUNT, SRAY_TRACE_WAVE, SRAY_TRACE_SPARSE = 0, 1, 2, 4, 8
SRAY_RENDER_REF_Q1 = 1
SRAY_NOALLOC_MAX_QUERIES = 1 << 20
SRAY_RENDER_COUNT = 2
SRAY_PMAP_CAUSTIC, SRAY_PMAP_GLOBAL = 0, 1
SRAY_GATHER_IRRADIANCE, SRAY_GATHER_RADIANCE = 0, 1
SRAY_GATHER_GRID = 0x10
SRAY_REQUEST_DEVICE_OUT = 0x1

RAY_MAGNITUDE = 10000.0
ERROR_MARGIN = 1e-6
XSMALL_NUMBER = 1e-8

f32, i32, u32, u64, f64 = C.c_float, C.c_int32, C.c_uint32, C.c_uint64, C.c_double


class ObjectDesc(C.Structure):
    _fields_ = [("type", i32), ("material", i32), ("mesh", i32), ("radius", f32), ("start", f32),
                ("end", f32), ("xform", f32 * 12), ("inv_xform", f32 * 12), ("motion", f32 * 3),
                ("motion_t0", i32), ("motion_t1", i32), ("time_xforms", C.POINTER(f32))]


class MeshDesc(C.Structure):
    _fields_ = [("tri_start", u32), ("tri_count", u32)]


class MaterialDesc(C.Structure):
    _fields_ = [("diffuse", f32 * 4), ("specular", f32 * 4), ("shininess", f32), ("reflect", f32),
                ("refract", f32), ("ior", f32), ("flags", u32), ("normal", f32), ("map", i32 * 7), ("pad_", i32)]


class TextureDesc(C.Structure):
    _fields_ = [("width", i32), ("height", i32), ("texels", C.POINTER(f32)), ("filter", i32), ("wrap", i32),
                ("xform", f32 * 12)]


class LightDesc(C.Structure):
    _fields_ = [("type", i32), ("pos", f32 * 3), ("color", f32 * 4), ("radius", f32), ("dim", f32 * 3),
                ("xform", f32 * 12), ("time_pos", C.POINTER(f32))]


class CameraDesc(C.Structure):
    _fields_ = [("xform", f32 * 12), ("vfov", f64), ("max_dist", f32), ("shutter", i32),
                ("time_xforms", C.POINTER(f32))]


class SceneDesc(C.Structure):
    _fields_ = [("num_objects", u32), ("objects", C.POINTER(ObjectDesc)),
                ("num_meshes", u32), ("meshes", C.POINTER(MeshDesc)),
                ("num_triangles", u32), ("tri_pos", C.POINTER(f32)), ("tri_normal", C.POINTER(f32)),
                ("tri_tangent", C.POINTER(f32)), ("tri_texcoord", C.POINTER(f32)),
                ("num_materials", u32), ("materials", C.POINTER(MaterialDesc)),
                ("num_lights", u32), ("lights", C.POINTER(LightDesc)),
                ("camera", CameraDesc), ("env_ambient", f32 * 4), ("env_color", f32 * 4), ("env_ior", f32),
                ("scene_oct_depth", i32), ("scene_oct_items", i32), ("mesh_oct_depth", i32),
                ("mesh_oct_items", i32), ("frame_t0", i32), ("frame_t1", i32),
                ("num_textures", u32), ("textures", C.POINTER(TextureDesc))]


class Ray(C.Structure):
    _fields_ = [("origin", f32 * 3), ("dir", f32 * 3), ("time", i32), ("pad_", i32)]


class Hit(C.Structure):
    _fields_ = [("t", f32), ("obj", i32), ("prim", i32), ("rank", u32)]


class TraceCounts(C.Structure):
    _fields_ = [("nodes", u64), ("spheres", u64), ("cylinders", u64), ("triangles", u64),
                ("item_refs", u64), ("mesh_enters", u64),
                ("item_boxes", u64)]


class RenderParams(C.Structure):
    _fields_ = [("width", i32), ("height", i32), ("x0", i32), ("y0", i32), ("x1", i32), ("y1", i32),
                ("min_samples", i32), ("max_samples", i32), ("max_var", f32), ("min_energy", f32),
                ("iter", i32), ("shadow_samples", i32), ("diffuse_samples", i32), ("frame_time", i32),
                ("mblur", i32), ("caust_photons", i32), ("gi_photons", i32), ("gather_dist", f32),
                ("seed", u64), ("flags", u32), ("gather_k", i32)]


class Stats(C.Structure):
    _fields_ = [("rays_primary", u64), ("rays_shadow", u64), ("rays_reflect", u64), ("rays_refract", u64),
                ("rays_diffuse", u64), ("samples", u64), ("render_ms", f64), ("accel_ms", f64),
                ("closest", TraceCounts), ("shadow", TraceCounts)]

    def rays(self):
        return (self.rays_primary + self.rays_shadow + self.rays_reflect + self.rays_refract
                + self.rays_diffuse)


class Photon(C.Structure):
    _fields_ = [("pos", f32 * 3), ("dir", f32 * 3), ("norm", f32 * 3), ("col", f32 * 3)]


class PhotonParams(C.Structure):
    _fields_ = [("caust_photons", i32), ("gi_photons", i32), ("photon_energy", f32), ("iter", i32), ("t0", i32),
                ("t1", i32), ("seed", u64)]


class PhotonStats(C.Structure):
    _fields_ = [("shot", u64 * 2), ("stored", u64 * 2), ("segments", u64), ("trace_ms", f64), ("build_ms", f64)]


class Tuning(C.Structure):
    _fields_ = [("trace_block", i32), ("tile_w", i32), ("gather_lanes", i32), ("chunk_log2", i32),
                ("gather_cell", f32), ("graphs", i32), ("persistent", i32), ("lanes", i32),
                ("sparse_from", i32), ("chain", i32), ("sample_runs", i32)]


class GatherQuery(C.Structure):
    _fields_ = [("pos", f32 * 3), ("norm", f32 * 3), ("dir", f32 * 3), ("pad_", f32)]


class Request(C.Structure):
    _fields_ = [("scene_id", u32), ("params", RenderParams), ("camera", CameraDesc), ("use_camera", i32),
                ("flags", i32)]


class Result(C.Structure):
    _fields_ = [("rgba", C.POINTER(f32)), ("rays", u64), ("latency_ms", f64), ("queue_ms", f64)]


class OctreeInfo(C.Structure):
    _fields_ = [("num_nodes", u32), ("num_leaves", u32), ("num_item_refs", u32), ("height", u32),
                ("root_min", f32 * 3), ("root_max", f32 * 3)]


class KernelStat(C.Structure):
    _fields_ = [("name", C.c_char * 32), ("launches", u64), ("total_ms", f64)]


ABI_STRUCTS = [ObjectDesc, MeshDesc, MaterialDesc, TextureDesc, LightDesc, CameraDesc, SceneDesc, Ray, Hit, TraceCounts,
               RenderParams, Stats, Photon, GatherQuery, Request, Result, OctreeInfo, KernelStat, PhotonParams,
               PhotonStats, Tuning]

# every extern "C" entry point declared in include/sray_gpu.h
EXPORTS = [
    "sray_gpu_create", "sray_gpu_destroy", "sray_gpu_last_error", "sray_gpu_scene_upload",
    "sray_gpu_scene_release", "sray_gpu_scene_bytes", "sray_gpu_trace", "sray_gpu_trace_device",
    "sray_gpu_render", "sray_gpu_render_device", "sray_gpu_pmap_upload", "sray_gpu_pmap_release",
    "sray_gpu_pmap_gather", "sray_gpu_submit", "sray_gpu_submit_many", "sray_gpu_poll", "sray_gpu_set_admission_cap",
    "sray_scene_load_xml", "sray_scene_desc_of", "sray_scene_free", "sray_host_last_error",
    "sray_octree_build_host", "sray_scene_generate", "sray_scene_write_xml", "sray_abi_sizes",
    "sray_gpu_primary_rays_device", "sray_gpu_set_profiling", "sray_gpu_kernel_stats",
    "sray_gpu_malloc", "sray_gpu_free", "sray_gpu_memcpy", "sray_gpu_synchronize",
    "sray_gpu_shoot_photons", "sray_gpu_pmap_download", "sray_projmap_build", "sray_projmap_cell",
    "sray_gpu_set_tuning", "sray_gpu_get_tuning", "sray_gpu_autotune", "sray_gpu_autotune_octree",
    "sray_photon_dump", "sray_photon_restore", "sray_photon_free", "sray_image_write_pfm", "sray_image_write_png",
    "sray_status_record",
    "sray_io_last_error", "sray_gpu_set_status_fd", "sray_gpu_pmap_gather_device_noalloc",
    "sray_scene_set_frame", "sray_image_read", "sray_image_free",
]
SRAY_STATUS_RECORD = 32


def declare(lib):
    """Attach argtypes/restype to every entry point of the loaded library."""
    P = C.POINTER
    vp = C.c_void_p
    sig = {
        "sray_gpu_create": (i32, [C.c_int, u32, P(vp)]),
        "sray_gpu_destroy": (None, [vp]),
        "sray_gpu_last_error": (C.c_char_p, [vp]),
        "sray_gpu_scene_upload": (i32, [vp, P(SceneDesc), P(u32)]),
        "sray_gpu_scene_release": (i32, [vp, u32]),
        "sray_gpu_scene_bytes": (i32, [vp, u32, P(u64)]),
        "sray_gpu_trace": (i32, [vp, u32, P(Ray), u32, P(Hit), u32, P(TraceCounts)]),
        "sray_gpu_trace_device": (i32, [vp, u32, vp, u32, vp, u32]),
        "sray_gpu_render": (i32, [vp, u32, P(RenderParams), P(f32), P(Stats)]),
        "sray_gpu_render_device": (i32, [vp, u32, P(RenderParams), vp, P(Stats)]),
        "sray_gpu_primary_rays_device": (i32, [vp, u32, P(RenderParams), C.c_int, vp]),
        "sray_gpu_pmap_upload": (i32, [vp, u32, P(Photon), u32, C.c_int, P(u32)]),
        "sray_gpu_pmap_release": (i32, [vp, u32]),
        "sray_gpu_pmap_gather": (i32, [vp, u32, P(GatherQuery), u32, f32, C.c_int, C.c_int, P(f32)]),
        "sray_gpu_submit": (i32, [vp, P(Request), P(f32), P(u64)]),
        "sray_gpu_submit_many": (i32, [vp, P(Request), u32, P(P(f32)), P(u64)]),
        "sray_gpu_poll": (i32, [vp, u64, C.c_int, P(Result)]),
        "sray_gpu_set_admission_cap": (i32, [vp, u64]),
        "sray_scene_load_xml": (i32, [C.c_char_p, C.c_char_p, P(vp)]),
        "sray_scene_desc_of": (P(SceneDesc), [vp]),
        "sray_scene_free": (None, [vp]),
        "sray_host_last_error": (C.c_char_p, []),
        "sray_octree_build_host": (i32, [P(SceneDesc), C.c_int, P(OctreeInfo), P(f32), P(f32), P(u32), u32, u32]),
        "sray_scene_generate": (i32, [C.c_int, u64, C.c_int, P(vp)]),
        "sray_scene_write_xml": (i32, [vp, C.c_char_p, C.c_size_t, P(C.c_size_t)]),
        "sray_abi_sizes": (i32, [P(u32), C.c_int]),
        "sray_gpu_set_profiling": (i32, [vp, C.c_int]),
        "sray_gpu_malloc": (i32, [vp, C.c_size_t, P(vp)]),
        "sray_gpu_free": (i32, [vp, vp]),
        "sray_gpu_memcpy": (i32, [vp, vp, vp, C.c_size_t, C.c_int]),
        "sray_gpu_synchronize": (i32, [vp]),
        "sray_gpu_kernel_stats": (i32, [vp, P(KernelStat), C.c_int, P(C.c_int), C.c_int]),
        "sray_gpu_shoot_photons": (i32, [vp, u32, P(PhotonParams), P(u32), P(u32), P(PhotonStats)]),
        "sray_gpu_pmap_download": (i32, [vp, u32, vp, u32, P(u32)]),
        "sray_projmap_build": (i32, [P(SceneDesc), C.c_int, C.c_int, C.c_int, C.c_int, u64, vp, P(f64)]),
        "sray_projmap_cell": (i32, [f32, f32, f32, P(C.c_int), P(C.c_int)]),
        "sray_gpu_set_tuning": (i32, [vp, P(Tuning)]),
        "sray_gpu_get_tuning": (i32, [vp, P(Tuning)]),
        "sray_gpu_autotune": (i32, [vp, u32, P(RenderParams), C.c_int, P(Tuning), P(f64)]),
        "sray_gpu_autotune_octree": (i32, [vp, P(SceneDesc), P(RenderParams), C.c_int, P(i32), P(i32), C.c_int,
                                           P(C.c_int), P(f64)]),
        "sray_photon_dump": (i32, [C.c_char_p, P(Photon), u32, f32]),
        "sray_photon_restore": (i32, [C.c_char_p, f32, P(P(Photon)), P(u32)]),
        "sray_photon_free": (None, [P(Photon)]),
        "sray_image_write_pfm": (i32, [C.c_char_p, P(f32), C.c_int, C.c_int]),
        "sray_image_write_png": (i32, [C.c_char_p, P(f32), C.c_int, C.c_int]),
        "sray_status_record": (i32, [C.c_char, C.c_int, C.c_int, C.c_int, C.c_int, C.c_char_p]),
        "sray_io_last_error": (C.c_char_p, []),
        "sray_gpu_set_status_fd": (i32, [vp, C.c_int]),
        "sray_gpu_pmap_gather_device_noalloc": (i32, [vp, u32, vp, u32, f32, C.c_int, vp]),
        "sray_scene_set_frame": (i32, [vp, i32]),
        "sray_image_read": (i32, [C.c_char_p, P(P(f32)), P(C.c_int), P(C.c_int)]),
        "sray_image_free": (None, [P(f32)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib
