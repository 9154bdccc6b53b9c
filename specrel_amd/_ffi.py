"""ctypes mirrors of include/specrel_types.h and include/specrel_scene.h.

The structures are the boundary's byte layouts (GpuNode/GpuGeom/GpuAttr/GpuMaterial of
src/gpu.rs:42-100 and the per-view frame block); their sizes are asserted on import.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(_HERE, "lib")

F4 = C.c_float * 4


class Node(C.Structure):
    _fields_ = [("min", F4), ("max", F4), ("right_or_first", C.c_uint32), ("count", C.c_uint32),
                ("left", C.c_uint32), ("pad", C.c_uint32)]


class Geom(C.Structure):
    _fields_ = [("v0", F4), ("n", F4), ("b1", F4), ("b2", F4), ("b3", F4)]


class Attr(C.Structure):
    _fields_ = [(n, F4) for n in ("uv_u", "uv_v", "tb_x", "tb_y", "tb_z", "tau", "grad_u", "grad_v",
                                  "normal", "four_vel", "misc", "decided_at")]


class Material(C.Structure):
    _fields_ = [("base", F4), ("checker", F4), ("params", F4)]


class SceneArrays(C.Structure):
    _fields_ = [("nodes", C.c_void_p), ("num_nodes", C.c_uint32), ("num_prims", C.c_uint32),
                ("geoms", C.c_void_p), ("attrs", C.c_void_p), ("order", C.c_void_p),
                ("materials", C.c_void_p), ("num_materials", C.c_uint32), ("num_stars", C.c_uint32),
                ("star_words", C.c_void_p), ("num_star_words", C.c_uint32), ("pad", C.c_uint32),
                ("light_pos", F4)]


class Frame(C.Structure):
    _fields_ = [("cam_pos", F4), ("e", F4 * 4), ("half_wh", C.c_float * 2), ("size", C.c_uint32 * 2),
                ("mode", C.c_uint32), ("shadow_mode", C.c_uint32), ("slice_mode", C.c_uint32),
                ("num_flashes", C.c_uint32), ("frontier_origin", F4), ("frontier_tau", C.c_float),
                ("frontier_view_offset", C.c_float), ("prophecy_delta", C.c_float), ("num_haze", C.c_uint32),
                ("haze_origin", F4 * 5), ("flash_at", F4 * 16), ("flash_color_width", F4 * 16),
                ("flash_intensity", C.c_float * 16)]


class Stats(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("rays", "node_pops", "box_tests", "prim_tests", "hits", "shadow_rays",
                                          "shadow_node_pops", "shadow_box_tests", "shadow_prim_tests",
                                          "star_tests", "distinct_nodes", "distinct_prims", "distinct_hits")] + \
        [("reserved", C.c_uint64 * 3)]


class WlSample(C.Structure):
    _fields_ = [("center", F4), ("proper_time", C.c_float), ("phi", C.c_float * 3),
                ("proper_accel_local", C.c_float * 3), ("pad", C.c_float), ("e", F4 * 4)]


class Camera(C.Structure):
    _fields_ = [("position", F4), ("e", F4 * 4), ("proper_time", C.c_float), ("fov_y_rad", C.c_float),
                ("pad", C.c_float * 2)]


class RenderConfigC(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("mode", C.c_uint32), ("shadow_mode", C.c_uint32),
                ("slice_mode", C.c_uint32), ("prophecy_delta", C.c_float)]


class FlashC(C.Structure):
    _fields_ = [("at", F4), ("width", C.c_float), ("intensity", C.c_float), ("color", C.c_float * 3),
                ("pad", C.c_float)]


class DynamicsC(C.Structure):
    _fields_ = [("has_frontier", C.c_uint32), ("frontier_origin", F4), ("frontier_proper_time", C.c_float),
                ("frontier_lag", C.c_float), ("num_flashes", C.c_uint32), ("flashes", C.POINTER(FlashC)),
                ("num_remote", C.c_uint32), ("remote_origins", C.POINTER(F4))]


class Object(C.Structure):
    _fields_ = [("prim_first", C.c_uint32), ("prim_count", C.c_uint32), ("center", C.c_float * 4),
                ("material", C.c_uint32), ("pad", C.c_uint32)]


class GenParams(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("n_objects", C.c_uint32), ("slices", C.c_uint32), ("n_stars", C.c_uint32),
                ("n_clocks", C.c_uint32), ("icosphere_fraction", C.c_float), ("accel_fraction", C.c_float),
                ("accel_min", C.c_float), ("accel_max", C.c_float), ("t_min", C.c_float), ("t_max", C.c_float),
                ("first_object", C.c_uint32), ("pad", C.c_uint32)]


class Tube(C.Structure):
    """specrel_tube: a rest-frame triangle mesh swept along a sampled worldline."""
    _fields_ = [("vertices", C.POINTER(C.c_float)), ("tri_index", C.POINTER(C.c_uint32)),
                ("tri_uv", C.POINTER(C.c_float)), ("tri_normal", C.POINTER(C.c_float)),
                ("samples", C.POINTER(WlSample)), ("decided_at", C.POINTER(C.c_float)),
                ("n_vertices", C.c_uint32), ("n_triangles", C.c_uint32), ("n_samples", C.c_uint32),
                ("material", C.c_uint32)]


class FileInfo(C.Structure):
    _fields_ = [("version", C.c_uint32), ("flags", C.c_uint32), ("bvh_layout", C.c_uint64),
                ("bvh_edit", C.c_uint64), ("file_bytes", C.c_uint64), ("num_objects", C.c_uint32),
                ("reserved", C.c_uint32)]


class FileHeader(C.Structure):
    _fields_ = [("magic", C.c_char * 8), ("version", C.c_uint32), ("byte_order", C.c_uint32),
                ("header_bytes", C.c_uint32), ("num_sections", C.c_uint32), ("num_nodes", C.c_uint32),
                ("num_prims", C.c_uint32), ("num_materials", C.c_uint32), ("num_stars", C.c_uint32),
                ("num_star_words", C.c_uint32), ("num_objects", C.c_uint32), ("light_pos", C.c_float * 4),
                ("bvh_layout", C.c_uint64), ("bvh_edit", C.c_uint64), ("file_bytes", C.c_uint64),
                ("flags", C.c_uint32), ("reserved", C.c_uint32 * 9)]


class FileSection(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("elem_bytes", C.c_uint32), ("count", C.c_uint64), ("offset", C.c_uint64),
                ("checksum", C.c_uint32), ("reserved", C.c_uint32)]


assert C.sizeof(Tube) == 64 and C.sizeof(WlSample) == 112
assert C.sizeof(FileHeader) == 128 and C.sizeof(FileSection) == 32 and C.sizeof(FileInfo) == 40
assert C.sizeof(Node) == 48 and C.sizeof(Geom) == 80 and C.sizeof(Attr) == 192
assert C.sizeof(Material) == 48 and C.sizeof(Frame) == 800

# enums (include/specrel_types.h)
MODE_SHADED, MODE_ISOCHRONES, MODE_INSPECTOR, MODE_DECISION_AGE = 0, 1, 2, 3
SHADOW_OFF, SHADOW_NULL, SHADOW_CAMERA_SLICE = 0, 1, 2
VIEW_PAST_CONE, VIEW_SIMULTANEITY, VIEW_FRONTIER, VIEW_PROPHECY = 0, 1, 2, 3
MESH_CUBE, MESH_UV_SPHERE, MESH_GROUND, MESH_ICOSPHERE = 0, 1, 2, 3
WL_INERTIAL, WL_HYPERBOLIC, WL_TWIN, WL_CIRCLE, WL_BORN_RAMP, WL_OSCILLATE, WL_ORBIT = 0, 1, 2, 3, 4, 5, 6
NO_HIT = 0xFFFFFFFF

_P = C.c_void_p
_f = C.c_float
_u32 = C.c_uint32
_FP = C.POINTER(C.c_float)
_DP = C.POINTER(C.c_double)


def _scene_protos(lib):
    sig = {
        "specrel_scene_last_error": (C.c_char_p, []),
        "specrel_gen_defaults": (C.c_int, [C.c_char_p, C.POINTER(GenParams)]),
        "specrel_scene_generate": (C.c_int, [C.c_char_p, C.POINTER(GenParams), C.POINTER(_P)]),
        "specrel_scene_default_camera": (C.c_int, [C.c_char_p, C.POINTER(GenParams), C.POINTER(Camera)]),
        "specrel_scene_get_arrays": (C.c_int, [_P, C.POINTER(SceneArrays)]),
        "specrel_scene_star_params": (C.c_int, [_P, _FP, _u32]),
        "specrel_scene_objects": (_u32, [_P, C.POINTER(Object), _u32]),
        "specrel_scene_free": (None, [_P]),
        "specrel_builder_new": (_P, []),
        "specrel_builder_free": (None, [_P]),
        "specrel_builder_add_tube": (C.c_int, [_P, _u32, _FP, _u32, _DP, _f, _u32, _u32, _FP, _u32]),
        "specrel_builder_add_simplex": (C.c_int, [_P, _FP, _FP, _FP, _FP, _FP, _FP, _FP, _FP, _FP, _u32]),
        "specrel_builder_add_star": (C.c_int, [_P, _FP, _u32, _FP, _f, _f]),
        "specrel_builder_set_light": (C.c_int, [_P, _f, _f, _f]),
        "specrel_builder_set_materials": (C.c_int, [_P, C.POINTER(Material), _u32]),
        "specrel_builder_num_simplices": (_u32, [_P]),
        "specrel_builder_finish": (C.c_int, [_P, C.POINTER(_P)]),
        "specrel_worldline_sample": (C.c_int, [_u32, _DP, C.POINTER(WlSample), _u32, C.POINTER(_u32)]),
        "specrel_sync_proper_time": (C.c_int, [C.POINTER(WlSample), _u32, _f]),
        "specrel_extrude": (C.c_int, [_u32, _FP, C.POINTER(WlSample), _u32, _u32, C.POINTER(Geom),
                                      C.POINTER(Attr), _u32, C.POINTER(_u32)]),
        "specrel_tetrad_rest": (None, [_FP]),
        "specrel_tetrad_from_beta": (None, [_FP, _FP]),
        "specrel_tetrad_boost_local": (None, [_FP, _FP]),
        "specrel_tetrad_rotate_local": (None, [_FP, _f, _f, _f]),
        "specrel_tetrad_boost_to": (None, [_FP, _FP]),
        "specrel_frame_build": (C.c_int, [C.POINTER(Camera), C.POINTER(RenderConfigC), C.POINTER(DynamicsC),
                                          C.POINTER(Frame)]),
        "specrel_common_materials": (_u32, [C.POINTER(Material), _u32]),
        "specrel_build_tables": (None, [C.POINTER(_u32)]),
        "specrel_extrude_tubes_host": (C.c_int, [C.POINTER(Tube), _u32, C.POINTER(Geom), C.POINTER(Attr), _FP,
                                                 C.c_uint64, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
        "specrel_scene_generate_deferred": (C.c_int, [C.c_char_p, C.POINTER(GenParams), C.POINTER(_P)]),
        "specrel_builder_pending_tubes": (_u32, [_P, C.POINTER(Tube), _u32]),
        "specrel_builder_set_tube_records": (C.c_int, [_P, C.POINTER(Geom), C.POINTER(Attr), _FP, C.c_uint64,
                                                       C.POINTER(C.c_uint64)]),
        "specrel_builder_extrude_pending": (C.c_int, [_P]),
        "specrel_build_bvh_host": (C.c_int, [_FP, C.c_uint64, C.POINTER(Node), C.POINTER(C.c_uint32),
                                             C.POINTER(C.c_uint32)]),
        "specrel_builder_finish_with_tree": (C.c_int, [_P, C.POINTER(Node), C.c_uint32, C.POINTER(C.c_uint32),
                                                       C.POINTER(_P)]),
        "specrel_scene_file_write": (C.c_int, [C.c_char_p, C.POINTER(SceneArrays), C.POINTER(Object), _u32,
                                               C.c_uint64, C.c_uint64]),
        "specrel_scene_file_open": (C.c_int, [C.c_char_p, C.c_int, C.POINTER(_P), C.POINTER(SceneArrays),
                                              C.POINTER(FileInfo)]),
        "specrel_scene_file_objects": (C.POINTER(Object), [_P, C.POINTER(_u32)]),
        "specrel_scene_file_close": (None, [_P]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


_scene_lib = None


def scene_lib():
    """The scene producer library (built by specrel_amd.build; SPECREL_SCENE_LIB names another
    build of the same source, e.g. an address-sanitized one)."""
    global _scene_lib
    if _scene_lib is None:
        path = os.environ.get("SPECREL_SCENE_LIB") or os.path.join(LIB_DIR, "libspecrel_scene.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is missing: run specrel_amd.build.build_all() first")
        _scene_lib = _scene_protos(C.CDLL(path))
    return _scene_lib
