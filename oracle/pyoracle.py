"""ctypes binding of liboracle.so (oracle/oracle.c). TEST INFRASTRUCTURE ONLY: used by tests/,
__graft_entry__.smoke() and the CPU-baseline leg of bench.py, never by the product."""
import ctypes as C
import os

import numpy as np

from specrel_amd import _ffi as F

HERE = os.path.dirname(os.path.abspath(__file__))
# SPECREL_ORACLE_LIB names another build of the same source (tests/run_sanitized.sh's).
LIB_PATH = os.environ.get("SPECREL_ORACLE_LIB") or os.path.join(HERE, "liboracle.so")


class Counts(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("rays", "node_pops", "box_tests", "prim_tests", "hits", "shadow_rays",
                                          "shadow_node_pops", "shadow_box_tests", "shadow_prim_tests",
                                          "star_pairs", "distinct_nodes", "distinct_prims", "distinct_hits")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class Record(C.Structure):
    """oracle_record: a hit record given field by field (HitRecord4D, mesh4d.rs:382-425)."""
    _fields_ = [("lambda_", C.c_float), ("event", C.c_float * 4), ("tet_bary", C.c_float * 4),
                ("tri_bary", C.c_float * 3), ("uv", C.c_float * 2), ("proper_time", C.c_float),
                ("tau_uv_grad", C.c_float * 2), ("uv_world_scale", C.c_float * 2), ("normal", C.c_float * 4),
                ("four_velocity", C.c_float * 4), ("decided_at", C.c_float), ("material", C.c_uint32)]


class ShadeParts(C.Structure):
    _fields_ = [("albedo", C.c_float * 3), ("lighting", C.c_float), ("cos_light", C.c_float),
                ("cos_view", C.c_float), ("t_emit", C.c_float), ("dist", C.c_float),
                ("light_shadowed", C.c_int32), ("flash", C.c_float * 3), ("lit", C.c_float * 3)]

    def as_dict(self):
        return {n: (np.array(getattr(self, n)[:], np.float32) if n in ("albedo", "flash", "lit") else getattr(self, n))
                for n, _ in self._fields_}


_lib = None
_FP = C.POINTER(C.c_float)
_U32P = C.POINTER(C.c_uint32)
_P = C.c_void_p


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: build it with oracle/build.sh or __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        sig = {
            "oracle_render": (C.c_int, [C.POINTER(F.SceneArrays), C.POINTER(F.Frame), C.c_int, _U32P, _U32P, _FP,
                                        _FP, _FP, C.POINTER(Counts)]),
            "oracle_render_rows": (C.c_int, [C.POINTER(F.SceneArrays), C.POINTER(F.Frame), C.c_int, C.c_uint32,
                                             C.c_uint32, C.c_uint32, _FP, _U32P, _FP, _FP, C.POINTER(Counts)]),
            "oracle_bloom": (None, [_FP, C.c_uint32, C.c_uint32]),
            "oracle_pack": (None, [_FP, C.c_uint32, _U32P]),
            "oracle_gamma_lut": (None, [C.POINTER(C.c_uint8)]),
            "oracle_intersect": (C.c_int, [_P, _FP, _FP, C.c_float, C.c_float, _FP]),
            "oracle_trace": (C.c_int64, [C.POINTER(F.SceneArrays), _FP, _FP, C.c_float, C.c_float, _FP,
                                         C.POINTER(Counts)]),
            "oracle_trace_brute": (C.c_int64, [C.POINTER(F.SceneArrays), _FP, _FP, C.c_float, C.c_float, _FP]),
            "oracle_occluded": (C.c_int, [C.POINTER(F.SceneArrays), _FP, _FP, C.c_float, C.c_float]),
            "oracle_occluded_brute": (C.c_int, [C.POINTER(F.SceneArrays), _FP, _FP, C.c_float, C.c_float]),
            "oracle_star_light": (None, [_FP, C.c_uint32, _FP, _FP, _FP, C.c_float, C.c_float, C.c_float,
                                         C.c_float, _FP]),
            "oracle_cone_event": (C.c_int, [_FP, C.c_uint32, C.c_uint32, _FP, _FP, C.c_float, _FP, _FP]),
            "oracle_blackbody": (None, [C.c_float, _FP]),
            "oracle_color_norm": (C.c_float, [C.c_float, C.c_float]),
            "oracle_digital_tau": (C.c_int, [C.c_float, C.c_float, C.c_float]),
            "oracle_decision_age": (C.c_float, [_FP, _FP, _FP]),
            "oracle_shade_sky": (None, [_FP, _FP]),
            "oracle_haze": (None, [_FP, _FP, _FP, C.c_uint32, _FP]),
            "oracle_expf": (C.c_float, [C.c_float]),
            "oracle_sinf": (C.c_float, [C.c_float]),
            "oracle_cosf": (C.c_float, [C.c_float]),
            "oracle_atan2f": (C.c_float, [C.c_float, C.c_float]),
            "oracle_acosf": (C.c_float, [C.c_float]),
            "oracle_rem_euclid": (C.c_float, [C.c_float, C.c_float]),
            "oracle_pixel_states": (C.c_int, [C.POINTER(F.SceneArrays), C.POINTER(F.Frame), C.c_int, C.c_void_p]),
            "oracle_shade_record": (C.c_int, [C.POINTER(F.SceneArrays), C.POINTER(F.Frame), _FP,
                                              C.POINTER(Record), _FP, C.POINTER(ShadeParts)]),
            "oracle_pixel_ray": (None, [C.POINTER(F.Frame), C.c_uint32, C.c_uint32, _FP, _FP]),
            "oracle_sky_gradient": (None, [_FP, _FP]),
            "oracle_distance_fog": (None, [_FP, _FP, C.c_float, _FP]),
            "oracle_math_vec": (None, [C.c_int, _FP, _FP, _FP, C.c_uint64]),
            "oracle_adjoint_ref": (C.c_int, [C.POINTER(F.SceneArrays), C.POINTER(F.Frame), C.POINTER(C.c_double),
                                             _FP, C.POINTER(C.c_uint8), _U32P, _U32P, _FP, C.c_uint32, C.c_int, C.c_int,
                                             C.c_int,
                                             C.POINTER(C.c_double), C.POINTER(C.c_double),
                                             C.POINTER(C.c_uint64), C.POINTER(C.c_uint8), C.POINTER(C.c_double)]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def threads():
    """Worker threads of an oracle call: the cores this process may run on, at most 16 (the host
    share of one GPU)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def f32(a, n=None):
    a = np.ascontiguousarray(np.asarray(a, np.float32))
    return a, a.ctypes.data_as(_FP)


def render(scene, frame, nthreads=None, want_hdr=False, counts=False):
    """Full oracle frame: returns dict with rgb, hit, lam, tau (and hdr, counts if asked)."""
    w, h = frame.size[0], frame.size[1]
    n = w * h
    rgb = np.zeros(n, np.uint32)
    hit = np.zeros(n, np.uint32)
    lam = np.zeros(n, np.float32)
    tau = np.zeros(n, np.float32)
    hdr = np.zeros(n * 3, np.float32) if want_hdr else None
    cnt = Counts() if counts else None
    nthreads = nthreads or threads()
    lib().oracle_render(C.byref(scene.arrays), C.byref(frame), nthreads, rgb.ctypes.data_as(_U32P),
                        hit.ctypes.data_as(_U32P), lam.ctypes.data_as(_FP), tau.ctypes.data_as(_FP),
                        None if hdr is None else hdr.ctypes.data_as(_FP), None if cnt is None else C.byref(cnt))
    out = {"rgb": rgb.reshape(h, w), "hit": hit.reshape(h, w), "lam": lam.reshape(h, w), "tau": tau.reshape(h, w)}
    if hdr is not None:
        out["hdr"] = hdr.reshape(h, w, 3)
    if cnt is not None:
        out["counts"] = cnt.as_dict()
    return out


def render_rows(scene, frame, row0, step, nrows, nthreads=None, counts=False, hits=False):
    """Rows row0, row0 + step, ... of the frame: the pre-bloom HDR (nrows, w, 3), the counts when
    asked, and with hits a dict of the rows' hit index / lambda / tau (nrows, w)."""
    w = frame.size[0]
    hdr = np.zeros(nrows * w * 3, np.float32)
    cnt = Counts() if counts else None
    hit = np.zeros(nrows * w, np.uint32) if hits else None
    lam = np.zeros(nrows * w, np.float32) if hits else None
    tau = np.zeros(nrows * w, np.float32) if hits else None
    lib().oracle_render_rows(C.byref(scene.arrays), C.byref(frame), nthreads or threads(), row0, step,
                             nrows, hdr.ctypes.data_as(_FP), None if hit is None else hit.ctypes.data_as(_U32P),
                             None if lam is None else lam.ctypes.data_as(_FP),
                             None if tau is None else tau.ctypes.data_as(_FP), None if cnt is None else C.byref(cnt))
    out = (hdr.reshape(nrows, w, 3), (cnt.as_dict() if cnt is not None else None))
    if hits:
        out = out + ({"hit": hit.reshape(nrows, w), "lam": lam.reshape(nrows, w), "tau": tau.reshape(nrows, w)},)
    return out


def bloom(hdr):
    h, w, _ = hdr.shape
    a = np.ascontiguousarray(hdr, np.float32).copy()
    lib().oracle_bloom(a.ctypes.data_as(_FP), w, h)
    return a


def pack(hdr):
    a = np.ascontiguousarray(hdr, np.float32)
    n = a.size // 3
    out = np.zeros(n, np.uint32)
    lib().oracle_pack(a.ctypes.data_as(_FP), n, out.ctypes.data_as(_U32P))
    return out.reshape(a.shape[:-1])


def gamma_lut():
    out = (C.c_uint8 * 4096)()
    lib().oracle_gamma_lut(out)
    return np.frombuffer(out, np.uint8).copy()


def trace(scene, o, d, tmin=1e-3, tmax=180.0, brute=False):
    oo, op = f32(o)
    dd, dp = f32(d)
    out = np.zeros(9, np.float32)
    if brute:
        slot = lib().oracle_trace_brute(C.byref(scene.arrays), op, dp, tmin, tmax, out.ctypes.data_as(_FP))
    else:
        slot = lib().oracle_trace(C.byref(scene.arrays), op, dp, tmin, tmax, out.ctypes.data_as(_FP), None)
    return slot, out


def occluded(scene, o, d, tmin, tmax, brute=False):
    oo, op = f32(o)
    dd, dp = f32(d)
    fn = lib().oracle_occluded_brute if brute else lib().oracle_occluded
    return bool(fn(C.byref(scene.arrays), op, dp, tmin, tmax))


def intersect(geom20, o, d, tmin, tmax):
    g, gp = f32(geom20)
    oo, op = f32(o)
    dd, dp = f32(d)
    out = np.zeros(9, np.float32)
    ok = lib().oracle_intersect(g.ctypes.data, op, dp, tmin, tmax, out.ctypes.data_as(_FP))
    return (out if ok else None)


def star_light(words, num_stars, origin, d, e0, cone, width, spread, lambda_max):
    w, wp = f32(words)
    o, op = f32(origin)
    dd, dp = f32(d)
    e, ep = f32(e0)
    out = np.zeros(3, np.float32)
    lib().oracle_star_light(wp, num_stars, op, dp, ep, cone, width, spread, lambda_max, out.ctypes.data_as(_FP))
    return out


def cone_event(words, num_stars, star, origin, e0, cone):
    w, wp = f32(words)
    o, op = f32(origin)
    e, ep = f32(e0)
    d = np.zeros(4, np.float32)
    u = np.zeros(4, np.float32)
    ok = lib().oracle_cone_event(wp, num_stars, star, op, ep, cone, d.ctypes.data_as(_FP), u.ctypes.data_as(_FP))
    return (d, u) if ok else None


def adjoint_ref(scene, frame, theta, target, objects, mask_in=None, nthreads=None, frame_values=True, strict=False):
    """FP64 forward-mode gradient of the masked image loss (adjoint_ref.cpp) in the parameters
    [event(4), phi(3), rot(3), d_0(3), ...]; `frame` must be the f32 frame of `theta`. With
    frame_values the model is evaluated at the frame's f32 event and tetrad (the renderer's
    inputs), else at theta in double."""
    w, h = frame.size[0], frame.size[1]
    th = np.ascontiguousarray(theta, np.float64)
    tgt = np.ascontiguousarray(target, np.float32)
    n = len(objects)
    first = np.array([o[0] for o in objects] or [0], np.uint32)
    count = np.array([o[1] for o in objects] or [0], np.uint32)
    center = np.array([o[2] for o in objects] or [(0, 0, 0, 0)], np.float32)
    mi = None if mask_in is None else np.ascontiguousarray(mask_in, np.uint8)
    mo = np.zeros((h, w), np.uint8)
    loss = C.c_double()
    grad = np.zeros(10 + 3 * n, np.float64)
    cnt = np.zeros(6, np.uint64)
    err = C.c_double()
    u8 = C.POINTER(C.c_uint8)
    rc = lib().oracle_adjoint_ref(C.byref(scene.arrays), C.byref(frame), th.ctypes.data_as(C.POINTER(C.c_double)),
                                  tgt.ctypes.data_as(_FP), None if mi is None else mi.ctypes.data_as(u8),
                                  first.ctypes.data_as(_U32P), count.ctypes.data_as(_U32P),
                                  center.ctypes.data_as(_FP), n, int(frame_values), int(strict), nthreads or threads(),
                                  C.byref(loss),
                                  grad.ctypes.data_as(C.POINTER(C.c_double)),
                                  cnt.ctypes.data_as(C.POINTER(C.c_uint64)), mo.ctypes.data_as(u8), C.byref(err))
    if rc != 0:
        raise RuntimeError("oracle_adjoint_ref: unsupported view")
    names = ("valid", "miss", "starlight", "border", "silhouette", "edge")
    return {"loss": loss.value, "grad": grad, "counts": dict(zip(names, (int(c) for c in cnt))), "mask": mo,
            "max_rel_color_err": err.value}


PIXEL_STATE = np.dtype([("slot", np.int64), ("nw_albedo", np.float32, 3), ("col0", np.float32, 3),
                        ("dark", np.float32), ("pattern", np.int32), ("in_shadow", np.int32),
                        ("color", np.float32, 3), ("local_dir", np.float32, 3), ("material", np.uint32),
                        ("flags", np.uint32), ("mtb_k", np.int32)])


def pixel_states(scene, frame, nthreads=None):
    """The f32 forward state of every pixel (oracle_pixel_states) as a structured (h, w) array."""
    w, h = frame.size[0], frame.size[1]
    out = np.zeros(w * h, PIXEL_STATE)
    assert out.itemsize == 80
    if lib().oracle_pixel_states(C.byref(scene.arrays), C.byref(frame), nthreads or threads(),
                                 out.ctypes.data) != 0:
        raise RuntimeError("oracle_pixel_states: unsupported view")
    return out.reshape(h, w)


def record(event=(0, 0, 0, 0), lam=1.0, tet_bary=(0.25, 0.25, 0.25, 0.25), tri_bary=(0.3, 0.3, 0.4), uv=(0.5, 0.5),
           tau=0.0, tau_uv_grad=(0.0, 0.0), uv_world_scale=(1.0, 1.0), normal=(0, 0, 1, 0), four_velocity=(1, 0, 0, 0),
           decided_at=-1e9, material=0):
    r = Record()
    r.lambda_ = lam
    for name, val in (("event", event), ("tet_bary", tet_bary), ("tri_bary", tri_bary), ("uv", uv),
                      ("tau_uv_grad", tau_uv_grad), ("uv_world_scale", uv_world_scale), ("normal", normal),
                      ("four_velocity", four_velocity)):
        arr = getattr(r, name)
        for i, v in enumerate(val):
            arr[i] = v
    r.proper_time = tau
    r.decided_at = decided_at
    r.material = material
    return r


def shade_record(scene, frame, direction, rec):
    """shade_hit of a hand-built record: (colour before haze and stars, intermediate parts)."""
    d, dp = f32(direction)
    out = np.zeros(3, np.float32)
    parts = ShadeParts()
    if lib().oracle_shade_record(C.byref(scene.arrays), C.byref(frame), dp, C.byref(rec), out.ctypes.data_as(_FP),
                                 C.byref(parts)) != 0:
        raise ValueError("material out of range")
    return out, parts.as_dict()


def pixel_ray(frame, px, py):
    o = np.zeros(4, np.float32)
    d = np.zeros(4, np.float32)
    lib().oracle_pixel_ray(C.byref(frame), px, py, o.ctypes.data_as(_FP), d.ctypes.data_as(_FP))
    return o, d


def sky_gradient(direction):
    d, dp = f32(direction)
    out = np.zeros(3, np.float32)
    lib().oracle_sky_gradient(dp, out.ctypes.data_as(_FP))
    return out


def distance_fog(color, direction, lam):
    c, cp = f32(color)
    d, dp = f32(direction)
    out = np.zeros(3, np.float32)
    lib().oracle_distance_fog(cp, dp, lam, out.ctypes.data_as(_FP))
    return out


MATH_FNS = {"exp": 0, "sin": 1, "cos": 2, "atan2": 3, "acos": 4}


def math_vec(fn, a, b=None):
    """The oracle's f32 helper `fn` over arrays (atan2: a = y, b = x)."""
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(a if b is None else b, np.float32)
    out = np.empty_like(a)
    lib().oracle_math_vec(MATH_FNS[fn], a.ctypes.data_as(_FP), b.ctypes.data_as(_FP), out.ctypes.data_as(_FP), a.size)
    return out
