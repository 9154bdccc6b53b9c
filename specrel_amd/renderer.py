"""`GpuRenderer` for the gfx950 backend: the host-side mirror of src/gpu.rs:432-901.

    gpu = GpuRenderer([scene_a, scene_b])          # GpuRenderer::new (gpu.rs:472)
    gpu.render(0, camera, cfg, dynamics, fb)        # GpuRenderer::render (gpu.rs:798)
    gpu.update_scene(0, scene_a)                    # GpuRenderer::update_scene (gpu.rs:704)
    gpu.adapter_name(); gpu.last_timings            # gpu.rs:734, 451

Like the reference, construction raises (the Rust `Err`) when no usable device exists; the
caller then chooses a CPU renderer itself. There is no CPU path here: every render runs the
HIP kernels of libspecrel_hip.so.
"""
import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

from . import _ffi as F
from .scene import build_frame

OK, ENODEV, EINVAL, ENOMEM, EDEVICE = 0, -1, -2, -3, -4
OPT_STATS, OPT_STAR_CULL, OPT_STACK, OPT_PERSISTENT, OPT_SHADOW_QUEUE, OPT_GRAPH, OPT_MERGED_LOADS = 1, 2, 3, 4, 5, 6, 7
OPT_SHADOW_IN_TRACE, OPT_TRACE_WAVES, OPT_STAR_OVERLAP, OPT_WAVE_SHAPE, OPT_TILE_ORDER = 8, 9, 10, 11, 12
OPT_LEAF_POSTPONE = 13
OPT_ADJOINT_MASK = 14
OPT_WAVE_CLOCK = 15
OPT_NODE_WIDTH = 16
OPT_LDS_TOP = 17
OPT_FUSED_SHADE = 18
OPT_STAGE_EVENTS = 19
# the library's defaults for SPECREL_OPT_TRACE_WAVES and SPECREL_OPT_LDS_TOP
DEFAULT_TRACE_WAVES, DEFAULT_LDS_TOP = 1, 0
# specrel_timings.trace_variant: the primary-ray kernel's traversal
TRACE_PLAIN, TRACE_MERGED, TRACE_WIDE4, TRACE_LDS_TOP = 0, 1, 2, 3
# specrel_timings.shadow_path: where the last render traced its shadow queries
SHADOWS_NONE, SHADOWS_IN_SHADE, SHADOWS_QUEUED, SHADOWS_IN_TRACE = 0, 1, 2, 3


class Timings(C.Structure):
    _fields_ = [(n, C.c_float) for n in ("submit_ms", "wait_ms", "kernel_ms", "readback_ms", "star_prep_ms",
                                         "trace_ms", "shade_ms", "bloom_ms", "composite_ms", "adjoint_ms")] + \
        [("shadow_path", C.c_uint32), ("extrude_ms", C.c_float), ("trace_block_waves", C.c_uint32),
         ("trace_variant", C.c_uint32)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_ if n != "reserved"}


class ViewOut(C.Structure):
    _fields_ = [("rgb", C.POINTER(C.c_uint32)), ("hit", C.POINTER(C.c_uint32)), ("lambda_", C.POINTER(C.c_float)),
                ("tau", C.POINTER(C.c_float)), ("len", C.c_uint64)]


class AdjObject(C.Structure):
    _fields_ = [("prim_first", C.c_uint32), ("prim_count", C.c_uint32), ("center", C.c_float * 4)]


class AdjointOut(C.Structure):
    _fields_ = [("loss", C.c_double), ("d_event", C.c_double * 4), ("d_tetrad", (C.c_double * 4) * 4)] + \
        [(n, C.c_uint64) for n in ("valid", "miss", "starlight", "border", "silhouette", "edge")]


class Range(C.Structure):
    _fields_ = [("first", C.c_uint32), ("count", C.c_uint32)]


class WaveClock(C.Structure):
    _fields_ = [("start", C.c_uint64), ("end", C.c_uint64), ("tile", C.c_uint32), ("quarter", C.c_uint32),
                ("hw_id", C.c_uint32), ("xcc", C.c_uint32)]


WAVE_CLOCK_DTYPE = np.dtype([("start", "<u8"), ("end", "<u8"), ("tile", "<u4"), ("quarter", "<u4"),
                             ("hw_id", "<u4"), ("xcc", "<u4")])


class GpuUniforms(C.Structure):
    """GpuUniforms (gpu.rs:120-144), 768 B: the reference's per-view block."""
    _fields_ = [("cam_pos", C.c_float * 4), ("e", (C.c_float * 4) * 4), ("light_pos", C.c_float * 4),
                ("half_wh", C.c_float * 2), ("size", C.c_uint32 * 2), ("mode", C.c_uint32),
                ("shadow_mode", C.c_uint32), ("slice_mode", C.c_uint32), ("num_nodes", C.c_uint32),
                ("frontier_origin", C.c_float * 4), ("frontier_view_offset", C.c_float),
                ("num_flashes", C.c_uint32), ("frontier_tau", C.c_float), ("prophecy_delta", C.c_float),
                ("num_haze", C.c_uint32), ("num_stars", C.c_uint32), ("pad", C.c_uint32 * 2),
                ("haze_origin", (C.c_float * 4) * 5), ("flash_at", (C.c_float * 4) * 16),
                ("flash_col_width", (C.c_float * 4) * 16)]


def gpu_uniforms(frame, arrays) -> GpuUniforms:
    """GpuUniforms::build's block (gpu.rs:146-212) for a frame of specrel_amd.scene.build_frame and
    a scene's arrays (num_nodes, num_stars, light_pos): the same values, flash colours multiplied
    by their intensities in f32 as the reference does."""
    u = GpuUniforms()
    C.memmove(u.cam_pos, frame.cam_pos, 16)
    C.memmove(u.e, frame.e, 64)
    for i in range(4):
        u.light_pos[i] = arrays.light_pos[i]
    C.memmove(u.half_wh, frame.half_wh, 8)
    C.memmove(u.size, frame.size, 8)
    u.mode, u.shadow_mode, u.slice_mode = frame.mode, frame.shadow_mode, frame.slice_mode
    u.num_nodes, u.num_stars = arrays.num_nodes, arrays.num_stars
    C.memmove(u.frontier_origin, frame.frontier_origin, 16)
    u.frontier_view_offset, u.frontier_tau = frame.frontier_view_offset, frame.frontier_tau
    u.prophecy_delta, u.num_haze, u.num_flashes = frame.prophecy_delta, frame.num_haze, frame.num_flashes
    C.memmove(u.haze_origin, frame.haze_origin, 80)
    C.memmove(u.flash_at, frame.flash_at, 256)
    for i in range(frame.num_flashes):
        k = np.float32(frame.flash_intensity[i])
        for j in range(3):
            u.flash_col_width[i][j] = float(np.float32(frame.flash_color_width[i][j]) * k)
        u.flash_col_width[i][3] = frame.flash_color_width[i][3]
    return u


class BvhVersion(C.Structure):
    _fields_ = [("layout", C.c_uint64), ("edit", C.c_uint64)]


class Tuning(C.Structure):
    _fields_ = [("stack", C.c_int32), ("persistent", C.c_int32), ("shadow_queue", C.c_int32), ("graph", C.c_int32),
                ("candidates", C.c_uint32), ("best_ms", C.c_float), ("default_ms", C.c_float),
                ("shadow_in_trace", C.c_int32), ("wave_shape", C.c_int32), ("leaf_postpone", C.c_int32),
                ("trace_waves", C.c_int32), ("node_width", C.c_int32), ("lds_top", C.c_int32),
                ("fused_shade", C.c_int32)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_ if n != "reserved"}


class BvhStats(C.Structure):
    _fields_ = [("ms", C.c_float)] + [(n, C.c_uint32) for n in ("levels", "large_segments", "small_subtrees",
                                                                 "fallback_segments", "jump_passes")] + \
        [("reserved", C.c_uint32 * 2)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_ if n != "reserved"}


class SpecrelError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"specrel error {code}: {msg}")
        self.code = code


_lib = None
_P = C.c_void_p
_U32P = C.POINTER(C.c_uint32)
_FP = C.POINTER(C.c_float)

# Every symbol of include/specrel.h with its ctypes signature.
SIGNATURES = {
    "specrel_create": (C.c_int, [C.c_int, C.POINTER(_P)]),
    "specrel_destroy": (None, [_P]),
    "specrel_last_error": (C.c_char_p, [_P]),
    "specrel_device_name": (C.c_char_p, [_P]),
    "specrel_set_option": (C.c_int, [_P, C.c_int, C.c_int]),
    "specrel_upload_scene": (C.c_int, [_P, C.c_uint32, C.POINTER(F.SceneArrays)]),
    "specrel_upload_scene_file": (C.c_int, [_P, C.c_uint32, C.c_char_p, C.c_int]),
    "specrel_build_bvh": (C.c_int, [_P, _FP, C.c_uint64, C.POINTER(F.Node), _U32P, _U32P, C.POINTER(BvhStats)]),
    "specrel_extrude_tubes": (C.c_int, [_P, C.POINTER(F.Tube), C.c_uint32, C.POINTER(F.Geom), C.POINTER(F.Attr),
                                        _FP, C.c_uint64, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "specrel_patch_scene": (C.c_int, [_P, C.c_uint32, C.POINTER(F.SceneArrays), C.POINTER(Range), C.c_uint32,
                                     C.POINTER(Range), C.c_uint32, C.POINTER(C.c_int)]),
    "specrel_num_scenes": (C.c_uint32, [_P]),
    "specrel_upload_scene_versioned": (C.c_int, [_P, C.c_uint32, C.POINTER(F.SceneArrays), C.POINTER(BvhVersion)]),
    "specrel_update_scene": (C.c_int, [_P, C.c_uint32, C.POINTER(F.SceneArrays), C.POINTER(BvhVersion),
                                      C.POINTER(Range), C.c_uint32, C.POINTER(Range), C.c_uint32,
                                      C.POINTER(C.c_int)]),
    "specrel_render": (C.c_int, [_P, C.c_uint32, C.POINTER(F.Frame), _U32P, _U32P, _FP, _FP, C.c_uint64]),
    "specrel_render_resident": (C.c_int, [_P, C.c_uint32, C.POINTER(F.Frame), C.c_int]),
    "specrel_render_noalloc": (C.c_int, [_P, C.c_uint32, C.POINTER(F.Frame), C.c_void_p, C.c_void_p, C.c_void_p,
                                         C.c_void_p, C.c_uint64]),
    "specrel_sync": (C.c_int, [_P]),
    "specrel_download": (C.c_int, [_P, _U32P, _U32P, _FP, _FP, C.c_uint64]),
    "specrel_render_batch": (C.c_int, [_P, C.c_uint32, C.POINTER(F.Frame), C.c_uint32, C.POINTER(ViewOut)]),
    "specrel_render_batch_resident": (C.c_int, [_P, C.c_uint32, C.POINTER(F.Frame), C.c_uint32, C.c_int]),
    "specrel_download_view": (C.c_int, [_P, C.c_uint32, _U32P, _U32P, _FP, _FP, C.c_uint64]),
    "specrel_download_linear": (C.c_int, [_P, C.c_uint32, _FP, C.c_uint64]),
    "specrel_autotune": (C.c_int, [_P, C.c_uint32, C.POINTER(F.Frame), C.c_uint32, C.c_uint32, C.POINTER(Tuning)]),
    "specrel_adjoint": (C.c_int, [_P, _FP, C.POINTER(C.c_uint8), C.POINTER(C.c_uint8), C.c_uint64,
                                  C.POINTER(AdjObject), C.c_uint32, C.POINTER(AdjointOut), C.POINTER(C.c_double)]),
    "specrel_adjoint_batch": (C.c_int, [_P, _FP, C.POINTER(C.c_uint8), C.POINTER(C.c_uint8), C.c_uint64,
                                        C.POINTER(AdjObject), C.c_uint32, C.POINTER(AdjointOut),
                                        C.POINTER(C.c_double)]),
    "specrel_last_timings": (C.c_int, [_P, C.POINTER(Timings)]),
    "specrel_last_stats": (C.c_int, [_P, C.POINTER(F.Stats)]),
    "specrel_render_uniforms": (C.c_int, [_P, C.c_uint32, C.c_void_p, _U32P, C.c_uint64]),
    "specrel_last_wave_clock": (C.c_int, [_P, C.c_void_p, C.c_uint64, C.POINTER(C.c_uint64),
                                          C.POINTER(C.c_uint32)]),
}


def hip_lib_path():
    """The in-tree build, or the library SPECREL_HIP_LIB names (another build to compare against)."""
    return os.environ.get("SPECREL_HIP_LIB") or os.path.join(F.LIB_DIR, "libspecrel_hip.so")


def hip_lib():
    """Load libspecrel_hip.so; raises if it was not built (no silent fallback)."""
    global _lib
    if _lib is None:
        path = hip_lib_path()
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is missing: build it with `python -m specrel_amd.build`")
        lib = C.CDLL(path)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def _ptr(a, t):
    return None if a is None else a.ctypes.data_as(t)


class Context:
    """Thin owner of a specrel_ctx (one device)."""

    def __init__(self, device=0):
        lib = hip_lib()
        h = C.c_void_p()
        rc = lib.specrel_create(device, C.byref(h))
        if rc != OK:
            raise SpecrelError(rc, lib.specrel_last_error(None).decode())
        self._h = h
        self.lib = lib
        self._sizes = []  # (w, h) of every view of the last render, in view order

    def close(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self.lib.specrel_destroy(h)
            h.value = None  # (not a new c_void_p: at interpreter exit the module may be gone)

    def __del__(self):
        self.close()

    def _check(self, rc):
        if rc != OK:
            raise SpecrelError(rc, self.lib.specrel_last_error(self._h).decode())
        return rc

    @property
    def handle(self):
        return self._h

    def device_name(self):
        return self.lib.specrel_device_name(self._h).decode()

    def set_option(self, opt, value):
        self._check(self.lib.specrel_set_option(self._h, opt, int(value)))

    def upload(self, idx, scene):
        self._check(self.lib.specrel_upload_scene(self._h, idx, C.byref(scene.arrays)))

    def upload_versioned(self, idx, packed):
        """specrel_upload_scene_versioned of a PackedScene (its arrays and BVH version)."""
        v = BvhVersion(*packed.version)
        self._check(self.lib.specrel_upload_scene_versioned(self._h, idx, C.byref(packed.arrays), C.byref(v)))

    def update(self, idx, packed, node_ranges=None, prim_ranges=None):
        """specrel_update_scene of a PackedScene after an edit (its dirty ranges unless given):
        True when the ranges were patched in, False when the library uploaded it all."""
        nrs = packed.dirty_nodes if node_ranges is None else node_ranges
        prs = packed.dirty_prims if prim_ranges is None else prim_ranges
        nr = (Range * max(1, len(nrs)))(*[Range(int(a), int(c)) for a, c in nrs])
        pr = (Range * max(1, len(prs)))(*[Range(int(a), int(c)) for a, c in prs])
        v = BvhVersion(*packed.version)
        out = C.c_int(0)
        self._check(self.lib.specrel_update_scene(self._h, idx, C.byref(packed.arrays), C.byref(v), nr, len(nrs), pr,
                                                  len(prs), C.byref(out)))
        return out.value == 1

    def upload_file(self, idx, path, verify=True):
        """specrel_upload_scene_file: a scene file written by any producer (scene.Scene.save)."""
        self._check(self.lib.specrel_upload_scene_file(self._h, idx, os.fsencode(path), 1 if verify else 0))

    def extrude_tubes(self, tubes, download=True):
        """specrel_extrude_tubes over a sequence of F.Tube: (geoms (n,20), attrs (n,48), boxes (n,8)
        float32 arrays, per-tube counts), or (None, None, None, counts) with download=False."""
        arr = (F.Tube * max(1, len(tubes)))(*tubes)
        n = C.c_uint64(0)
        counts = np.zeros(max(1, len(tubes)), np.uint64)
        cptr = counts.ctypes.data_as(C.POINTER(C.c_uint64))
        self._check(self.lib.specrel_extrude_tubes(self._h, arr, len(tubes), None, None, None, 0, C.byref(n), cptr))
        if not download:
            return None, None, None, counts[:len(tubes)]
        total = n.value
        g = np.zeros((total, 20), np.float32)
        a = np.zeros((total, 48), np.float32)
        b = np.zeros((total, 8), np.float32)
        self._check(self.lib.specrel_extrude_tubes(
            self._h, arr, len(tubes), g.ctypes.data_as(C.POINTER(F.Geom)), a.ctypes.data_as(C.POINTER(F.Attr)),
            b.ctypes.data_as(_FP), total, C.byref(n), cptr))
        return g, a, b, counts[:len(tubes)]

    def build_bvh(self, boxes=None, n=None):
        """specrel_build_bvh: (nodes (m,12) float32 view of GpuNode rows, order (n,) uint32, stats) for
        boxes (n,8) float32, or for the last extrusion's records when boxes is None (n given)."""
        if boxes is not None:
            b = np.ascontiguousarray(boxes, np.float32)
            n = len(b)
            bp = b.ctypes.data_as(_FP)
        else:
            bp = None
        nodes = np.zeros((max(1, 2 * n), 12), np.float32)
        order = np.zeros(max(1, n), np.uint32)
        m = C.c_uint32(0)
        st = BvhStats()
        self._check(self.lib.specrel_build_bvh(self._h, bp, n, nodes.ctypes.data_as(C.POINTER(F.Node)),
                                               order.ctypes.data_as(_U32P), C.byref(m), C.byref(st)))
        return nodes[:m.value], order[:n], st.as_dict()

    def patch(self, idx, scene, node_ranges=(), prim_ranges=()):
        """specrel_patch_scene: returns True when the ranges were applied, False when the library
        uploaded the whole scene instead."""
        nr = (Range * max(1, len(node_ranges)))(*[Range(int(a), int(c)) for a, c in node_ranges])
        pr = (Range * max(1, len(prim_ranges)))(*[Range(int(a), int(c)) for a, c in prim_ranges])
        full = C.c_int(0)
        self._check(self.lib.specrel_patch_scene(self._h, idx, C.byref(scene.arrays), nr, len(node_ranges), pr,
                                                 len(prim_ranges), C.byref(full)))
        return full.value == 0

    def _outputs(self, w, h, want_hits, rgb=None):
        n = w * h
        rgb = np.zeros(n, np.uint32) if rgb is None else rgb
        hit = np.zeros(n, np.uint32) if want_hits else None
        lam = np.zeros(n, np.float32) if want_hits else None
        tau = np.zeros(n, np.float32) if want_hits else None
        return rgb, hit, lam, tau

    @staticmethod
    def _result(w, h, rgb, hit, lam, tau):
        out = {"rgb": rgb.reshape(h, w)}
        if hit is not None:
            out.update(hit=hit.reshape(h, w), lam=lam.reshape(h, w), tau=tau.reshape(h, w))
        return out

    def render_frame(self, idx, frame, want_hits=True, rgb=None):
        """specrel_render of one view. rgb: optional caller-owned uint32 buffer to render into (the
        framebuffer of GpuRenderer::render); its element count is passed along and checked."""
        w, h = frame.size[0], frame.size[1]
        rgb, hit, lam, tau = self._outputs(w, h, want_hits, rgb)
        self._check(self.lib.specrel_render(self._h, idx, C.byref(frame), _ptr(rgb, _U32P), _ptr(hit, _U32P),
                                            _ptr(lam, _FP), _ptr(tau, _FP), rgb.size))
        self._sizes = [(w, h)]
        return self._result(w, h, rgb, hit, lam, tau)

    def render_uniforms(self, idx, u, rgb=None):
        """specrel_render_uniforms: one view from the reference's GpuUniforms block; the RGB8 image."""
        w, h = u.size[0], u.size[1]
        rgb = np.zeros(w * h, np.uint32) if rgb is None else rgb
        self._check(self.lib.specrel_render_uniforms(self._h, idx, C.byref(u), _ptr(rgb, _U32P), rgb.size))
        self._sizes = [(w, h)]
        return rgb.reshape(h, w)

    def render_resident(self, idx, frame, want_hits=False):
        self._check(self.lib.specrel_render_resident(self._h, idx, C.byref(frame), int(want_hits)))
        self._sizes = [(frame.size[0], frame.size[1])]

    def render_noalloc(self, idx, frame, d_rgb, d_hit=None, d_lambda=None, d_tau=None):
        """specrel_render_noalloc: one view into caller-owned device buffers (device addresses as
        ints), every intermediate in the library's fixed per-GPU scratch area; enqueued only."""
        n = frame.size[0] * frame.size[1]
        self._check(self.lib.specrel_render_noalloc(self._h, idx, C.byref(frame), d_rgb, d_hit, d_lambda, d_tau, n))

    def sync(self):
        self._check(self.lib.specrel_sync(self._h))

    def view_size(self, view=0):
        """(w, h) of view `view` of the last render made through this object."""
        if view >= len(self._sizes):
            raise SpecrelError(EINVAL, "view index out of range")
        return self._sizes[view]

    def download(self, want_hits=False):
        """Outputs of the last single-view render (its size is the context's, not the caller's)."""
        return self.download_view(0, want_hits)

    @staticmethod
    def _frames(frames):
        arr = (F.Frame * len(frames))()
        for i, f in enumerate(frames):
            arr[i] = f
        return arr

    def render_batch(self, idx, frames, want_hits=True):
        """Every view of `frames` in one batched render; returns one dict per view like render_frame."""
        arr = self._frames(frames)
        outs = (ViewOut * len(frames))()
        res = []
        for i, f in enumerate(frames):
            w, h = f.size[0], f.size[1]
            rgb, hit, lam, tau = self._outputs(w, h, want_hits)
            outs[i].rgb = _ptr(rgb, _U32P)
            outs[i].hit = _ptr(hit, _U32P)
            outs[i].lambda_ = _ptr(lam, _FP)
            outs[i].tau = _ptr(tau, _FP)
            outs[i].len = w * h
            res.append(self._result(w, h, rgb, hit, lam, tau))
        self._check(self.lib.specrel_render_batch(self._h, idx, arr, len(frames), outs))
        self._sizes = [(f.size[0], f.size[1]) for f in frames]
        return res

    def render_batch_resident(self, idx, frames, want_hits=False):
        arr = frames if isinstance(frames, C.Array) else self._frames(frames)
        self._check(self.lib.specrel_render_batch_resident(self._h, idx, arr, len(arr), int(want_hits)))
        self._sizes = [(f.size[0], f.size[1]) for f in arr]

    def download_view(self, view, want_hits=False):
        w, h = self.view_size(view)
        rgb, hit, lam, tau = self._outputs(w, h, want_hits)
        self._check(self.lib.specrel_download_view(self._h, view, _ptr(rgb, _U32P), _ptr(hit, _U32P),
                                                   _ptr(lam, _FP), _ptr(tau, _FP), w * h))
        return self._result(w, h, rgb, hit, lam, tau)

    def download_linear(self, view=0):
        """The view's pre-bloom linear HDR image of the last render, (h, w, 3) float32."""
        w, h = self.view_size(view)
        out = np.zeros((h, w, 3), np.float32)
        self._check(self.lib.specrel_download_linear(self._h, view, out.ctypes.data_as(_FP), out.size))
        return out

    def autotune(self, idx, frames, iters=3):
        arr = frames if isinstance(frames, C.Array) else self._frames(frames)
        t = Tuning()
        self._check(self.lib.specrel_autotune(self._h, idx, arr, len(arr), iters, C.byref(t)))
        return t.as_dict()

    def adjoint(self, target, objects=(), mask_in=None, want_mask=False):
        """specrel_adjoint on the last render. target: (h, w, 3) float32 linear HDR; objects:
        (prim_first, prim_count, center[4]) tuples. Returns a dict with loss, d_event (4),
        d_tetrad (4x4), d_objects (n x 3), the pixel-class counts and optionally the mask."""
        tgt = np.ascontiguousarray(target, np.float32)
        h, w = tgt.shape[0], tgt.shape[1]
        objs = (AdjObject * max(1, len(objects)))()
        for i, (first, count, center) in enumerate(objects):
            objs[i].prim_first, objs[i].prim_count = int(first), int(count)
            for k in range(4):
                objs[i].center[k] = center[k]
        mi = None if mask_in is None else np.ascontiguousarray(mask_in, np.uint8)
        mo = np.zeros((h, w), np.uint8) if want_mask else None
        out = AdjointOut()
        dobj = np.zeros((max(1, len(objects)), 3), np.float64)
        u8 = C.POINTER(C.c_uint8)
        self._check(self.lib.specrel_adjoint(self._h, tgt.ctypes.data_as(_FP), _ptr(mi, u8), _ptr(mo, u8), h * w,
                                             objs, len(objects), C.byref(out),
                                             dobj.ctypes.data_as(C.POINTER(C.c_double))))
        res = {"loss": out.loss, "d_event": np.array(out.d_event[:]),
               "d_tetrad": np.array([[out.d_tetrad[a][k] for k in range(4)] for a in range(4)]),
               "d_objects": dobj[:len(objects)].copy()}
        res.update({n: getattr(out, n) for n in ("valid", "miss", "starlight", "border", "silhouette", "edge")})
        if want_mask:
            res["mask"] = mo
        return res

    def adjoint_batch(self, targets, objects=(), masks_in=None, want_mask=False):
        """specrel_adjoint_batch on the last (batched) render: targets and masks are lists with one
        (h, w, 3) / (h, w) array per view. Returns one result dict per view, as adjoint()."""
        tg = np.ascontiguousarray(np.concatenate([np.asarray(t, np.float32).reshape(-1) for t in targets]))
        shapes = [np.asarray(t).shape[:2] for t in targets]
        objs = (AdjObject * max(1, len(objects)))()
        for i, (first, count, center) in enumerate(objects):
            objs[i].prim_first, objs[i].prim_count = int(first), int(count)
            for k in range(4):
                objs[i].center[k] = center[k]
        mi = None if masks_in is None else np.ascontiguousarray(
            np.concatenate([np.asarray(m, np.uint8).reshape(-1) for m in masks_in]))
        npx = sum(h * w for h, w in shapes)
        mo = np.zeros(npx, np.uint8) if want_mask else None
        outs = (AdjointOut * len(targets))()
        dobj = np.zeros((len(targets), max(1, len(objects)), 3), np.float64)
        u8 = C.POINTER(C.c_uint8)
        self._check(self.lib.specrel_adjoint_batch(self._h, tg.ctypes.data_as(_FP), _ptr(mi, u8), _ptr(mo, u8), npx,
                                                   objs, len(objects), outs,
                                                   dobj.ctypes.data_as(C.POINTER(C.c_double))))
        res, off = [], 0
        for v, (h, w) in enumerate(shapes):
            o = outs[v]
            r = {"loss": o.loss, "d_event": np.array(o.d_event[:]),
                 "d_tetrad": np.array([[o.d_tetrad[a][k] for k in range(4)] for a in range(4)]),
                 "d_objects": dobj[v, :len(objects)].copy()}
            r.update({n: getattr(o, n) for n in ("valid", "miss", "starlight", "border", "silhouette", "edge")})
            if want_mask:
                r["mask"] = mo[off:off + h * w].reshape(h, w)
            off += h * w
            res.append(r)
        return res

    def timings(self):
        t = Timings()
        self._check(self.lib.specrel_last_timings(self._h, C.byref(t)))
        return t.as_dict()

    def stats(self):
        s = F.Stats()
        self._check(self.lib.specrel_last_stats(self._h, C.byref(s)))
        return {n: getattr(s, n) for n, _ in F.Stats._fields_ if n != "reserved"}

    def wave_clock(self):
        """The last render's per-wave records (SPECREL_OPT_WAVE_CLOCK) as a structured array, and
        the clock rate in kHz."""
        n, khz = C.c_uint64(0), C.c_uint32(0)
        self._check(self.lib.specrel_last_wave_clock(self._h, None, 0, C.byref(n), C.byref(khz)))
        out = np.zeros(n.value, WAVE_CLOCK_DTYPE)
        if n.value:
            self._check(self.lib.specrel_last_wave_clock(self._h, out.ctypes.data, n.value, C.byref(n),
                                                         C.byref(khz)))
        return out, khz.value


@dataclass
class GpuTimings:  # gpu.rs:411-420
    submit_ms: float = 0.0
    wait_ms: float = 0.0
    kernel_ms: float = 0.0
    readback_ms: float = 0.0


class GpuRenderer:
    """Drop-in for the reference's GpuRenderer (gpu.rs:433-452)."""

    def __init__(self, scenes, device=0):
        try:
            self.ctx = Context(device)
        except SpecrelError as e:
            raise RuntimeError(f"no GPU adapter: {e}") from e
        from .scene import PackedScene
        for i, s in enumerate(scenes):
            if isinstance(s, PackedScene):
                self.ctx.upload_versioned(i, s)
            else:
                self.ctx.upload(i, s)
        self.last_timings = GpuTimings()

    def adapter_name(self):
        return self.ctx.device_name()

    def update_scene(self, scene_idx, scene, node_ranges=None, prim_ranges=None):
        """gpu.rs:704-731. A PackedScene (versioned, edited in place) is patched with its latest
        edit's dirty ranges when the device holds the version before it and the arrays fit, else
        uploaded afresh; a producer's scene with explicit ranges goes through specrel_patch_scene
        (same shape), without them it is uploaded again. Returns whether ranges were patched."""
        from .scene import PackedScene
        if isinstance(scene, PackedScene):
            return self.ctx.update(scene_idx, scene, node_ranges, prim_ranges)
        if node_ranges is None and prim_ranges is None:
            self.ctx.upload(scene_idx, scene)
            return False
        return self.ctx.patch(scene_idx, scene, node_ranges or (), prim_ranges or ())

    def render(self, scene_idx, camera, cfg, dynamics, framebuffer):
        """Renders into `framebuffer` (a contiguous uint32 array of cfg.width * cfg.height pixels,
        written in place); a size mismatch raises SpecrelError(EINVAL) from the library, as the
        reference asserts (gpu.rs:808)."""
        if not (isinstance(framebuffer, np.ndarray) and framebuffer.dtype == np.uint32
                and framebuffer.flags.c_contiguous):
            raise TypeError("framebuffer must be a contiguous uint32 numpy array")
        frame = build_frame(camera, cfg, dynamics)
        self.ctx.render_frame(scene_idx, frame, want_hits=False, rgb=framebuffer.reshape(-1))
        t = self.ctx.timings()
        self.last_timings = GpuTimings(t["submit_ms"], t["wait_ms"], t["kernel_ms"], t["readback_ms"])
