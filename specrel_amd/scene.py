"""Host-side scene, camera and frame objects (the caller side of the renderer boundary).

Mirrors the reference's Rust types that feed `GpuRenderer`: `Scene4D` (src/scene.rs:150-161),
`Camera4D` (src/renderer.rs:45-65), `RenderConfig` (renderer.rs:160-168), `FrameDynamics`
(renderer.rs:205-218) and `Tetrad` (src/minkowski.rs:399-609). The heavy lifting (meshes,
worldlines, extrusion, SAH BVH, packing) runs in libspecrel_scene.so.
"""
import ctypes as C
import os
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

from . import _ffi as F

RelativisticShaded, LookbackIsochrones, Simplex4DInspector, DecisionAge = range(4)
ShadowOff, ShadowLorentzInvariant, ShadowCameraSlice = range(3)
PastLightCone, CameraSimultaneity, Frontier, ProphecySlice = range(4)
PROPHECY_DELTA_DEFAULT = 3.5


def _check(rc):
    if rc < 0:
        raise RuntimeError(F.scene_lib().specrel_scene_last_error().decode())
    return rc


def _fptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class Tetrad:
    """Orthonormal Lorentz frame e[0..4] (minkowski.rs:399)."""

    def __init__(self, e=None):
        self.e = np.zeros((4, 4), np.float32)
        if e is None:
            F.scene_lib().specrel_tetrad_rest(_fptr(self.e))
        else:
            self.e[:] = np.asarray(e, np.float32).reshape(4, 4)

    @classmethod
    def from_beta(cls, beta):
        t = cls()
        b = np.asarray(beta, np.float32)
        F.scene_lib().specrel_tetrad_from_beta(_fptr(t.e), _fptr(b))
        return t

    def boost_local(self, d_rapidity):
        r = np.asarray(d_rapidity, np.float32)
        F.scene_lib().specrel_tetrad_boost_local(_fptr(self.e), _fptr(r))
        return self

    def rotate_local(self, yaw, pitch, roll):
        F.scene_lib().specrel_tetrad_rotate_local(_fptr(self.e), yaw, pitch, roll)
        return self

    def boost_to(self, u):
        uu = np.asarray(u, np.float32)
        F.scene_lib().specrel_tetrad_boost_to(_fptr(self.e), _fptr(uu))
        return self

    def copy(self):
        return Tetrad(self.e.copy())


@dataclass
class Camera4D:
    position: Sequence[float] = (0.0, 0.0, 0.0, 0.0)
    tetrad: Tetrad = field(default_factory=Tetrad)
    proper_time: float = 0.0
    fov_y_rad: float = float(np.float32(65.0) * np.float32(np.float32(np.pi) / np.float32(180.0)))

    @classmethod
    def new(cls, position, fov_deg=65.0):
        # f32::to_radians multiplies by PI/180 rounded to f32.
        rad = np.float32(fov_deg) * (np.float32(np.pi) / np.float32(180.0))
        return cls(tuple(position), Tetrad(), 0.0, float(np.float32(rad)))

    def to_c(self):
        c = F.Camera()
        for i in range(4):
            c.position[i] = self.position[i]
        for k in range(4):
            for i in range(4):
                c.e[k][i] = float(self.tetrad.e[k, i])
        c.proper_time = self.proper_time
        c.fov_y_rad = self.fov_y_rad
        return c


@dataclass
class RenderConfig:
    width: int
    height: int
    mode: int = RelativisticShaded
    shadow_mode: int = ShadowLorentzInvariant
    slice_mode: int = PastLightCone
    prophecy_delta: float = PROPHECY_DELTA_DEFAULT


@dataclass
class FlashEvent:
    at: Sequence[float]
    width: float
    intensity: float
    color: Sequence[float]


@dataclass
class FrameDynamics:
    frontier: Optional[tuple] = None  # (origin[4], proper_time)
    flashes: List[FlashEvent] = field(default_factory=list)
    frontier_lag: float = 0.0
    remote_origins: List[Sequence[float]] = field(default_factory=list)


def build_frame(camera: Camera4D, cfg: RenderConfig, dyn: Optional[FrameDynamics] = None) -> F.Frame:
    """The per-view block (GpuUniforms::build, gpu.rs:147-212)."""
    cc = camera.to_c()
    rc = F.RenderConfigC(cfg.width, cfg.height, cfg.mode, cfg.shadow_mode, cfg.slice_mode, cfg.prophecy_delta)
    d = F.DynamicsC()
    keep = []
    if dyn is not None:
        if dyn.frontier is not None:
            d.has_frontier = 1
            for i in range(4):
                d.frontier_origin[i] = dyn.frontier[0][i]
            d.frontier_proper_time = dyn.frontier[1]
        d.frontier_lag = dyn.frontier_lag
        if dyn.flashes:
            arr = (F.FlashC * len(dyn.flashes))()
            for i, fl in enumerate(dyn.flashes):
                for k in range(4):
                    arr[i].at[k] = fl.at[k]
                arr[i].width = fl.width
                arr[i].intensity = fl.intensity
                for k in range(3):
                    arr[i].color[k] = fl.color[k]
            keep.append(arr)
            d.num_flashes = len(dyn.flashes)
            d.flashes = C.cast(arr, C.POINTER(F.FlashC))
        if dyn.remote_origins:
            ro = (F.F4 * len(dyn.remote_origins))()
            for i, o in enumerate(dyn.remote_origins):
                for k in range(4):
                    ro[i][k] = o[k]
            keep.append(ro)
            d.num_remote = len(dyn.remote_origins)
            d.remote_origins = C.cast(ro, C.POINTER(F.F4))
    out = F.Frame()
    _check(F.scene_lib().specrel_frame_build(C.byref(cc), C.byref(rc), C.byref(d), C.byref(out)))
    return out


class Scene:
    """A packed scene (PackedSceneGpuData, gpu.rs:232-244) owned by the scene library."""

    def __init__(self, handle, name=""):
        self._h = C.c_void_p(handle)
        self.name = name
        self.arrays = F.SceneArrays()
        F.scene_lib().specrel_scene_get_arrays(self._h, C.byref(self.arrays))

    @classmethod
    def generate(cls, name, params: Optional[F.GenParams] = None, **overrides):
        lib = F.scene_lib()
        if params is None and overrides:
            params = gen_params(name, **overrides)
        h = C.c_void_p()
        _check(lib.specrel_scene_generate(name.encode(), C.byref(params) if params is not None else None,
                                          C.byref(h)))
        s = cls(h.value, name)
        s.params = params
        return s

    def default_camera(self):
        return default_camera(self.name, getattr(self, "params", None))

    def __del__(self):
        if getattr(self, "_h", None) and self._h.value:
            F.scene_lib().specrel_scene_free(self._h)
            self._h = C.c_void_p()

    @property
    def num_prims(self):
        return self.arrays.num_prims

    @property
    def num_nodes(self):
        return self.arrays.num_nodes

    @property
    def num_stars(self):
        return self.arrays.num_stars

    def objects(self):
        """Worldtubes of the scene: list of (prim_first, prim_count, center[t,x,y,z], material), with
        primitives as input indices (order[] maps leaf slots to them)."""
        lib = F.scene_lib()
        n = lib.specrel_scene_objects(self._h, None, 0)
        arr = (F.Object * max(n, 1))()
        lib.specrel_scene_objects(self._h, arr, n)
        return [(o.prim_first, o.prim_count, tuple(o.center), o.material) for o in arr[:n]]

    def _view(self, ptr, n, dtype, width):
        if n == 0:
            return np.zeros((0, width), dtype)
        buf = (C.c_char * (n * width * np.dtype(dtype).itemsize)).from_address(ptr)
        buf._owner = self  # the view keeps the scene (and the memory it points into) alive
        return np.frombuffer(buf, dtype).reshape(n, width)

    def nodes(self):
        return self._view(self.arrays.nodes, self.num_nodes, np.float32, 12)

    def geoms(self):
        return self._view(self.arrays.geoms, self.num_prims, np.float32, 20)

    def attrs(self):
        return self._view(self.arrays.attrs, self.num_prims, np.float32, 48)

    def order(self):
        return self._view(self.arrays.order, self.num_prims, np.uint32, 1)[:, 0]

    def star_words(self):
        return self._view(self.arrays.star_words, self.arrays.num_star_words, np.float32, 4)

    def save(self, path, bvh_layout=0, bvh_edit=0):
        """Writes the versioned scene file (include/specrel_file.h) with the object table."""
        lib = F.scene_lib()
        n = lib.specrel_scene_objects(self._h, None, 0) if self._h.value else 0
        objs = (F.Object * max(n, 1))()
        if n:
            lib.specrel_scene_objects(self._h, objs, n)
        _check(lib.specrel_scene_file_write(os.fsencode(path), C.byref(self.arrays), objs, n, bvh_layout, bvh_edit))

    def star_params(self):
        out = np.zeros((self.num_stars, 2), np.float32)
        if self.num_stars:
            F.scene_lib().specrel_scene_star_params(self._h, _fptr(out), self.num_stars)
        return out


class SceneFile(Scene):
    """A scene read from a scene file: the arrays point into a read-only mapping of the file,
    kept until the object is closed or collected. Usable wherever a Scene is (upload, oracle)."""

    def __init__(self, path, verify=True):
        self.name = os.path.basename(os.fspath(path))
        self.arrays = F.SceneArrays()
        self.info = F.FileInfo()
        h = C.c_void_p()
        _check(F.scene_lib().specrel_scene_file_open(os.fsencode(path), 1 if verify else 0, C.byref(h),
                                                     C.byref(self.arrays), C.byref(self.info)))
        self._f = h

    def close(self):
        if getattr(self, "_f", None) and self._f.value:
            F.scene_lib().specrel_scene_file_close(self._f)
            self._f = C.c_void_p()

    def __del__(self):
        self.close()

    def objects(self):
        n = C.c_uint32(0)
        p = F.scene_lib().specrel_scene_file_objects(self._f, C.byref(n))
        return [(o.prim_first, o.prim_count, tuple(o.center), o.material) for o in p[:n.value]] if n.value else []

    def save(self, path, bvh_layout=None, bvh_edit=None):
        objs = self.objects()
        arr = (F.Object * max(len(objs), 1))()
        for i, (a, c, cen, m) in enumerate(objs):
            arr[i] = F.Object(a, c, (C.c_float * 4)(*cen), m, 0)
        _check(F.scene_lib().specrel_scene_file_write(
            os.fsencode(path), C.byref(self.arrays), arr, len(objs),
            self.info.bvh_layout if bvh_layout is None else bvh_layout,
            self.info.bvh_edit if bvh_edit is None else bvh_edit))

    def star_params(self):
        raise NotImplementedError("a scene file holds packed stars only")


def gen_params(name, **overrides) -> F.GenParams:
    p = F.GenParams()
    _check(F.scene_lib().specrel_gen_defaults(name.encode(), C.byref(p)))
    for k, v in overrides.items():
        setattr(p, k, v)
    return p


def default_camera(name, params=None) -> Camera4D:
    c = F.Camera()
    _check(F.scene_lib().specrel_scene_default_camera(name.encode(), C.byref(params) if params is not None else None,
                                                      C.byref(c)))
    t = Tetrad(np.array([[c.e[k][i] for i in range(4)] for k in range(4)], np.float32))
    return Camera4D(tuple(c.position[i] for i in range(4)), t, c.proper_time, c.fov_y_rad)


_NEXT_LAYOUT = [1]


def new_layout():
    """A fresh BVH layout number (bvh4d.rs:57-61): numbered from 1; 0 stands for none."""
    v = _NEXT_LAYOUT[0]
    _NEXT_LAYOUT[0] += 1
    return v


class PackedScene:
    """The packed arrays of a scene (PackedSceneGpuData, gpu.rs:232-289) in numpy, with the version
    of the BVH they come from and the node and primitive ranges its latest edit wrote (BvhVersion
    and Bvh4D::dirty, bvh4d.rs:47-61, 150-215): what GpuRenderer::update_scene consumes. As an
    arena, it is edited in place the way Bvh4D is -- begin_edit, then write_part (a whole subtree
    and its primitives into free ranges or past the ends, links moved with it) and write_nodes --
    so that a copy of the arrays (the device's) can be patched with the dirty ranges."""

    def __init__(self, nodes, geoms, attrs, order, materials, star_words, num_stars, light, layout=None, edit=0):
        self.nodes = np.ascontiguousarray(nodes, np.float32).reshape(-1, 12)
        self.geoms = np.ascontiguousarray(geoms, np.float32).reshape(-1, 20)
        self.attrs = np.ascontiguousarray(attrs, np.float32).reshape(-1, 48)
        self.order = np.ascontiguousarray(order, np.uint32).reshape(-1)
        self.materials = np.ascontiguousarray(materials, np.float32).reshape(-1, 12)
        self.star_words = np.ascontiguousarray(star_words, np.float32).reshape(-1, 4)
        self.num_stars = int(num_stars)
        self.light = tuple(float(x) for x in light)
        self.layout = new_layout() if layout is None else int(layout)
        self.edit = int(edit)
        self.dirty_nodes, self.dirty_prims = [], []

    @classmethod
    def from_scene(cls, scene, layout=None):
        """A copy of a producer's scene, as a fresh layout (edit 0)."""
        a = scene.arrays
        mats = np.frombuffer((C.c_char * (a.num_materials * 48)).from_address(a.materials), np.float32).copy()
        return cls(scene.nodes().copy(), scene.geoms().copy(), scene.attrs().copy(), scene.order().copy(), mats,
                   scene.star_words().copy(), a.num_stars, tuple(a.light_pos[i] for i in range(4)), layout)

    def copy(self):
        """The same arrays and version (a second, independent copy)."""
        return PackedScene(self.nodes.copy(), self.geoms.copy(), self.attrs.copy(), self.order.copy(),
                           self.materials.copy(), self.star_words.copy(), self.num_stars, self.light, self.layout,
                           self.edit)

    @classmethod
    def arena(cls, scene):
        """An empty arena with `scene`'s materials, stars and light (Bvh4D::reset_arena,
        bvh4d.rs:169-179): a new layout holding only the root, a tree over nothing."""
        p = scene.copy() if isinstance(scene, PackedScene) else cls.from_scene(scene)
        p.reset_arena()
        return p

    def reset_arena(self):
        self.nodes = NOTHING.reshape(1, 12).copy()
        self.geoms = np.zeros((0, 20), np.float32)
        self.attrs = np.zeros((0, 48), np.float32)
        self.order = np.zeros(0, np.uint32)
        self.layout, self.edit = new_layout(), 0
        self.dirty_nodes, self.dirty_prims = [], []

    @property
    def num_nodes(self):
        return len(self.nodes)

    @property
    def num_prims(self):
        return len(self.geoms)

    @property
    def version(self):
        return (self.layout, self.edit)

    @property
    def arrays(self) -> F.SceneArrays:
        """specrel_scene_arrays over the current numpy arrays (valid while they are not replaced)."""
        a = F.SceneArrays()
        a.nodes = self.nodes.ctypes.data
        a.num_nodes = self.num_nodes
        a.num_prims = self.num_prims
        a.geoms = self.geoms.ctypes.data
        a.attrs = self.attrs.ctypes.data
        a.order = self.order.ctypes.data
        a.materials = self.materials.ctypes.data
        a.num_materials = len(self.materials)
        a.num_stars = self.num_stars
        a.star_words = self.star_words.ctypes.data
        a.num_star_words = len(self.star_words)
        for i in range(4):
            a.light_pos[i] = self.light[i]
        self._keep = a
        return a

    def begin_edit(self):
        """Bvh4D::begin_edit (bvh4d.rs:181-186): the next edit in place."""
        self.edit += 1
        self.dirty_nodes, self.dirty_prims = [], []

    @staticmethod
    def _put(arr, at, rows):
        n = at + len(rows)
        if n > len(arr):
            grown = np.zeros((n,) + arr.shape[1:], arr.dtype)
            grown[:len(arr)] = arr
            arr = grown
        arr[at:n] = rows
        return arr

    def write_part(self, src, nodes, prims, node_at, prim_at):
        """Bvh4D::write_part (bvh4d.rs:188-210): src's nodes `nodes` and primitives `prims` (ranges,
        a whole subtree and its primitives) to nodes from node_at and primitives from prim_at,
        links moved with them; the primitives in leaf order."""
        dn = np.uint32((node_at - nodes.start) & 0xFFFFFFFF)
        dp = np.uint32((prim_at - prims.start) & 0xFFFFFFFF)
        moved = src.nodes[nodes.start:nodes.stop].copy()
        u = moved.view(np.uint32)
        leaf = u[:, 9] > 0
        u[leaf, 8] += dp
        u[~leaf, 8] += dn
        u[~leaf, 10] += dn
        self.nodes = self._put(self.nodes, node_at, moved)
        self.geoms = self._put(self.geoms, prim_at, src.geoms[prims.start:prims.stop])
        self.attrs = self._put(self.attrs, prim_at, src.attrs[prims.start:prims.stop])
        self.order = self._put(self.order, prim_at, np.arange(prim_at, prim_at + len(prims), dtype=np.uint32))
        self.dirty_nodes.append((node_at, len(nodes)))
        self.dirty_prims.append((prim_at, len(prims)))

    def write_nodes(self, at, rows):
        """Bvh4D::write_nodes (bvh4d.rs:212-215): nodes (12 floats each, links as they are) from at."""
        rows = np.asarray(rows, np.float32).reshape(-1, 12)
        self.nodes = self._put(self.nodes, at, rows)
        self.dirty_nodes.append((at, len(rows)))


def node_row(lo, hi, left=0, right_or_first=0, count=0):
    """One GpuNode (gpu.rs:44-51) as 12 floats: min[4], max[4], right_or_first, count, left, pad."""
    r = np.zeros(12, np.float32)
    r[0:4], r[4:8] = lo, hi
    r.view(np.uint32)[8:12] = (right_or_first, count, left, 0)
    return r


# BvhNode4D::NOTHING (bvh4d.rs:31): no primitives, an empty box (Aabb4D::empty, mesh4d.rs:71-76).
NOTHING = node_row([np.inf] * 4, [-np.inf] * 4)


class Builder:
    """Incremental scene construction (worldtubes, raw simplices, stars)."""

    def __init__(self):
        self._b = C.c_void_p(F.scene_lib().specrel_builder_new())

    def __del__(self):
        if getattr(self, "_b", None) and self._b.value:
            F.scene_lib().specrel_builder_free(self._b)

    def add_tube(self, mesh_kind, mesh_params, wl_kind, wl_params, material, sync_t=None, decided_at=None):
        mp = np.asarray(mesh_params, np.float32)
        wp = np.asarray(wl_params, np.float64)
        dec = None if decided_at is None else np.asarray(decided_at, np.float32)
        _check(F.scene_lib().specrel_builder_add_tube(
            self._b, mesh_kind, _fptr(mp), wl_kind, wp.ctypes.data_as(C.POINTER(C.c_double)),
            0.0 if sync_t is None else sync_t, 0 if sync_t is None else 1, material,
            None if dec is None else _fptr(dec), 0 if dec is None else len(dec)))
        return self

    def add_simplex(self, verts, uv=None, tri_bary=None, tau=None, grad=None, uvws=(1.0, 1.0), normal=(0, 0, 0, 1),
                    four_vel=(1, 0, 0, 0), decided_at=None, material=0):
        z = lambda x, n: np.zeros(n, np.float32) if x is None else np.asarray(x, np.float32).ravel()
        args = [z(verts, 16), z(uv, 8), z(tri_bary, 12), z(tau, 4), z(grad, 8), z(uvws, 2), z(normal, 4),
                z(four_vel, 4)]
        dec = None if decided_at is None else np.asarray(decided_at, np.float32)
        rc = F.scene_lib().specrel_builder_add_simplex(self._b, *[_fptr(a) for a in args],
                                                       None if dec is None else _fptr(dec), material)
        return _check(rc) == 1

    def add_star(self, pos_or_events, temperature, intensity, beta=(0, 0, 0)):
        ev = np.asarray(pos_or_events, np.float32)
        n = 1 if ev.ndim == 1 else ev.shape[0]
        b = np.asarray(beta, np.float32)
        _check(F.scene_lib().specrel_builder_add_star(self._b, _fptr(ev), n, _fptr(b), temperature, intensity))
        return self

    def set_light(self, x, y, z):
        F.scene_lib().specrel_builder_set_light(self._b, x, y, z)
        return self

    def __len__(self):
        return F.scene_lib().specrel_builder_num_simplices(self._b)

    def finish(self, name="custom") -> Scene:
        h = C.c_void_p()
        _check(F.scene_lib().specrel_builder_finish(self._b, C.byref(h)))
        return Scene(h.value, name)


class DeferredScene:
    """A named scene generated with its worldtubes recorded, not extruded
    (specrel_scene_generate_deferred): tubes() are the descriptors, whose records come back from
    the GPU (specrel_extrude_tubes) or the host (extrude_on_host), then finish() builds the BVH."""

    def __init__(self, name, params: Optional[F.GenParams] = None):
        self.name = name
        self.params = params
        h = C.c_void_p()
        _check(F.scene_lib().specrel_scene_generate_deferred(name.encode(), C.byref(params) if params is not None
                                                             else None, C.byref(h)))
        self._b = h

    def __del__(self):
        if getattr(self, "_b", None) and self._b.value:
            F.scene_lib().specrel_builder_free(self._b)

    def tubes(self):
        lib = F.scene_lib()
        n = lib.specrel_builder_pending_tubes(self._b, None, 0)
        arr = (F.Tube * max(1, n))()
        lib.specrel_builder_pending_tubes(self._b, arr, n)
        return list(arr[:n])

    def set_records(self, geoms, attrs, boxes, counts):
        g = np.ascontiguousarray(geoms, np.float32)
        a = np.ascontiguousarray(attrs, np.float32)
        b = np.ascontiguousarray(boxes, np.float32)
        c = np.ascontiguousarray(counts, np.uint64)
        _check(F.scene_lib().specrel_builder_set_tube_records(
            self._b, g.ctypes.data_as(C.POINTER(F.Geom)), a.ctypes.data_as(C.POINTER(F.Attr)), _fptr(b), len(g),
            c.ctypes.data_as(C.POINTER(C.c_uint64))))

    def finish_with_tree(self, nodes, order) -> Scene:
        """specrel_builder_finish_with_tree: the scene over a tree built elsewhere."""
        nd = np.ascontiguousarray(nodes, np.float32)
        od = np.ascontiguousarray(order, np.uint32)
        h = C.c_void_p()
        _check(F.scene_lib().specrel_builder_finish_with_tree(self._b, nd.ctypes.data_as(C.POINTER(F.Node)), len(nd),
                                                              od.ctypes.data_as(C.POINTER(C.c_uint32)), C.byref(h)))
        s = Scene(h.value, self.name)
        s.params = self.params
        return s

    def extrude_on_host(self):
        _check(F.scene_lib().specrel_builder_extrude_pending(self._b))

    def finish(self) -> Scene:
        h = C.c_void_p()
        _check(F.scene_lib().specrel_builder_finish(self._b, C.byref(h)))
        s = Scene(h.value, self.name)
        s.params = self.params
        return s


def extrude_tubes_host(tubes):
    """specrel_extrude_tubes_host: (geoms (n,20), attrs (n,48), boxes (n,8), per-tube counts)."""
    lib = F.scene_lib()
    arr = (F.Tube * max(1, len(tubes)))(*tubes)
    n = C.c_uint64(0)
    counts = np.zeros(max(1, len(tubes)), np.uint64)
    cptr = counts.ctypes.data_as(C.POINTER(C.c_uint64))
    _check(lib.specrel_extrude_tubes_host(arr, len(tubes), None, None, None, 0, C.byref(n), cptr))
    total = n.value
    g = np.zeros((total, 20), np.float32)
    a = np.zeros((total, 48), np.float32)
    b = np.zeros((total, 8), np.float32)
    _check(lib.specrel_extrude_tubes_host(arr, len(tubes), g.ctypes.data_as(C.POINTER(F.Geom)),
                                          a.ctypes.data_as(C.POINTER(F.Attr)), _fptr(b), total, C.byref(n), cptr))
    return g, a, b, counts[:len(tubes)]


def build_bvh_host(boxes):
    """specrel_build_bvh_host: (nodes (m,12) float32 rows of GpuNode, order (n,) uint32)."""
    b = np.ascontiguousarray(boxes, np.float32)
    n = len(b)
    nodes = np.zeros((max(1, 2 * n), 12), np.float32)
    order = np.zeros(max(1, n), np.uint32)
    m = C.c_uint32(0)
    _check(F.scene_lib().specrel_build_bvh_host(_fptr(b), n, nodes.ctypes.data_as(C.POINTER(F.Node)),
                                                order.ctypes.data_as(C.POINTER(C.c_uint32)), C.byref(m)))
    return nodes[:m.value], order[:n]


def worldline(kind, params, cap=4096):
    out = (F.WlSample * cap)()
    n = C.c_uint32()
    wp = np.asarray(params, np.float64)
    _check(F.scene_lib().specrel_worldline_sample(kind, wp.ctypes.data_as(C.POINTER(C.c_double)), out, cap,
                                                  C.byref(n)))
    return out, n.value


def build_tables():
    out = (C.c_uint32 * (4096 + 98))()
    F.scene_lib().specrel_build_tables(out)
    return np.frombuffer(out, np.uint32).copy()


def scene_box(scene: "Scene"):
    """The root hyperbox of the scene's BVH as (min[t,x,y,z], max[t,x,y,z])."""
    n = scene.nodes()
    if len(n) == 0:
        return np.zeros(4, np.float32), np.zeros(4, np.float32)
    return n[0, 0:4].copy(), n[0, 4:8].copy()


def batch_observers(box, n, seed, w_min=160, w_max=1920, rapidity_max=2.0, modes=(PastLightCone, CameraSimultaneity),
                    shade=RelativisticShaded, shadows=ShadowLorentzInvariant):
    """`n` observers for a batched render: width uniform in [w_min, w_max] with height = width*5/8
    (160x100 .. 1920x1200), events uniform in the box, boost rapidity |phi| uniform in
    [0, rapidity_max] along a uniform random direction (Tetrad::boost_local), a random
    rotate_local (yaw, roll uniform in [-pi, pi), pitch in [-pi/2, pi/2)), and the slice mode
    drawn from `modes` (real camera = PastLightCone, God's eye = CameraSimultaneity).
    numpy's PCG64 with `seed` draws everything; returns (Camera4D, RenderConfig) pairs."""
    rng = np.random.default_rng(seed)
    lo, hi = np.asarray(box[0], np.float64), np.asarray(box[1], np.float64)
    out = []
    for _ in range(n):
        w = int(rng.integers(w_min, w_max + 1))
        h = max(1, (w * 5) // 8)
        ev = lo + (hi - lo) * rng.random(4)
        d = rng.normal(size=3)
        d /= max(np.linalg.norm(d), 1e-12)
        phi = d * rng.uniform(0.0, rapidity_max)
        yaw, roll = rng.uniform(-np.pi, np.pi, 2)
        pitch = rng.uniform(-np.pi / 2, np.pi / 2)
        t = Tetrad().boost_local(phi.astype(np.float32)).rotate_local(float(yaw), float(pitch), float(roll))
        cam = Camera4D.new(tuple(float(x) for x in ev))
        cam.tetrad = t
        mode = modes[int(rng.integers(0, len(modes)))]
        out.append((cam, RenderConfig(w, h, shade, shadows, mode)))
    return out
