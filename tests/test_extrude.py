"""Deferred extrusion on the host: a scene generated with its worldtubes recorded, extruded
through the tube descriptors (specrel_extrude_tubes_host) and finished, equals the scene
generated directly, byte for byte; degenerate prisms, empty tubes, decided_at and the Rindler
check behave as extend_worldtube_simplices (mesh4d.rs:721-854). CPU only; the device version is
checked against this one in test_gpu_extrude.py."""
import ctypes as C

import numpy as np
import pytest

from specrel_amd import _ffi as F
from specrel_amd import scene as S

SCENES = [("twin", None), ("decision_age", None), ("cube", None),
          ("large", dict(n_objects=6, slices=12, n_stars=40)),
          ("accel_clocks", dict(n_objects=6, n_clocks=2, slices=24, n_stars=10))]


def params(name, over):
    return S.gen_params(name, **over) if over else None


def arrays(sc):
    return [sc.nodes().tobytes(), sc.geoms().tobytes(), sc.attrs().tobytes(), sc.order().tobytes(),
            sc.star_words().tobytes(), list(sc.arrays.light_pos)]


@pytest.mark.parametrize("name,over", SCENES)
def test_deferred_host_extrusion_equals_direct_generation(name, over):
    p = params(name, over)
    direct = S.Scene.generate(name, p)
    d = S.DeferredScene(name, p)
    tubes = d.tubes()
    assert len(tubes) == len(direct.objects()) > 0
    g, a, b, counts = S.extrude_tubes_host(tubes)
    assert int(counts.sum()) == len(g) == direct.num_prims
    d.set_records(g, a, b, counts)
    got = d.finish()
    assert arrays(got) == arrays(direct)
    assert got.objects() == direct.objects()
    # the builder's own host path gives the same scene
    d2 = S.DeferredScene(name, p)
    d2.extrude_on_host()
    assert arrays(d2.finish()) == arrays(direct)


def test_finish_refuses_pending_tubes():
    d = S.DeferredScene("twin")
    with pytest.raises(RuntimeError, match="waiting for their records"):
        d.finish()


class TubeData:
    """A tube built by hand: keeps its arrays alive next to the descriptor."""

    def __init__(self, verts, tris, uvs, normals, samples, material=3, decided=None):
        self.v = np.asarray(verts, np.float32).ravel()
        self.i = np.asarray(tris, np.uint32).ravel()
        self.uv = np.asarray(uvs, np.float32).ravel()
        self.n = np.asarray(normals, np.float32).ravel()
        self.s, ns = samples
        self.d = None if decided is None else np.asarray(decided, np.float32)
        self.t = F.Tube(self.v.ctypes.data_as(C.POINTER(C.c_float)), self.i.ctypes.data_as(C.POINTER(C.c_uint32)),
                        self.uv.ctypes.data_as(C.POINTER(C.c_float)), self.n.ctypes.data_as(C.POINTER(C.c_float)),
                        C.cast(self.s, C.POINTER(F.WlSample)),
                        None if self.d is None else self.d.ctypes.data_as(C.POINTER(C.c_float)),
                        len(self.v) // 3, len(self.i) // 3, ns, material)


def hand_tubes(steps=6):
    """A good triangle and a degenerate one (its three corners at one point: every tetrahedron
    of its prism has a zero edge, so all three are dropped), an inertial and a hyperbolic
    worldline, one with decided_at, and a tube with a single sample (nothing to extrude)."""
    verts = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (1, 0, 0), (1, 0, 0)]
    tris = [(0, 1, 2), (1, 3, 4)]  # vertices 1, 3, 4 coincide
    uvs = [(0, 0, 1, 0, 0, 1), (1, 0, 1, 0, 0, 1)]
    normals = [(0, 0, 1), (0, 0, 1)]
    inertial = S.worldline(F.WL_INERTIAL, [0, 0, 5, 0.3, 0.1, 0, -4, 4, steps])
    hyper = S.worldline(F.WL_HYPERBOLIC, [2, 0, 0, 1, 0, 0, 0.2, 0.0, 1.5, -5, 5, steps])
    single = S.worldline(F.WL_INERTIAL, [0, 0, 0, 0, 0, 0, 0, 1, 0])
    good = TubeData(verts, tris[:1], uvs[:1], normals[:1], inertial)
    mixed = TubeData(verts, tris, uvs, normals, inertial)
    dec = np.linspace(-3, 3, hyper[1])
    accel = TubeData(verts, tris, uvs, normals, hyper, material=5, decided=dec)
    empty = TubeData(verts, tris, uvs, normals, single)
    return good, mixed, accel, empty


def test_degenerate_prisms_are_dropped_in_order():
    good, mixed, accel, empty = hand_tubes()
    g1, a1, b1, c1 = S.extrude_tubes_host([good.t])
    g2, a2, b2, c2 = S.extrude_tubes_host([mixed.t])
    assert c1[0] == c2[0] == 3 * 6  # the degenerate triangle adds nothing
    assert g1.tobytes() == g2.tobytes() and a1.tobytes() == a2.tobytes() and b1.tobytes() == b2.tobytes()
    g, a, b, c = S.extrude_tubes_host([good.t, empty.t, mixed.t, accel.t])
    assert list(c) == [18, 0, 18, 18]
    assert g[:18].tobytes() == g1.tobytes() and g[18:36].tobytes() == g1.tobytes()
    # decided_at of the accelerated tube: per vertex the value of its sample
    dec = a[36:, 44:48]
    assert set(np.unique(dec)) <= set(accel.d.tolist())
    # material bits in misc[2]
    assert (a[36:, 42].view(np.uint32) == 5).all() and (a[:36, 42].view(np.uint32) == 3).all()
    # records are valid simplices: the hyperplane normal is a unit vector
    n = g[:, 4:8]
    assert np.allclose((n * n).sum(1), 1.0, atol=1e-5)
    # boxes hold their vertices: v0 inside the padded box
    v0 = g[:, 0:4]
    assert ((b[:, 0:4] <= v0) & (v0 <= b[:, 4:8])).all()


def test_rindler_horizon_and_bad_indices_are_refused():
    verts = [(0, 0, 0), (6, 0, 0), (0, 1, 0)]  # 6 units from the centre
    hyper = S.worldline(F.WL_HYPERBOLIC, [2, 0, 0, -1, 0, 0, 0.25, 0.0, 1.0, -2, 2, 4])
    t = TubeData(verts, [(0, 1, 2)], [(0, 0, 1, 0, 0, 1)], [(0, 0, 1)], hyper)
    with pytest.raises(RuntimeError, match="Rindler horizon"):
        S.extrude_tubes_host([t.t])
    bad = TubeData([(0, 0, 0), (1, 0, 0), (0, 1, 0)], [(0, 1, 7)], [(0, 0, 1, 0, 0, 1)], [(0, 0, 1)],
                   S.worldline(F.WL_INERTIAL, [0, 0, 0, 0, 0, 0, 0, 1, 2]))
    with pytest.raises(RuntimeError, match="out of range"):
        S.extrude_tubes_host([bad.t])
