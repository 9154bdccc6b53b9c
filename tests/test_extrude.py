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


@pytest.mark.parametrize("name,over", SCENES)
def test_host_tree_api_and_finish_with_tree(name, over):
    # specrel_build_bvh_host over the records' bounds gives the scene's own tree, and a scene
    # finished with that tree equals the generated one
    p = params(name, over)
    direct = S.Scene.generate(name, p)
    d = S.DeferredScene(name, p)
    g, a, b, counts = S.extrude_tubes_host(d.tubes())
    nodes, order = S.build_bvh_host(b)
    assert nodes.tobytes() == direct.nodes().tobytes() and order.tobytes() == direct.order().tobytes()
    d.set_records(g, a, b, counts)
    assert arrays(d.finish_with_tree(nodes, order)) == arrays(direct)


def test_finish_with_tree_checks_the_order():
    d = S.DeferredScene("cube")
    g, a, b, counts = S.extrude_tubes_host(d.tubes())
    nodes, order = S.build_bvh_host(b)
    d.set_records(g, a, b, counts)
    bad = order.copy()
    bad[0] = bad[1]
    with pytest.raises(RuntimeError, match="permutation"):
        d.finish_with_tree(nodes, bad)


def synthetic_boxes(seed, n, kind):
    """Bounds that stress the build's corner cases: clustered centroids with tiny extents
    (midpoint splits and large leaves of equal centroids), or uniform random boxes."""
    rng = np.random.default_rng(seed)
    if kind == "uniform":
        c = rng.uniform(-50, 50, (n, 4)).astype(np.float32)
        h = rng.uniform(0.01, 2, (n, 4)).astype(np.float32)
    else:
        base = rng.uniform(-5, 5, (8, 4)).astype(np.float32)
        c = base[rng.integers(0, 8, n)]
        jitter = rng.choice([0.0, 2e-5, 6e-5], size=(n, 4)).astype(np.float32)
        c = (c + jitter).astype(np.float32)
        h = rng.uniform(0.01, 0.5, (n, 4)).astype(np.float32)
    return np.concatenate([c - h, c + h], axis=1).astype(np.float32)


@pytest.mark.parametrize("kind", ["uniform", "clustered"])
def test_host_tree_of_synthetic_bounds_is_a_valid_bvh(kind):
    b = synthetic_boxes(3, 5000, kind)
    nodes, order = S.build_bvh_host(b)
    assert sorted(order.tolist()) == list(range(len(b)))
    u = nodes.view(np.uint32)
    # every record lies in its leaf's bounds; leaves tile the leaf slots in order
    covered = 0
    for k in range(len(nodes)):
        if u[k, 9] > 0:
            first, cnt = int(u[k, 8]), int(u[k, 9])
            assert first == covered
            covered += cnt
            bx = b[order[first:first + cnt]]
            assert (nodes[k, 0:4] <= bx[:, 0:4]).all() and (bx[:, 4:8] <= nodes[k, 4:8]).all()
    assert covered == len(b)
