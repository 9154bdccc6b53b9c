"""The scene producer (libspecrel_scene.so) against the reference's physics tests
(src/minkowski.rs:615-691, src/mesh4d.rs:1095-1322) and structural invariants of the packed
BVH (src/bvh4d.rs). CPU only."""
import ctypes as C

import numpy as np
import pytest

from oracle import pyoracle as O
from specrel_amd import _ffi as F
from specrel_amd import scene as S

ETA = np.diag([-1.0, 1.0, 1.0, 1.0])


def mdot(a, b):
    return float(np.asarray(a, np.float64) @ ETA @ np.asarray(b, np.float64))


def test_tetrad_orthonormal_null_rays_and_inverse():
    t = S.Tetrad.from_beta((0.6, -0.3, 0.5))
    t.rotate_local(0.4, -0.25, 0.1)
    t.boost_local((0.2, 0.1, 0.8))
    e = t.e.astype(np.float64)
    assert abs(mdot(e[0], e[0]) + 1.0) < 1e-5
    for a in range(1, 4):
        assert abs(mdot(e[a], e[a]) - 1.0) < 1e-5
        assert abs(mdot(e[0], e[a])) < 1e-5
        for b in range(a + 1, 4):
            assert abs(mdot(e[a], e[b])) < 1e-5
    vl = np.array([2.3, -1.1, 0.7, 3.4])
    vw = vl[0] * e[0] + vl[1] * e[1] + vl[2] * e[2] + vl[3] * e[3]
    rec = np.array([-mdot(vw, e[0]), mdot(vw, e[1]), mdot(vw, e[2]), mdot(vw, e[3])])
    assert np.all(np.abs(rec - vl) < 1e-5)
    n = np.array([0.3, -0.4, 1.0])
    n /= np.linalg.norm(n)
    k = -e[0] + e[1] * n[0] + e[2] * n[1] + e[3] * n[2]
    assert abs(mdot(k, k)) < 1e-5 and k[0] < 0


def test_relativistic_aberration():
    beta = 0.75
    t = S.Tetrad.from_beta((0.0, 0.0, beta))
    e = t.e.astype(np.float64)
    for deg in (0.0, 30.0, 60.0, 90.0, 120.0, 150.0, 180.0):
        th = np.radians(deg)
        k = np.array([-1.0, np.sin(th), 0.0, np.cos(th)])
        vp = k + e[0] * mdot(k, e[0])
        cosp = mdot(vp, e[3]) / np.sqrt(mdot(vp, vp))
        expected = (np.cos(th) + beta) / (1.0 + beta * np.cos(th))
        assert abs(cosp - expected) < 1e-5, deg


def test_simplex_normal_and_dual_basis():
    rng = np.random.default_rng(7)
    b = S.Builder()
    verts = []
    for _ in range(50):
        v = rng.normal(size=(4, 4)).astype(np.float32)
        if b.add_simplex(verts=v):
            verts.append(v.astype(np.float64))
    sc = b.finish()
    g = sc.geoms().astype(np.float64)
    order = sc.order()
    for slot in range(sc.num_prims):
        v = verts[order[slot]]
        E = v[1:] - v[0]
        N, B = g[slot, 4:8], g[slot, 8:20].reshape(3, 4)
        assert abs(np.linalg.norm(N) - 1.0) < 1e-5
        scale = np.abs(E).max()
        assert np.all(np.abs(E @ N) < 1e-4 * scale)          # N annihilates the tet's span
        assert np.allclose(B @ E.T, np.eye(3), atol=2e-3)    # B_i(E_j) = delta_ij
        assert np.all(np.abs(B @ N) < 1e-4)                  # B_i(N) = 0


def _extrude(mesh_kind, mp, samples, n):
    cap = 200000
    geoms = (F.Geom * cap)()
    attrs = (F.Attr * cap)()
    cnt = C.c_uint32()
    mpa = np.asarray(mp, np.float32)
    rc = F.scene_lib().specrel_extrude(mesh_kind, mpa.ctypes.data_as(C.POINTER(C.c_float)), samples, n, 0, geoms,
                                       attrs, cap, C.byref(cnt))
    assert rc == 0
    m = cnt.value
    G = np.frombuffer(geoms, np.float32).reshape(cap, 20)[:m].copy()
    A = np.frombuffer(attrs, np.float32).reshape(cap, 48)[:m].copy()
    return G, A


def test_length_contraction_in_worldtube():
    s, n = S.worldline(F.WL_INERTIAL, [0, 0, 0, 0.8, 0, 0, -20.0, 20.0, 4])
    G, _ = _extrude(F.MESH_CUBE, [1.0, 1.0, 1.0], s, n)
    xs = []
    for g in G:
        r = O.intersect(g, (0.0, -5.0, 0.1, 0.1), (0.0, 1.0, 0.0, 0.0), 0.0, 100.0)
        if r is not None:
            xs.append(r[2])
    assert abs((max(xs) - min(xs)) - 1.2) < 1e-4


def test_thomas_precession_one_orbit():
    beta, radius = 0.8, 4.0
    gamma = 1.0 / np.sqrt(1.0 - beta * beta)
    period = 2 * np.pi / (beta / radius)
    s, n = S.worldline(F.WL_CIRCLE, [radius, beta, 0.0, float(np.float32(period)), 180])
    angles = []
    for i in (0, n - 1):
        t = S.Tetrad(np.array([[s[i].e[k][j] for j in range(4)] for k in range(4)], np.float32))
        t.boost_to((1.0, 0.0, 0.0, 0.0))
        angles.append(np.arctan2(np.float32(t.e[1, 3]), np.float32(t.e[1, 1])))
    measured = np.float32(angles[0] - angles[1]) % np.float32(2 * np.pi)
    expected = ((gamma - 1.0) * 2 * np.pi) % (2 * np.pi)
    assert abs(measured - expected) < 5e-3


def test_born_rigid_front_back_offset_and_central_tau():
    s, n = S.worldline(F.WL_BORN_RAMP, [0.8, 0.0, 20.0, 100])
    F.scene_lib().specrel_sync_proper_time(s, n, 0.0)
    mid = s[50]
    assert abs(mid.phi[0] - np.arctanh(0.8)) < 2e-3
    assert abs(s[100].phi[0]) < 2e-3
    sub = (F.WlSample * 3)(s[49], s[50], s[51])
    G, A = _extrude(F.MESH_CUBE, [1.0, 1.0, 1.0], sub, 3)
    e = np.array([[mid.e[k][j] for j in range(4)] for k in range(4)], np.float32)
    c = np.array(mid.center, np.float32)
    for xl in (-0.75, -0.25, 0.25, 0.75):
        o = c + e[1] * xl + e[2] * np.float32(0.2) + e[3] * np.float32(-3.0)
        d = e[3].copy()
        best = None
        for i, g in enumerate(G):
            r = O.intersect(g, o, d, 0.0, 2.5)
            if r is not None and (best is None or r[0] < best[1][0]):
                best = (i, r)
        assert best is not None
        a = A[best[0]].reshape(12, 4)
        bary = best[1][5:9]
        uvu, uvv, tau = a[0] @ bary, a[1] @ bary, a[5] @ bary
        gu, gv = a[6] @ bary, a[7] @ bary
        assert abs(a[10, 0] - 2.0) < 1e-4 and abs(a[10, 1] - 2.0) < 1e-4  # uv_world_scale
        central = tau - (uvu - 0.5) * gu - (uvv - 0.5) * gv
        assert abs(central - mid.proper_time) < 1e-3


def test_decided_at_follows_worldline_steps():
    b = S.Builder()
    b.add_tube(F.MESH_CUBE, [0.5, 0.5, 0.5], F.WL_INERTIAL, [0, 0, 0, 0.3, 0, 0, 0.0, 4.0, 2], 0,
               decided_at=[10.0, 20.0, 30.0])
    sc = b.finish()
    dec = sc.attrs()[:, 44:48]
    steps = set()
    for row in dec:
        lo, hi = row.min(), row.max()
        assert (lo, hi) in ((10.0, 20.0), (20.0, 30.0))
        steps.add(lo)
    assert steps == {10.0, 20.0}
    st = S.Builder()
    st.add_tube(F.MESH_CUBE, [0.5, 0.5, 0.5], F.WL_INERTIAL, [0, 0, 0, 0.3, 0, 0, 0.0, 4.0, 2], 0)
    assert np.all(st.finish().attrs()[:, 44:48] == np.float32(-1.0e9))


def test_inertial_proper_time_is_t_over_gamma():
    s, n = S.worldline(F.WL_INERTIAL, [0, 0, 0, 0.6, 0, 0, -10.0, 10.0, 8])
    for i in range(n):
        assert abs(s[i].proper_time - s[i].center[0] * 0.8) < 1e-5


def test_config_tet_counts():
    assert S.Scene.generate("cube").num_prims == 12 * 64 * 3
    p = S.gen_params("large", n_objects=5, slices=7, n_stars=3, accel_fraction=0.0)
    sc = S.Scene.generate("large", p)
    assert sc.num_prims == 5 * 320 * 3 * 7
    assert sc.num_stars == 3


@pytest.mark.parametrize("name", ["cube", "twin", "decision_age"])
def test_bvh_structure(name):
    sc = S.Scene.generate(name)
    nodes = sc.nodes()
    mn, mx = nodes[:, 0:4], nodes[:, 4:8]
    u = nodes[:, 8:12].view(np.uint32)
    rof, cnt, left = u[:, 0], u[:, 1], u[:, 2]
    order = sc.order()
    assert np.array_equal(np.sort(order), np.arange(sc.num_prims, dtype=np.uint32))
    covered = np.zeros(sc.num_prims, np.int32)
    for i in range(len(nodes)):
        if cnt[i] > 0:
            # a leaf holds its own range; more than 3 prims only where every centroid coincides
            # (bvh4d.rs:322-324, 396-405), which axis-aligned cube faces do produce
            covered[rof[i]:rof[i] + cnt[i]] += 1
        else:
            assert left[i] == i + 1 and rof[i] > i
            for c in (left[i], rof[i]):
                assert np.all(mn[c] >= mn[i]) and np.all(mx[c] <= mx[i])
    assert np.all(covered == 1)
    # every leaf's box holds its simplices' vertices (v0 is vertex 0 of each tet)
    g = sc.geoms()
    for i in np.nonzero(cnt)[0][:200]:
        v0 = g[rof[i]:rof[i] + cnt[i], 0:4]
        assert np.all(v0 >= mn[i]) and np.all(v0 <= mx[i])
