"""Host side of the image adjoint (specrel_adjoint): the observer parameters and their tetrad,
the objects of a scene, the gradient in those parameters, and in-place Lorentz boosts of whole
objects (patched onto the device) for finite differences.

Parameters of an observer: event E = (t, x, y, z), rapidity vector phi and rotation vector w.
Its tetrad is e_a = columns of B(phi) R(w), with B(phi) = exp(K(phi)) the pure boost of rapidity
|phi| along phi and R(w) the rotation of the spatial axes by |w| about w (Rodrigues): e0 is the
boosted time axis, e1..e3 the boosted, rotated spatial axes (the camera looks along e3, up e2),
as Tetrad::boost_local followed by a rotation in the local frame (minkowski.rs:424-596).

Parameters of object j: the rapidity vector d_j of a Lorentz boost of its whole worldtube about
its reference event c_j: X -> c_j + B(d_j)(X - c_j). Records, normals and four-velocities follow
(hyperplane normals and barycentric rows as covectors). At d_j = 0 the gradient is the
derivative with respect to the object's velocity at c_j (rapidity, so dv = dd / gamma^2 along v).
"""
import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import scene as S


def boost_matrix(phi):
    """exp(K(phi)) as a 4x4 matrix on (t, x, y, z); works on complex input (complex step)."""
    phi = np.asarray(phi)
    dt = np.result_type(phi.dtype, np.float64)
    K = np.zeros((4, 4), dt)
    K[0, 1:] = phi
    K[1:, 0] = phi
    r2 = np.sum(phi * phi)
    if abs(r2) < 1e-12:
        c1 = 1.0 + r2 / 6.0 + r2 * r2 / 120.0
        c2 = 0.5 + r2 / 24.0 + r2 * r2 / 720.0
    else:
        r = np.sqrt(r2)
        c1 = np.sinh(r) / r
        c2 = (np.cosh(r) - 1.0) / r2
    return np.eye(4, dtype=dt) + c1 * K + c2 * (K @ K)


def rotation_matrix(w):
    """The rotation by rotation vector w as a 4x4 matrix acting on the spatial axes."""
    w = np.asarray(w)
    dt = np.result_type(w.dtype, np.float64)
    W = np.zeros((4, 4), dt)
    W[1, 2], W[1, 3] = -w[2], w[1]
    W[2, 1], W[2, 3] = w[2], -w[0]
    W[3, 1], W[3, 2] = -w[1], w[0]
    a2 = np.sum(w * w)
    if abs(a2) < 1e-12:
        s1 = 1.0 - a2 / 6.0 + a2 * a2 / 120.0
        s2 = 0.5 - a2 / 24.0 + a2 * a2 / 720.0
    else:
        a = np.sqrt(a2)
        s1 = np.sin(a) / a
        s2 = (1.0 - np.cos(a)) / a2
    return np.eye(4, dtype=dt) + s1 * W + s2 * (W @ W)


def observer_tetrad(phi, w):
    """e[a] = leg a as (t, x, y, z): the columns of B(phi) R(w)."""
    return (boost_matrix(phi) @ rotation_matrix(w)).T


def tetrad_jacobian(phi, w, h=1e-30):
    """d e[a][i] / d (phi, w) by complex step: shape (4, 4, 6)."""
    J = np.zeros((4, 4, 6))
    base = np.concatenate([np.asarray(phi, np.float64), np.asarray(w, np.float64)])
    for p in range(6):
        x = base.astype(np.complex128)
        x[p] += 1j * h
        J[:, :, p] = observer_tetrad(x[:3], x[3:]).imag / h
    return J


@dataclass
class Observer:
    event: np.ndarray
    phi: np.ndarray
    rot: np.ndarray = field(default_factory=lambda: np.zeros(3))
    fov_deg: float = 65.0

    @property
    def theta(self):
        return np.concatenate([self.event, self.phi, self.rot]).astype(np.float64)

    @classmethod
    def from_theta(cls, theta, fov_deg=65.0):
        t = np.asarray(theta, np.float64)
        return cls(t[:4].copy(), t[4:7].copy(), t[7:10].copy(), fov_deg)

    def camera(self):
        cam = S.Camera4D.new(tuple(float(x) for x in self.event), self.fov_deg)
        cam.tetrad = S.Tetrad(observer_tetrad(self.phi, self.rot).astype(np.float32))
        return cam

    def frame(self, w, h, shadows=S.ShadowLorentzInvariant):
        return S.build_frame(self.camera(), S.RenderConfig(w, h, S.RelativisticShaded, shadows, S.PastLightCone))


def scene_objects(scene, limit=None):
    """(prim_first, prim_count, center) of the scene's worldtubes, sorted by prim_first."""
    objs = sorted((f, c, tuple(ctr)) for f, c, ctr, _m in scene.objects())
    return objs if limit is None else objs[:limit]


def observer_gradient(res, obs: Observer):
    """The gradient in (event, phi, rot) from specrel_adjoint's event and tetrad-leg gradients."""
    J = tetrad_jacobian(obs.phi, obs.rot)
    g6 = np.einsum("ak,akp->p", np.asarray(res["d_tetrad"]), J)
    return np.concatenate([np.asarray(res["d_event"]), g6])


def gradient(ctx, obs: Observer, target, objects, mask_in=None, want_mask=False):
    """Adjoint of the last render (which must be obs's frame): the full gradient vector
    [event(4), phi(3), rot(3), d_0(3), d_1(3), ...] and the library's result dict."""
    res = ctx.adjoint(target, objects, mask_in=mask_in, want_mask=want_mask)
    g = np.concatenate([observer_gradient(res, obs), np.asarray(res["d_objects"]).reshape(-1)])
    return g, res


def _ranges(idx):
    idx = np.unique(np.asarray(idx, np.int64))
    if idx.size == 0:
        return []
    cut = np.nonzero(np.diff(idx) != 1)[0] + 1
    starts = np.concatenate([[0], cut])
    ends = np.concatenate([cut, [idx.size]])
    return [(int(idx[a]), int(idx[b - 1] - idx[a] + 1)) for a, b in zip(starts, ends)]


class ObjectBooster:
    """Lorentz-boosts one object's worldtube in the scene's packed arrays (in place) and grows the
    BVH bounds of its leaves and their ancestors to hold it; undo() restores the arrays. Each call
    returns the dirty node and primitive ranges for specrel_patch_scene."""

    def __init__(self, scene):
        self.scene = scene
        nodes = scene.nodes()
        nu = nodes.view(np.uint32)
        nn = nodes.shape[0]
        right, count, left = nu[:, 8].astype(np.int64), nu[:, 9].astype(np.int64), nu[:, 10].astype(np.int64)
        interior = np.nonzero(count == 0)[0]
        self.parent = np.full(nn, -1, np.int64)
        self.parent[left[interior]] = interior
        self.parent[right[interior]] = interior
        leaves = np.nonzero(count > 0)[0]
        cnt = count[leaves]
        first = right[leaves]
        total = int(cnt.sum())
        self.leaf_of_slot = np.empty(scene.num_prims, np.int64)
        offs = np.arange(total) - np.repeat(np.cumsum(cnt) - cnt, cnt)
        self.leaf_of_slot[np.repeat(first, cnt) + offs] = np.repeat(leaves, cnt)
        order = scene.order()
        self.slot_of_input = np.empty(scene.num_prims, np.int64)
        self.slot_of_input[order] = np.arange(scene.num_prims)
        self._saved = None

    def boost(self, prim_first, prim_count, center, rapidity):
        assert self._saved is None, "undo() the previous boost first"
        sc = self.scene
        slots = np.sort(self.slot_of_input[prim_first:prim_first + prim_count])
        geoms, attrs, nodes = sc.geoms(), sc.attrs(), sc.nodes()
        g = geoms[slots].astype(np.float64)
        L = boost_matrix(np.asarray(rapidity, np.float64))
        Li = boost_matrix(-np.asarray(rapidity, np.float64))
        c = np.asarray(center, np.float64)
        v0, N, B = g[:, 0:4], g[:, 4:8], g[:, 8:20].reshape(-1, 3, 4)
        # vertices v0 + e_j with B_i . e_j = delta_ij and N . e_j = 0
        M = np.concatenate([B, N[:, None, :]], axis=1)
        E = np.linalg.inv(M)[:, :, :3]  # columns e_1..e_3
        verts = np.concatenate([v0[:, None, :], v0[:, None, :] + np.transpose(E, (0, 2, 1))], axis=1)
        verts2 = c + (verts - c) @ L.T
        N2 = N @ Li.T
        N2 /= np.linalg.norm(N2, axis=1, keepdims=True)
        B2 = B @ Li.T
        ng = np.concatenate([verts2[:, 0], N2, B2.reshape(-1, 12)], axis=1).astype(np.float32)
        a = attrs[slots].copy()
        a[:, 32:36] = (a[:, 32:36].astype(np.float64) @ L.T).astype(np.float32)
        a[:, 36:40] = (a[:, 36:40].astype(np.float64) @ L.T).astype(np.float32)
        pad = 1e-4 + 1e-6 * np.abs(verts2)
        bmin = (verts2 - pad).min(axis=1).astype(np.float32)
        bmax = (verts2 + pad).max(axis=1).astype(np.float32)
        # dirty nodes: the leaves and every ancestor
        touched = [np.unique(self.leaf_of_slot[slots])]
        cur = touched[0]
        while cur.size:
            p = self.parent[cur]
            cur = np.unique(p[p >= 0])
            touched.append(cur)
        dirty = np.unique(np.concatenate(touched))
        self._saved = (slots, geoms[slots].copy(), attrs[slots].copy(), dirty, nodes[dirty].copy())
        geoms[slots] = ng
        attrs[slots] = a
        mn, mx = nodes[:, 0:4], nodes[:, 4:8]
        leaf = self.leaf_of_slot[slots]
        np.minimum.at(mn, leaf, bmin)
        np.maximum.at(mx, leaf, bmax)
        cur = np.unique(leaf)
        while cur.size:
            p = self.parent[cur]
            keep = p >= 0
            np.minimum.at(mn, p[keep], mn[cur[keep]])
            np.maximum.at(mx, p[keep], mx[cur[keep]])
            cur = np.unique(p[keep])
        return _ranges(dirty), _ranges(slots)

    def undo(self):
        slots, g, a, dirty, nd = self._saved
        self.scene.geoms()[slots] = g
        self.scene.attrs()[slots] = a
        self.scene.nodes()[dirty] = nd
        self._saved = None
        return _ranges(dirty), _ranges(slots)


def fd_gradient(ctx, scene, obs: Observer, w, h, target, objects, mask, params, step=1e-3, slot=0):
    """Central differences of the loss over the fixed pixel set `mask` for the parameter indices
    `params` of [event(4), phi(3), rot(3), d_0(3), ...]; the scene is patched for object
    parameters and restored afterwards."""
    out = {}
    booster = None
    for p in params:
        vals = []
        for sgn in (1.0, -1.0):
            if p < 10:
                th = obs.theta.copy()
                th[p] += sgn * step
                o2 = Observer.from_theta(th, obs.fov_deg)
                ctx.render_resident(slot, o2.frame(w, h))
            else:
                booster = booster or ObjectBooster(scene)
                j, ax = divmod(p - 10, 3)
                r = np.zeros(3)
                r[ax] = sgn * step
                first, count, center = objects[j]
                nr, pr = booster.boost(first, count, center, r)
                assert ctx.patch(slot, scene, nr, pr)
                ctx.render_resident(slot, obs.frame(w, h))
            ctx.sync()
            vals.append(ctx.adjoint(target, objects, mask_in=mask)["loss"])
            if p >= 10:
                nr, pr = booster.undo()
                assert ctx.patch(slot, scene, nr, pr)
        out[p] = (vals[0] - vals[1]) / (2.0 * step)
    return out


def framing_observer(scene, objects, k, distance=10.0, phi=(0.0, 0.0, 0.0)):
    """An observer at lab time 0, `distance` behind object k along -z, turned (rotation vector)
    to look at where its past light cone meets the object: the object's records whose event
    time is about -(distance to them) give that place, found by a few fixed-point steps."""
    first, count, center = objects[k]
    g = scene.geoms()
    slot_of_input = np.empty(scene.num_prims, np.int64)
    slot_of_input[scene.order()] = np.arange(scene.num_prims)
    v0 = g[slot_of_input[first:first + count], 0:4].astype(np.float64)
    P = np.array(center[1:], np.float64) - np.array([0.0, 0.0, distance])
    Q = np.array(center[1:], np.float64)
    for _ in range(4):
        r = np.linalg.norm(Q - P)
        near = np.abs(v0[:, 0] + r) < 2.0
        if near.any():
            Q = v0[near, 1:].mean(axis=0)
    u = (Q - P) / np.linalg.norm(Q - P)
    axis = np.cross([0.0, 0.0, 1.0], u)
    s = np.linalg.norm(axis)
    rot = np.zeros(3) if s < 1e-12 else axis / s * np.arctan2(s, u[2])
    return Observer(np.concatenate([[0.0], P]), np.asarray(phi, np.float64), rot)


def busiest_object(objects, radius=20.0):
    """The index of the object with the most other objects' reference events within `radius`."""
    c = np.array([o[2][1:] for o in objects], np.float64)
    counts = [(np.linalg.norm(c - c[i], axis=1) < radius).sum() for i in range(len(objects))]
    return int(np.argmax(counts))


def adjoint_problem(scene, seed=7, distance=8.0):
    """The measured adjoint set-up on a scene: the initial observer at rest at lab time 0, 8 units
    behind the busiest object and turned toward it; the hidden observer boosted to rapidity 0.8
    along a seeded direction with its event offset by N(0,1)^4 (seeded)."""
    objs = scene_objects(scene)
    obs = framing_observer(scene, objs, busiest_object(objs), distance)
    rng = np.random.default_rng(seed)
    d = rng.normal(size=3)
    hidden = Observer(obs.event + rng.normal(size=4), 0.8 * d / np.linalg.norm(d), obs.rot.copy())
    return obs, hidden, objs


def batch_problem(scene, n_views, seed, w_min=128, w_max=1024, distance=(6.0, 14.0)):
    """Observers for a batched adjoint, each framing an object of the scene (object v mod count,
    at a seeded distance, with a small seeded boost), with a width uniform in [w_min, w_max]
    (height width*5/8), and for each a hidden observer: the event offset by N(0, 0.5)^4 and the
    rapidity by N(0, 0.1)^3. Returns (observers, hidden observers, sizes, objects)."""
    objs = scene_objects(scene)
    rng = np.random.default_rng(seed)
    obs, hidden, sizes = [], [], []
    for v in range(n_views):
        wv = int(rng.integers(w_min, w_max + 1))
        o = framing_observer(scene, objs, v % len(objs), float(rng.uniform(*distance)), phi=rng.normal(size=3) * 0.05)
        obs.append(o)
        hidden.append(Observer(o.event + rng.normal(size=4) * 0.5, o.phi + rng.normal(size=3) * 0.1, o.rot.copy()))
        sizes.append((wv, max(1, (wv * 5) // 8)))
    return obs, hidden, sizes, objs


def batch_targets(ctx, slot, hidden, sizes):
    """Renders the hidden observers as one batch and returns their linear images."""
    frames = [o.frame(w, h) for o, (w, h) in zip(hidden, sizes)]
    ctx.render_batch_resident(slot, frames)
    ctx.sync()
    return [ctx.download_linear(v) for v in range(len(sizes))]


def batch_gradients(ctx, slot, obs, sizes, targets, objects, masks_in=None, want_mask=False):
    """Batched forward render of the observers, then the batched adjoint: one gradient vector
    [event(4), phi(3), rot(3), d_0(3), ...] and result dict per observer."""
    frames = [o.frame(w, h) for o, (w, h) in zip(obs, sizes)]
    ctx.render_batch_resident(slot, frames)
    res = ctx.adjoint_batch(targets, objects, masks_in=masks_in, want_mask=want_mask)
    grads = [np.concatenate([observer_gradient(r, o), np.asarray(r["d_objects"]).reshape(-1)])
             for r, o in zip(res, obs)]
    return grads, res
