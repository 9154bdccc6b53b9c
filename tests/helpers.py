"""Shared helpers for the parity tests: run the HIP renderer and the oracle on one view."""
import numpy as np

from specrel_amd import scene as S
from oracle import pyoracle as O


def render_both(ctx, scene, frame, slot=0, upload=True):
    if upload:
        ctx.upload(slot, scene)
    g = ctx.render_frame(slot, frame, want_hits=True)
    o = O.render(scene, frame)
    return g, o


def assert_same(g, o, what=""):
    """Bit-exact RGB8 and hit index; lambda and tau within 1e-6 relative (here: identical bits)."""
    for k in ("rgb", "hit"):
        bad = np.count_nonzero(g[k] != o[k])
        assert bad == 0, f"{what}: {k} differs at {bad} of {g[k].size} pixels; first at {np.argwhere(g[k] != o[k])[:5].tolist()}"
    for k in ("lam", "tau"):
        a, b = g[k], o[k]
        same = (a == b) | (np.isnan(a) & np.isnan(b))
        if not same.all():
            fin = np.isfinite(a) & np.isfinite(b)
            rel = np.abs(a[fin] - b[fin]) / np.maximum(np.abs(b[fin]), 1e-30)
            assert rel.max(initial=0.0) <= 1e-6 and (~same & ~fin).sum() == 0, f"{what}: {k} max rel {rel.max()}"


def frame_for(scene_name, w, h, mode=S.RelativisticShaded, shadow=S.ShadowLorentzInvariant,
              view=S.PastLightCone, camera=None, dynamics=None, params=None):
    cam = camera or S.default_camera(scene_name, params)
    return S.build_frame(cam, S.RenderConfig(w, h, mode, shadow, view), dynamics)


# gpu_matches_cpu_reference's 9 (mode, shadow, view) configurations (gpu.rs:1268-1278).
HARNESS_CONFIGS = [
    (S.RelativisticShaded, S.ShadowLorentzInvariant, S.PastLightCone),
    (S.RelativisticShaded, S.ShadowCameraSlice, S.CameraSimultaneity),
    (S.LookbackIsochrones, S.ShadowOff, S.PastLightCone),
    (S.Simplex4DInspector, S.ShadowOff, S.PastLightCone),
    (S.DecisionAge, S.ShadowOff, S.PastLightCone),
    (S.DecisionAge, S.ShadowOff, S.ProphecySlice),
    (S.RelativisticShaded, S.ShadowLorentzInvariant, S.Frontier),
    (S.RelativisticShaded, S.ShadowCameraSlice, S.ProphecySlice),
    (S.LookbackIsochrones, S.ShadowOff, S.Frontier),
]
DEMO_SCENES = ["twin", "barn", "penrose", "train", "carousel"]


def harness_frame(scene_name, config, w=192, h=120):
    """The harness's view of a demo scene (gpu.rs:1283-1297): the scene's initial camera, boosted
    by (0.2, 0, 0.6) and rotated by (0.15, -0.05, 0) in its own frame."""
    cam = S.default_camera(scene_name)
    cam.tetrad.boost_local((0.2, 0.0, 0.6))
    cam.tetrad.rotate_local(0.15, -0.05, 0.0)
    mode, shadow, view = HARNESS_CONFIGS[config]
    return S.build_frame(cam, S.RenderConfig(w, h, mode, shadow, view))


# Arenas edited in place (the scene cache of Dynamic Mode, scene_cache.rs:280-336), built from
# the subtrees of a producer's tree.

def links(p, i):
    u = p.nodes.view(np.uint32)[i]
    return int(u[10]), int(u[8]), int(u[9])  # left, right_or_first, count


def subtree(p, i):
    """The node and primitive ranges of the subtree at node i of a built tree (Bvh4D::build lays
    a subtree out contiguously, left child after its parent, primitives in leaf order)."""
    stack, nodes, lo, hi = [i], 0, None, None
    while stack:
        j = stack.pop()
        nodes += 1
        left, right, count = links(p, j)
        if count:
            lo = right if lo is None else min(lo, right)
            hi = right + count if hi is None else max(hi, right + count)
        else:
            stack += [left, right]
    return range(i, i + nodes), range(lo, hi)


def cut(p, depth):
    """The subtrees `depth` levels below the root (leaves above that depth stand alone), left to
    right, and the interior nodes above them in preorder."""
    parts, tops = [], []

    def go(j, d):
        left, right, count = links(p, j)
        if count or d == depth:
            parts.append(j)
            return
        tops.append(j)
        go(left, d + 1)
        go(right, d + 1)

    go(0, 0)
    return parts, tops


def box(p, i):
    return p.nodes[i, 0:4].copy(), p.nodes[i, 4:8].copy()


def top_tree(p, roots, base):
    """Rows of a tree over the subtrees at nodes `roots` of p (split in halves, in the given
    order): row 0 is the root, for node 0, the rest for nodes base, base + 1, ... (top_tree's
    numbering, bvh4d.rs:598-644)."""
    rows = []

    def go(items):
        if len(items) == 1:
            return (items[0],) + box(p, items[0])
        k = len(rows)
        rows.append(None)
        mid = len(items) // 2
        (left, llo, lhi), (right, rlo, rhi) = go(items[:mid]), go(items[mid:])
        lo, hi = np.minimum(llo, rlo), np.maximum(lhi, rhi)
        rows[k] = S.node_row(lo, hi, left=left, right_or_first=right)
        return (0 if k == 0 else base + k - 1), lo, hi

    go(list(roots))
    return rows


def place_parts(arena, src, parts, node_at, prim_at=None):
    """write_part each of src's subtrees `parts` into arena: nodes from node_at on, primitives
    from prim_at on (or at their places in src when None). Returns the arena nodes of their roots
    and the next free node."""
    roots = []
    for j in parts:
        nr, pr = subtree(src, j)
        pa = pr.start if prim_at is None else prim_at
        arena.write_part(src, nr, pr, node_at, pa)
        roots.append(node_at)
        node_at += len(nr)
        if prim_at is not None:
            prim_at += len(pr)
    return roots, node_at


def relaid(src, depth, seed):
    """src's tree in an arena, laid out afresh: the subtrees `depth` levels down written as parts
    in shuffled node places (primitives where src has them), the nodes above rewritten over node
    0 and the free range after the parts, with src's own boxes and shape. Rays meet exactly what
    they meet in src, in the same order."""
    parts, tops = cut(src, depth)
    arena = S.PackedScene.arena(src)
    arena.begin_edit()
    perm = np.random.default_rng(seed).permutation(len(parts))
    at = {}
    node_at = 1
    for k in perm:
        roots, node_at = place_parts(arena, src, [parts[k]], node_at)
        at[parts[k]] = roots[0]
    base = node_at
    top_at = {j: (0 if n == 0 else base + n - 1) for n, j in enumerate(tops)}
    rows = []
    for j in tops:
        left, right, _ = links(src, j)
        where = lambda c: top_at[c] if c in top_at else at[c]
        lo, hi = box(src, j)
        rows.append(S.node_row(lo, hi, left=where(left), right_or_first=where(right)))
    arena.write_nodes(0, rows[:1])
    if len(rows) > 1:
        arena.write_nodes(base, rows[1:])
    return arena
