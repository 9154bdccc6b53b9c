"""GpuRenderer::update_scene (gpu.rs:704-731) for scenes edited in place (Dynamic Mode's arena,
scene_cache.rs:280-336): when the device holds the BVH version just before the edit and the
grown arrays fit the room it was given, only the ranges the edit wrote are patched in; otherwise
the scene is uploaded afresh. Either way every render equals the oracle's on the same arrays and a
fresh upload of them, bit for bit."""
import numpy as np
import pytest

from specrel_amd import scene as S
from oracle import pyoracle as O
from tests import helpers as H

pytestmark = pytest.mark.gpu

NO_HIT = 0xFFFFFFFF
W, H_ = 160, 100
# Scene slots no other test uploads to: their buffers start at the sizes these tests give them.
SLOT, FULL, SRC, RELAID = 40, 41, 42, 43


@pytest.fixture(scope="module")
def twin():
    return S.PackedScene.from_scene(S.Scene.generate("twin"))


@pytest.fixture(scope="module")
def frames():
    a = H.frame_for("twin", W, H_)
    cam = S.default_camera("twin")
    cam.tetrad.boost_local((0.25, 0.0, 0.3))
    b = S.build_frame(cam, S.RenderConfig(W, H_))
    return a, b


def check(ctx, arena, frames, what, full_slot=FULL):
    """Slot SLOT (as updated) renders as the oracle does on the arena's arrays, and as a fresh upload
    of them (slot full_slot)."""
    print(what, arena.version, arena.num_nodes, arena.num_prims, flush=True)
    ctx.upload(full_slot, arena)
    out = []
    for fr in frames:
        got = ctx.render_frame(SLOT, fr)
        H.assert_same(got, O.render(arena, fr), what)
        fresh = ctx.render_frame(full_slot, fr)
        for k in ("rgb", "hit", "lam", "tau"):
            assert np.array_equal(got[k].view(np.uint32), fresh[k].view(np.uint32)), (what, k)
        out.append(got)
    return out


def hits(renders):
    return sum(int((r["hit"] != NO_HIT).sum()) for r in renders)


def same_images(x, y):
    return all(np.array_equal(a["rgb"], b["rgb"]) for a, b in zip(x, y))


def capacity(n):
    return n + n // 2  # the room an edited scene's arrays get at upload


def test_edits_patch_in_place_or_upload_afresh(gpu_ctx, twin, frames):
    run_edits(gpu_ctx, twin, frames)


def test_edits_with_the_top_levels_in_lds(twin, frames):
    # The same edits traced by kernels that copy the tree's top 8 levels into LDS: the copy's node
    # list and child entries are rebuilt with each update, so every render stays the oracle's. A
    # context of its own, so that its scene slots start with the sizes the edits give them.
    from specrel_amd.renderer import OPT_LDS_TOP, TRACE_LDS_TOP, Context
    ctx = Context(0)
    try:
        ctx.set_option(OPT_LDS_TOP, 8)
        run_edits(ctx, twin, frames)
        assert ctx.timings()["trace_variant"] == TRACE_LDS_TOP
    finally:
        ctx.close()


def run_edits(gpu_ctx, twin, frames):
    parts, _ = H.cut(twin, 4)
    sizes = [tuple(map(len, H.subtree(twin, j))) for j in parts]
    big = [j for j, (n, p) in zip(parts, sizes) if p >= 1000]
    small = [j for j, (n, p) in zip(parts, sizes) if p < 1000]
    assert len(big) >= 4 and len(small) >= 4

    # An empty arena (edit 0): a tree over nothing, every ray misses.
    arena = S.PackedScene.arena(twin)
    gpu_ctx.upload_versioned(SLOT, arena)
    assert hits(check(gpu_ctx, arena, frames, "empty arena")) == 0

    # Edit 1: the large parts and a tree over them. The device holds one node: no room.
    arena.begin_edit()
    roots, end = H.place_parts(arena, twin, big, 1, 0)
    top = H.top_tree(arena, roots, end)
    arena.write_nodes(0, top[:1])
    arena.write_nodes(end, top[1:])
    assert not gpu_ctx.update(SLOT, arena)
    assert hits(check(gpu_ctx, arena, frames, "edit 1 (uploaded)")) > 1000
    cap_nodes, cap_prims = capacity(arena.num_nodes), capacity(arena.num_prims)

    # Edit 2: the small parts appended, a new tree over all of them past the ends (the old one's
    # nodes left as free room): grown within the room, patched.
    arena.begin_edit()
    old_top = (end, len(top) - 1)
    more, end2 = H.place_parts(arena, twin, small, arena.num_nodes, arena.num_prims)
    top2 = H.top_tree(arena, roots + more, end2)
    arena.write_nodes(0, top2[:1])
    arena.write_nodes(end2, top2[1:])
    assert arena.num_nodes <= cap_nodes and arena.num_prims <= cap_prims
    assert gpu_ctx.update(SLOT, arena)
    r2 = check(gpu_ctx, arena, frames, "edit 2 (grown in place)")

    # Edit 3: in place -- parts dropped from the tree, written over the first tree's nodes.
    same = [(a, b) for i, a in enumerate(big) for b in big[i + 1:] if sizes[parts.index(a)] == sizes[parts.index(b)]]
    a, b = same[0]
    ra, rb = roots[big.index(a)], roots[big.index(b)]
    arena.begin_edit()
    keep = [ra, rb] + [r for r in roots[:3] if r not in (ra, rb)] + more[:4]
    top3 = H.top_tree(arena, keep, old_top[0])
    assert len(top3) - 1 <= old_top[1]
    arena.write_nodes(0, top3[:1])
    arena.write_nodes(old_top[0], top3[1:])
    assert gpu_ctx.update(SLOT, arena)
    r3 = check(gpu_ctx, arena, frames, "edit 3 (parts dropped)")
    assert hits(r3) > 0 and not same_images(r2, r3)

    # Edit 4: two parts of the same shape swap places (primitives and nodes rewritten where they
    # lie, the tree above them untouched but for their boxes): the geometry is the same, the
    # slots it lies in are not.
    arena.begin_edit()
    na, pa_ = H.subtree(arena, ra)
    nb, pb_ = H.subtree(arena, rb)
    arena.write_part(twin, *H.subtree(twin, a), nb.start, pb_.start)
    arena.write_part(twin, *H.subtree(twin, b), na.start, pa_.start)
    assert gpu_ctx.update(SLOT, arena)
    r4 = check(gpu_ctx, arena, frames, "edit 4 (parts swapped)")
    assert same_images(r3, r4) and any(not np.array_equal(x["hit"], y["hit"]) for x, y in zip(r3, r4))

    # Edit 5: beyond the room -- a part written far past the primitives' end.
    arena.begin_edit()
    j = big[0]
    nr, pr = H.subtree(twin, j)
    arena.write_part(twin, nr, pr, arena.num_nodes, cap_prims + 16)
    top5 = H.top_tree(arena, keep + [arena.num_nodes - len(nr)], arena.num_nodes)
    at5 = arena.num_nodes
    arena.write_nodes(0, top5[:1])
    arena.write_nodes(at5, top5[1:])
    assert arena.num_prims > cap_prims
    assert not gpu_ctx.update(SLOT, arena)
    check(gpu_ctx, arena, frames, "edit 5 (beyond the room, uploaded)")

    # Edit 6: everything removed (scene_cache.rs:324): a tree over nothing again, patched.
    arena.begin_edit()
    arena.write_nodes(0, [S.NOTHING])
    assert gpu_ctx.update(SLOT, arena)
    assert hits(check(gpu_ctx, arena, frames, "edit 6 (nothing)")) == 0

    # Edit 7: one part only: its root copied over node 0 (scene_cache.rs:325).
    arena.begin_edit()
    arena.write_nodes(0, [arena.nodes[keep[0]]])
    assert gpu_ctx.update(SLOT, arena)
    assert hits(check(gpu_ctx, arena, frames, "edit 7 (one part)")) > 0


def test_versions_that_do_not_follow_upload_afresh(gpu_ctx, twin, frames):
    arena = H.relaid(twin, 3, 5)
    gpu_ctx.upload_versioned(SLOT, arena)
    roots = [H.links(arena, 0)[0], H.links(arena, 0)[1]]

    # An edit skipped: the device's copy is two edits behind.
    arena.begin_edit()
    arena.begin_edit()
    arena.write_nodes(0, [arena.nodes[roots[0]]])
    assert not gpu_ctx.update(SLOT, arena)
    check(gpu_ctx, arena, frames, "version gap")

    # The next edit follows again.
    arena.begin_edit()
    arena.write_nodes(0, H.top_tree(arena, [roots[1], roots[0]], arena.num_nodes)[:1])
    assert gpu_ctx.update(SLOT, arena)
    check(gpu_ctx, arena, frames, "following edit")

    # Another layout (arrays laid out afresh), edit numbers aside: uploaded.
    other = arena.copy()
    other.layout = S.new_layout()
    other.begin_edit()
    other.write_nodes(0, [arena.nodes[roots[1]]])
    assert not gpu_ctx.update(SLOT, other)
    check(gpu_ctx, other, frames, "new layout")


def test_links_outside_the_ranges_are_patched_too(gpu_ctx, twin, frames):
    """A node whose links the edit rewrote but the caller's ranges leave out: the device compares
    every reached node's links with its own, so that node's row is rewritten as well."""
    arena = H.relaid(twin, 3, 6)
    gpu_ctx.upload_versioned(SLOT, arena)
    before = check(gpu_ctx, arena, frames, "relaid")
    left, right, _ = H.links(arena, 0)
    # one interior node rewritten in place: the root's left child takes over its own right
    # child's links, dropping its left subtree (in view) ...
    arena.begin_edit()
    row = arena.nodes[left].copy()
    row.view(np.uint32)[8:11] = arena.nodes[H.links(arena, left)[1]].view(np.uint32)[8:11]
    arena.write_nodes(left, [row])
    # ... but the ranges name only the root
    assert gpu_ctx.update(SLOT, arena, node_ranges=[(0, 1)], prim_ranges=[])
    after = check(gpu_ctx, arena, frames, "incomplete ranges")
    assert not same_images(before, after)
    # the same on the right, with the edit's own ranges
    arena.begin_edit()
    row = arena.nodes[right].copy()
    row.view(np.uint32)[8:11] = arena.nodes[H.links(arena, right)[1]].view(np.uint32)[8:11]
    arena.write_nodes(right, [row])
    assert gpu_ctx.update(SLOT, arena)
    check(gpu_ctx, arena, frames, "complete ranges")


def test_relaid_tree_renders_as_its_source(gpu_ctx, twin, frames):
    """The same tree in other node places renders exactly as the producer's on the device, hit
    for hit (through the producer's leaf order)."""
    arena = H.relaid(twin, 4, 7)
    gpu_ctx.upload(SRC, twin)
    gpu_ctx.upload_versioned(RELAID, arena)
    for fr in frames:
        x, y = gpu_ctx.render_frame(SRC, fr), gpu_ctx.render_frame(RELAID, fr)
        for k in ("rgb", "lam", "tau"):
            assert np.array_equal(x[k].view(np.uint32), y[k].view(np.uint32)), k
        hit = y["hit"] != NO_HIT
        assert np.array_equal(x["hit"] != NO_HIT, hit) and hit.sum() > 1000
        assert np.array_equal(x["hit"][hit], twin.order[y["hit"][hit]])
