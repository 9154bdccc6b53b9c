"""Scenes edited in place as arenas (Bvh4D::begin_edit / write_part / write_nodes, bvh4d.rs:181-215,
as the scene cache of Dynamic Mode uses them, scene_cache.rs:280-336), on the CPU: the versions and
dirty ranges an edit leaves, links moved with the parts, and trees laid out afresh that the oracle
renders exactly as the tree they came from."""
import numpy as np
import pytest

from specrel_amd import scene as S
from oracle import pyoracle as O
from tests import helpers as H

NO_HIT = 0xFFFFFFFF


@pytest.fixture(scope="module")
def twin():
    return S.PackedScene.from_scene(S.Scene.generate("twin"))


def test_versions_and_dirty_ranges(twin):
    a = S.PackedScene.arena(twin)
    b = S.PackedScene.arena(twin)
    assert a.layout >= 1 and b.layout > a.layout and a.edit == 0  # new_layout, bvh4d.rs:57-61
    assert a.num_nodes == 1 and a.num_prims == 0
    assert np.array_equal(a.nodes[0].view(np.uint32), S.NOTHING.view(np.uint32))
    a.begin_edit()
    assert a.version == (a.layout, 1) and a.dirty_nodes == [] and a.dirty_prims == []
    nr, pr = H.subtree(twin, 1)
    a.write_part(twin, nr, pr, 5, 7)
    a.write_nodes(0, [S.node_row(*H.box(a, 5), left=1, right_or_first=5)])
    assert a.dirty_nodes == [(5, len(nr)), (0, 1)] and a.dirty_prims == [(7, len(pr))]
    assert a.num_nodes == 5 + len(nr) and a.num_prims == 7 + len(pr)
    assert np.all(a.nodes[1:5] == 0) and np.all(a.geoms[:7] == 0)  # the holes before them: zeros
    a.begin_edit()  # the next edit starts with nothing dirty
    assert a.edit == 2 and a.dirty_nodes == [] and a.dirty_prims == []
    a.reset_arena()
    assert a.edit == 0 and a.layout > b.layout and a.num_nodes == 1


def test_write_part_moves_links(twin):
    """Every node of a part keeps its shape and box; interior links move by the node offset and
    leaf ranges by the primitive offset (BvhNode4D::moved, bvh4d.rs:35-44); primitives land in
    leaf order, with an identity order over their new slots."""
    a = S.PackedScene.arena(twin)
    a.begin_edit()
    j = H.links(twin, 0)[1]  # the root's right subtree
    nr, pr = H.subtree(twin, j)
    node_at, prim_at = 3, 11
    a.write_part(twin, nr, pr, node_at, prim_at)
    dn, dp = node_at - nr.start, prim_at - pr.start
    for i in nr:
        left, right, count = H.links(twin, i)
        l2, r2, c2 = H.links(a, i + dn)
        assert c2 == count and np.array_equal(a.nodes[i + dn, :8], twin.nodes[i, :8])
        if count:
            assert r2 == right + dp
        else:
            assert (l2, r2) == (left + dn, right + dn)
    assert np.array_equal(a.geoms[prim_at:], twin.geoms[pr.start:pr.stop])
    assert np.array_equal(a.attrs[prim_at:], twin.attrs[pr.start:pr.stop])
    assert np.array_equal(a.order[prim_at:], np.arange(prim_at, prim_at + len(pr)))
    # moved down (a negative offset) wraps the same way as u32 arithmetic
    b = S.PackedScene.arena(twin)
    b.begin_edit()
    b.write_part(twin, nr, pr, 1, 0)
    assert H.links(b, 1)[0] == H.links(twin, j)[0] - j + 1


@pytest.mark.parametrize("name,depth,seed", [("twin", 3, 1), ("twin", 4, 2), ("carousel", 3, 3), ("barn", 4, 4)])
def test_relaid_tree_renders_as_its_source(name, depth, seed):
    """The same tree in other node places (and primitives in their own): the oracle's image, hit
    depths and proper times are bit-identical, and every hit is the same simplex (through the
    source's leaf order)."""
    src = S.PackedScene.from_scene(S.Scene.generate(name))
    arena = H.relaid(src, depth, seed)
    assert arena.edit == 1 and arena.num_nodes == src.num_nodes and arena.num_prims == src.num_prims
    assert sorted(arena.dirty_prims) != [] and sum(c for _, c in arena.dirty_nodes) == src.num_nodes
    placed = [a for a, _ in arena.dirty_prims]  # the parts in the order they were written
    assert placed != sorted(placed)  # shuffled: their nodes lie in another order than in src
    fr = H.frame_for(name, 160, 100)
    x, y = O.render(src, fr), O.render(arena, fr)
    for k in ("rgb", "lam", "tau"):
        assert np.array_equal(x[k].view(np.uint32), y[k].view(np.uint32)), k
    hit = y["hit"] != NO_HIT
    assert np.array_equal(x["hit"] != NO_HIT, hit) and hit.sum() > 1000
    assert np.array_equal(x["hit"][hit], src.order[y["hit"][hit]])


def test_tree_over_nothing_misses_everywhere(twin):
    """An arena's root before anything is written, or after everything is removed, is NOTHING:
    the oracle renders it as the scene with no nodes at all (sky and stars only)."""
    fr = H.frame_for("twin", 96, 60)
    a = S.PackedScene.arena(twin)
    empty = a.copy()
    empty.nodes = np.zeros((0, 12), np.float32)
    x, y = O.render(a, fr), O.render(empty, fr)
    assert np.all(x["hit"] == NO_HIT)
    for k in ("rgb", "hit", "lam", "tau"):
        assert np.array_equal(x[k].view(np.uint32), y[k].view(np.uint32)), k
