"""GPU parity of batched rendering: every view of a batch (uneven sizes, different frames and
modes, one launch per stage, graph replay) is bit-identical to the single-view render of its
frame, which is itself bit-exact against the oracle; a few views are checked against the oracle
directly. No counterpart in the reference (its callers loop over GpuRenderer::render)."""
import numpy as np
import pytest

from specrel_amd import scene as S
from specrel_amd.renderer import DEFAULT_LDS_TOP, DEFAULT_TRACE_WAVES
from specrel_amd.renderer import (OPT_FUSED_SHADE, OPT_GRAPH, OPT_LDS_TOP, OPT_LEAF_POSTPONE, OPT_NODE_WIDTH,
                                  OPT_PERSISTENT,
                                  OPT_SHADOW_IN_TRACE, OPT_SHADOW_QUEUE, OPT_STACK, OPT_STATS, OPT_TRACE_WAVES,
                                  OPT_WAVE_SHAPE)
from oracle import pyoracle as O
from tests.helpers import assert_same

pytestmark = pytest.mark.gpu

SIZES = [(1, 1), (17, 9), (160, 100), (100, 160), (64, 64), (33, 65), (200, 125), (129, 1)]
MODES = [
    (S.RelativisticShaded, S.ShadowLorentzInvariant, S.PastLightCone),
    (S.RelativisticShaded, S.ShadowCameraSlice, S.CameraSimultaneity),
    (S.LookbackIsochrones, S.ShadowOff, S.PastLightCone),
    (S.Simplex4DInspector, S.ShadowOff, S.PastLightCone),
    (S.DecisionAge, S.ShadowOff, S.ProphecySlice),
    (S.RelativisticShaded, S.ShadowLorentzInvariant, S.Frontier),
    (S.RelativisticShaded, S.ShadowOff, S.PastLightCone),
    (S.LookbackIsochrones, S.ShadowOff, S.Frontier),
]


@pytest.fixture(scope="module")
def twin():
    return S.Scene.generate("twin")


def twin_frames(scene, seed):
    """Observers near the twin scene's default camera: jittered events, boosts and rotations."""
    rng = np.random.default_rng(seed)
    base = scene.default_camera()
    cam_flash = base.position
    dyn = S.FrameDynamics(flashes=[S.FlashEvent((cam_flash[0] - 6.0, cam_flash[1], cam_flash[2], cam_flash[3]),
                                                1.2, 2.0, (1.0, 0.3, 0.2))])
    frames = []
    for i, ((w, h), (mode, shadow, view)) in enumerate(zip(SIZES, MODES)):
        cam = scene.default_camera()
        cam.position = tuple(float(p + d) for p, d in zip(base.position, rng.normal(0.0, 0.4, 4)))
        cam.tetrad.boost_local(rng.uniform(-0.5, 0.5, 3).astype(np.float32))
        cam.tetrad.rotate_local(*[float(x) for x in rng.uniform(-0.2, 0.2, 3)])
        frames.append(S.build_frame(cam, S.RenderConfig(w, h, mode, shadow, view), dyn if i == 0 else None))
    return frames


def singles(ctx, frames):
    return [ctx.render_frame(0, f, want_hits=True) for f in frames]


def assert_views_equal(batch, ref, what):
    for i, (b, r) in enumerate(zip(batch, ref)):
        for k in ("rgb", "hit", "lam", "tau"):
            assert b[k].shape == r[k].shape, f"{what} view {i} {k} shape"
            same = (b[k] == r[k]) | (np.isnan(b[k]) & np.isnan(r[k])) if b[k].dtype == np.float32 else b[k] == r[k]
            assert same.all(), f"{what}: view {i} {k} differs at {np.count_nonzero(~same)} pixels"


def test_batch_matches_single_views_and_oracle(gpu_ctx, twin):
    gpu_ctx.upload(0, twin)
    frames = twin_frames(twin, 7)
    ref = singles(gpu_ctx, frames)
    for graph in (0, 1):
        gpu_ctx.set_option(OPT_GRAPH, graph)
        got = gpu_ctx.render_batch(0, frames)
        assert_views_equal(got, ref, f"graph={graph}")
    gpu_ctx.set_option(OPT_GRAPH, 1)
    for i in (0, 2, 6):
        assert_same(got[i], O.render(twin, frames[i]), f"view {i} vs oracle")
    assert any((r["hit"] != 0xFFFFFFFF).sum() > 100 for r in ref)


def test_graph_replay_tracks_new_frames_and_shapes(gpu_ctx, twin):
    gpu_ctx.upload(0, twin)
    gpu_ctx.set_option(OPT_GRAPH, 1)
    for seed in (11, 12, 13):  # same sizes: the captured graph is replayed with new frame contents
        frames = twin_frames(twin, seed)
        assert_views_equal(gpu_ctx.render_batch(0, frames), singles(gpu_ctx, frames), f"seed {seed}")
    frames = twin_frames(twin, 14)[2:6]  # other sizes: captured again
    assert_views_equal(gpu_ctx.render_batch(0, frames), singles(gpu_ctx, frames), "reshaped")
    frames = twin_frames(twin, 15)
    gpu_ctx.render_batch_resident(0, frames, want_hits=True)
    gpu_ctx.sync()
    got = {i: gpu_ctx.download_view(i, want_hits=True) for i in (1, 6)}
    for i, d in got.items():  # (a single render replaces the batch's resident outputs)
        assert_views_equal([d], [gpu_ctx.render_frame(0, frames[i], want_hits=True)], f"download view {i}")


@pytest.mark.parametrize("stack,persistent,queue,fused", [(8, 1, 0, 1), (16, 0, 1, 1), (16, 1, 1, 1), (-8, 0, 0, 1),
                                                           (16, 0, 0, 2), (8, 0, 0, 2), (16, 0, 0, 0)])
def test_batch_launch_variants(gpu_ctx, twin, stack, persistent, queue, fused):
    # (fused 2: every view of the batch shaded by the primary-ray kernel; 0: single views too by the
    # shading pass)
    gpu_ctx.upload(0, twin)
    frames = twin_frames(twin, 21)
    ref = singles(gpu_ctx, frames)
    try:
        gpu_ctx.set_option(OPT_STACK, stack)
        gpu_ctx.set_option(OPT_PERSISTENT, persistent)
        gpu_ctx.set_option(OPT_SHADOW_QUEUE, queue)
        gpu_ctx.set_option(OPT_FUSED_SHADE, fused)
        what = f"stack={stack} p={persistent} q={queue} fused={fused}"
        assert_views_equal(gpu_ctx.render_batch(0, frames), ref, what)
        assert_views_equal(singles(gpu_ctx, frames), ref, what + ", single views")
    finally:
        gpu_ctx.set_option(OPT_STACK, 16)
        gpu_ctx.set_option(OPT_PERSISTENT, 0)
        gpu_ctx.set_option(OPT_SHADOW_QUEUE, 0)
        gpu_ctx.set_option(OPT_FUSED_SHADE, 1)


def test_star_pass_on_a_second_stream_gives_the_same_views(gpu_ctx, twin):
    from specrel_amd.renderer import OPT_STAR_OVERLAP
    gpu_ctx.upload(0, twin)
    frames = twin_frames(twin, 23)
    ref = singles(gpu_ctx, frames)  # star pass before the primary rays (default)
    try:
        gpu_ctx.set_option(OPT_STAR_OVERLAP, 1)
        assert_views_equal(gpu_ctx.render_batch(0, frames), ref, "star pass on a second stream")
        assert_views_equal(singles(gpu_ctx, frames), ref, "single views, star pass on a second stream")
    finally:
        gpu_ctx.set_option(OPT_STAR_OVERLAP, 0)


def test_batch_stats_are_the_sum_of_single_views(gpu_ctx, twin):
    gpu_ctx.upload(0, twin)
    frames = twin_frames(twin, 31)
    gpu_ctx.set_option(OPT_STATS, 1)
    try:
        total = None
        most = None
        for f in frames:
            gpu_ctx.render_frame(0, f)
            st = gpu_ctx.stats()
            total = st if total is None else {k: total[k] + st[k] for k in st}
            most = st if most is None else {k: max(most[k], st[k]) for k in st}
        gpu_ctx.render_batch(0, frames, want_hits=False)
        got = gpu_ctx.stats()
        distinct = ("distinct_nodes", "distinct_prims", "distinct_hits")
        assert {k: v for k, v in got.items() if k not in distinct} == \
            {k: v for k, v in total.items() if k not in distinct}
        # distinct records: the union over the views, between the largest view's and the sum
        for k in distinct:
            assert most[k] <= got[k] <= total[k], (k, most[k], got[k], total[k])
    finally:
        gpu_ctx.set_option(OPT_STATS, 0)


@pytest.mark.parametrize("fused", [1, 2])
def test_tile_order_over_many_tiles(gpu_ctx, twin, fused):
    # 16,676 tiles: past one block's share (4,096 tiles), the next render's tile order is sorted
    # by blocks of 1,024 tiles in two launches. Renders in that order, and in row order, are the same views
    # with the same counters; also with the views shaded by the primary-ray kernel (fused 2).
    from specrel_amd.renderer import OPT_TILE_ORDER
    gpu_ctx.upload(0, twin)
    rng = np.random.default_rng(51)
    base = twin.default_camera()
    frames = []
    for w, h in [(1280, 800), (1280, 800), (1280, 800), (1300, 900), (17, 9)]:
        cam = twin.default_camera()
        cam.position = tuple(float(p + d) for p, d in zip(base.position, rng.normal(0.0, 0.4, 4)))
        cam.tetrad.boost_local(rng.uniform(-0.5, 0.5, 3).astype(np.float32))
        frames.append(S.build_frame(cam, S.RenderConfig(w, h)))
    try:
        gpu_ctx.set_option(OPT_FUSED_SHADE, fused)
        gpu_ctx.set_option(OPT_TILE_ORDER, 0)
        ref = gpu_ctx.render_batch(0, frames)
        assert_views_equal([ref[0], ref[4]], [gpu_ctx.render_frame(0, frames[i], want_hits=True) for i in (0, 4)],
                           "row order vs single views")
        stats = {}
        for order in (0, 1):
            gpu_ctx.set_option(OPT_TILE_ORDER, order)
            for i in range(3):
                assert_views_equal(gpu_ctx.render_batch(0, frames), ref, f"order {order}, render {i}")
            gpu_ctx.set_option(OPT_STATS, 1)
            gpu_ctx.render_batch(0, frames, want_hits=False)
            stats[order] = gpu_ctx.stats()
            gpu_ctx.set_option(OPT_STATS, 0)
        assert stats[0] == stats[1] and stats[0]["rays"] == sum(w * h for w, h in (f.size for f in frames))
    finally:
        gpu_ctx.set_option(OPT_STATS, 0)
        gpu_ctx.set_option(OPT_TILE_ORDER, 1)
        gpu_ctx.set_option(OPT_FUSED_SHADE, 1)


def test_autotune_picks_a_valid_configuration(gpu_ctx, twin):
    gpu_ctx.upload(0, twin)
    frames = twin_frames(twin, 41)
    ref = singles(gpu_ctx, frames)
    try:
        t = gpu_ctx.autotune(0, frames, iters=2)
        assert t["stack"] in (8, 16) and t["persistent"] in (0, 1) and t["shadow_queue"] in (0, 1)
        assert t["wave_shape"] in (0, 1, 2) and not (t["persistent"] and t["wave_shape"])
        assert t["leaf_postpone"] in (0, 1)
        # 48 + the current options + the postponement flip, + the two other block sizes when the
        # pick is tiled with an LDS stack, + the other node fetch width and 4, 6 or 8 top levels
        # in LDS, + the shading fused into the primary rays of every view
        tiled = t["stack"] > 0 and not t["persistent"]
        assert t["trace_waves"] in ((1, 2, 4) if tiled else (4,))
        assert t["node_width"] in (2, 4) and t["lds_top"] in (0, 4, 6, 8) and t["fused_shade"] in (1, 2)
        assert t["candidates"] == (57 if tiled else 50) and 0 < t["best_ms"] <= t["default_ms"]
        assert_views_equal(gpu_ctx.render_batch(0, frames), ref, "autotuned")
    finally:
        gpu_ctx.set_option(OPT_NODE_WIDTH, 2)
        gpu_ctx.set_option(OPT_LDS_TOP, DEFAULT_LDS_TOP)
        gpu_ctx.set_option(OPT_FUSED_SHADE, 1)
        gpu_ctx.set_option(OPT_STACK, 16)
        gpu_ctx.set_option(OPT_PERSISTENT, 0)
        gpu_ctx.set_option(OPT_SHADOW_QUEUE, 0)
        gpu_ctx.set_option(OPT_GRAPH, 1)
        gpu_ctx.set_option(OPT_SHADOW_IN_TRACE, 1)
        gpu_ctx.set_option(OPT_WAVE_SHAPE, 0)
        gpu_ctx.set_option(OPT_LEAF_POSTPONE, 1)
        gpu_ctx.set_option(OPT_TRACE_WAVES, DEFAULT_TRACE_WAVES)


def test_large_scene_random_observers(gpu_ctx):
    p = S.gen_params("large", n_objects=8, slices=16, n_stars=500)
    sc = S.Scene.generate("large", p)
    gpu_ctx.upload(0, sc)
    obs = S.batch_observers(S.scene_box(sc), 6, seed=5, w_min=16, w_max=96)
    frames = [S.build_frame(c, r) for c, r in obs]
    assert_views_equal(gpu_ctx.render_batch(0, frames), singles(gpu_ctx, frames), "large random observers")
    assert_same(gpu_ctx.render_frame(0, frames[0], want_hits=True), O.render(sc, frames[0]), "view 0 vs oracle")


def test_batch_argument_errors(gpu_ctx, twin):
    from specrel_amd.renderer import SpecrelError
    gpu_ctx.upload(0, twin)
    with pytest.raises(SpecrelError):
        gpu_ctx.render_batch(0, [])
    f = twin_frames(twin, 1)[2]
    with pytest.raises(SpecrelError):
        gpu_ctx.render_batch(5, [f])
    gpu_ctx.render_batch(0, [f], want_hits=False)
    with pytest.raises(SpecrelError):
        gpu_ctx.download_view(1)
