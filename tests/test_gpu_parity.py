"""GPU parity: the HIP renderer against the oracle (CPU restatement of render_frame) on the same
packed scene and view. RGB8 and hit index must match bit for bit; lambda and tau within 1e-6
relative (they come out identical). Mirrors the reference's GPU-vs-CPU image tests
(src/gpu.rs:1015-1311) with zero tolerance instead of "<0.5% of pixels off by >4 levels"."""
import numpy as np
import pytest

from specrel_amd import scene as S
from specrel_amd.renderer import DEFAULT_LDS_TOP, DEFAULT_TRACE_WAVES
from oracle import pyoracle as O
from tests.helpers import DEMO_SCENES, HARNESS_CONFIGS, assert_same, frame_for, harness_frame, render_both

pytestmark = pytest.mark.gpu

# gpu_matches_cpu_reference's 9 (mode, shadow, view) configurations (gpu.rs:1268-1278).
CONFIGS = [
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


@pytest.fixture(scope="module")
def twin():
    return S.Scene.generate("twin")


def test_cube_config_bit_exact(gpu_ctx):
    sc = S.Scene.generate("cube")
    fr = frame_for("cube", 320, 200)
    g, o = render_both(gpu_ctx, sc, fr)
    assert (o["hit"] != 0xFFFFFFFF).sum() > 500
    assert_same(g, o, "cube 320x200")


@pytest.mark.parametrize("mode,shadow,view", CONFIGS)
def test_twin_scene_all_modes(gpu_ctx, twin, mode, shadow, view):
    cam = twin.default_camera()
    cam.tetrad.boost_local((0.2, 0.0, 0.6))
    cam.tetrad.rotate_local(0.15, -0.05, 0.0)
    fr = S.build_frame(cam, S.RenderConfig(192, 120, mode, shadow, view))
    g, o = render_both(gpu_ctx, twin, fr)
    assert len(np.unique(o["rgb"])) > 50
    assert_same(g, o, f"twin {mode}/{shadow}/{view}")


@pytest.fixture(scope="module")
def demo_scenes():
    return {n: S.Scene.generate(n) for n in DEMO_SCENES}


@pytest.mark.parametrize("name", DEMO_SCENES)
def test_demo_scenes_under_the_reference_harness(gpu_ctx, demo_scenes, name):
    # gpu_matches_cpu_reference (gpu.rs:1256-1311) over all five demo scenes and its 9
    # configurations, boosted and rotated observer, 192x120 -- with zero tolerance
    sc = demo_scenes[name]
    gpu_ctx.upload(0, sc)
    for c in range(len(HARNESS_CONFIGS)):
        fr = harness_frame(name, c)
        g = gpu_ctx.render_frame(0, fr, want_hits=True)
        o = O.render(sc, fr)
        assert len(np.unique(o["rgb"])) > 20, (name, c)
        assert_same(g, o, f"{name} config {c} {HARNESS_CONFIGS[c]}")


def test_flashes(gpu_ctx, twin):
    # gpu_matches_cpu_with_flashes (gpu.rs:1015-1082)
    cam = twin.default_camera()
    p = cam.position
    at = lambda r: (p[0] - 2.0 * r, p[1], p[2], p[3])
    dyn = S.FrameDynamics(flashes=[S.FlashEvent(at(3.0), 1.2, 2.0, (1.0, 0.3, 0.2)),
                                   S.FlashEvent(at(6.0), 1.5, 3.0, (0.2, 0.6, 1.0))])
    for shadow in (S.ShadowLorentzInvariant, S.ShadowOff):
        fr = S.build_frame(cam, S.RenderConfig(192, 120, S.RelativisticShaded, shadow, S.PastLightCone), dyn)
        g, o = render_both(gpu_ctx, twin, fr)
        assert_same(g, o, f"flashes shadow={shadow}")


def test_decision_age_with_frontier(gpu_ctx):
    # gpu_matches_cpu_decision_age_with_frontier (gpu.rs:1147-1230)
    sc = S.Scene.generate("decision_age")
    cam = S.Camera4D.new((0.0, 0.0, 0.0, -4.0), 65.0)
    dyn = S.FrameDynamics(frontier=((-0.5, 0.2, 0.0, -4.0), 24.0),
                          remote_origins=[(-3.0, 2.0, 0.0, 2.0)])
    for view in (S.PastLightCone, S.CameraSimultaneity, S.ProphecySlice):
        fr = S.build_frame(cam, S.RenderConfig(192, 120, S.DecisionAge, S.ShadowOff, view), dyn)
        g, o = render_both(gpu_ctx, sc, fr)
        assert len(np.unique(o["rgb"])) > 50
        assert_same(g, o, f"decision age view={view}")


def test_large_scene_reduced(gpu_ctx):
    p = S.gen_params("large", n_objects=10, slices=24, n_stars=800)
    sc = S.Scene.generate("large", p)
    cam = S.default_camera("large", p)
    fr = S.build_frame(cam, S.RenderConfig(320, 200))
    g, o = render_both(gpu_ctx, sc, fr)
    assert_same(g, o, "large reduced")


def test_star_culling_is_exact(gpu_ctx, twin):
    from specrel_amd.renderer import OPT_STAR_CULL
    cam = twin.default_camera()
    cam.tetrad.boost_local((0.0, 0.3, 0.9))
    fr = S.build_frame(cam, S.RenderConfig(256, 160))
    gpu_ctx.upload(0, twin)
    a = gpu_ctx.render_frame(0, fr)
    gpu_ctx.set_option(OPT_STAR_CULL, 0)
    try:
        b = gpu_ctx.render_frame(0, fr)
    finally:
        gpu_ctx.set_option(OPT_STAR_CULL, 1)
    assert np.array_equal(a["rgb"], b["rgb"])


def test_traversal_counters_match_oracle(gpu_ctx, twin):
    from oracle import pyoracle as O
    from specrel_amd.renderer import OPT_STATS
    fr = S.build_frame(twin.default_camera(), S.RenderConfig(160, 100))
    gpu_ctx.upload(0, twin)
    gpu_ctx.set_option(OPT_STATS, 1)
    try:
        gpu_ctx.render_frame(0, fr)
        st = gpu_ctx.stats()
    finally:
        gpu_ctx.set_option(OPT_STATS, 0)
    oc = O.render(twin, fr, counts=True)["counts"]
    for k in ("rays", "node_pops", "box_tests", "prim_tests", "hits", "shadow_rays", "shadow_node_pops",
              "shadow_box_tests", "shadow_prim_tests"):
        assert st[k] == oc[k], (k, st[k], oc[k])


@pytest.mark.parametrize("depth", [8, 10, 12, -4, -8, -16])
def test_stack_variants_are_exact(gpu_ctx, twin, depth):
    # LDS stacks (K > 0) and register windows (-N): entries beyond either live in private memory;
    # a 4-entry window sends most rays there.
    from specrel_amd.renderer import OPT_STACK
    cam = twin.default_camera()
    cam.tetrad.boost_local((0.3, 0.0, 0.5))
    fr = S.build_frame(cam, S.RenderConfig(192, 120))
    gpu_ctx.set_option(OPT_STACK, depth)
    try:
        g, o = render_both(gpu_ctx, twin, fr)
    finally:
        gpu_ctx.set_option(OPT_STACK, 16)
    assert_same(g, o, f"lds stack {depth}")


@pytest.mark.parametrize("persistent,queue,in_trace", [(0, 0, 0), (0, 0, 1), (1, 0, 1), (0, 1, 0)])
def test_traversal_variants_are_exact(gpu_ctx, twin, persistent, queue, in_trace):
    # tile-launched vs persistent primary rays; shadow rays traced by the shading pass, queued, or
    # by the primary-ray kernel (with flashes: several queries per pixel)
    from specrel_amd.renderer import (OPT_PERSISTENT, OPT_SHADOW_IN_TRACE, OPT_SHADOW_QUEUE, SHADOWS_IN_SHADE,
                                      SHADOWS_IN_TRACE, SHADOWS_QUEUED)
    cam = twin.default_camera()
    p = cam.position
    dyn = S.FrameDynamics(flashes=[S.FlashEvent((p[0] - 6.0, p[1], p[2], p[3]), 1.2, 2.0, (1.0, 0.3, 0.2))])
    fr = S.build_frame(cam, S.RenderConfig(192, 120), dyn)
    gpu_ctx.set_option(OPT_PERSISTENT, persistent)
    gpu_ctx.set_option(OPT_SHADOW_QUEUE, queue)
    gpu_ctx.set_option(OPT_SHADOW_IN_TRACE, in_trace)
    try:
        g, o = render_both(gpu_ctx, twin, fr)
        path = gpu_ctx.timings()["shadow_path"]
    finally:
        gpu_ctx.set_option(OPT_PERSISTENT, 0)
        gpu_ctx.set_option(OPT_SHADOW_QUEUE, 0)
        gpu_ctx.set_option(OPT_SHADOW_IN_TRACE, 1)
    assert_same(g, o, f"persistent={persistent} queue={queue} in_trace={in_trace}")
    assert path == (SHADOWS_QUEUED if queue else SHADOWS_IN_TRACE if in_trace and not persistent else SHADOWS_IN_SHADE)


@pytest.mark.parametrize("waves,depth", [(1, 16), (2, 16), (1, 8), (2, 8)])
def test_primary_blocks_of_fewer_waves_are_exact(gpu_ctx, twin, waves, depth):
    from specrel_amd.renderer import OPT_STACK, OPT_STATS, OPT_TRACE_WAVES
    cam = twin.default_camera()
    p = cam.position
    dyn = S.FrameDynamics(flashes=[S.FlashEvent((p[0] - 6.0, p[1], p[2], p[3]), 1.2, 2.0, (1.0, 0.3, 0.2))])
    fr = S.build_frame(cam, S.RenderConfig(200, 120), dyn)
    gpu_ctx.set_option(OPT_TRACE_WAVES, waves)
    gpu_ctx.set_option(OPT_STACK, depth)
    try:
        g, o = render_both(gpu_ctx, twin, fr)
        gpu_ctx.set_option(OPT_STATS, 1)
        gpu_ctx.render_frame(0, fr)
        st = gpu_ctx.stats()
    finally:
        gpu_ctx.set_option(OPT_STATS, 0)
        gpu_ctx.set_option(OPT_TRACE_WAVES, DEFAULT_TRACE_WAVES)
        gpu_ctx.set_option(OPT_STACK, 16)
    assert_same(g, o, f"{waves} waves per block, stack {depth}")
    oc = O.render(twin, fr, counts=True)["counts"]
    for k in ("rays", "node_pops", "box_tests", "prim_tests", "shadow_rays", "shadow_node_pops"):
        assert st[k] == oc[k], (k, st[k], oc[k])


def test_counters_with_shadows_in_the_shading_pass(gpu_ctx, twin):
    from oracle import pyoracle as O
    from specrel_amd.renderer import OPT_SHADOW_IN_TRACE, OPT_STATS
    fr = S.build_frame(twin.default_camera(), S.RenderConfig(160, 100))
    gpu_ctx.upload(0, twin)
    gpu_ctx.set_option(OPT_STATS, 1)
    gpu_ctx.set_option(OPT_SHADOW_IN_TRACE, 0)
    try:
        gpu_ctx.render_frame(0, fr)
        st = gpu_ctx.stats()
    finally:
        gpu_ctx.set_option(OPT_STATS, 0)
        gpu_ctx.set_option(OPT_SHADOW_IN_TRACE, 1)
    oc = O.render(twin, fr, counts=True)["counts"]
    for k in ("shadow_rays", "shadow_node_pops", "shadow_box_tests", "shadow_prim_tests"):
        assert st[k] == oc[k], (k, st[k], oc[k])


@pytest.mark.parametrize("shape,waves", [(1, 4), (2, 4), (1, 1), (2, 2)])
def test_wave_shapes_are_exact(gpu_ctx, twin, shape, waves):
    # A wave takes a 16x4 or 4x16 band of its tile instead of an 8x8 quarter: only which pixels
    # share a wave changes, so images and counters are the oracle's.
    from specrel_amd.renderer import OPT_STATS, OPT_TRACE_WAVES, OPT_WAVE_SHAPE
    cam = twin.default_camera()
    cam.tetrad.boost_local((0.1, -0.2, 0.3))
    p = cam.position
    dyn = S.FrameDynamics(flashes=[S.FlashEvent((p[0] - 6.0, p[1], p[2], p[3]), 1.2, 2.0, (1.0, 0.3, 0.2))])
    fr = S.build_frame(cam, S.RenderConfig(200, 124), dyn)
    gpu_ctx.set_option(OPT_WAVE_SHAPE, shape)
    gpu_ctx.set_option(OPT_TRACE_WAVES, waves)
    try:
        g, o = render_both(gpu_ctx, twin, fr)
        gpu_ctx.set_option(OPT_STATS, 1)
        gpu_ctx.render_frame(0, fr)
        st = gpu_ctx.stats()
    finally:
        gpu_ctx.set_option(OPT_STATS, 0)
        gpu_ctx.set_option(OPT_WAVE_SHAPE, 0)
        gpu_ctx.set_option(OPT_TRACE_WAVES, DEFAULT_TRACE_WAVES)
    assert_same(g, o, f"wave shape {shape}, {waves} waves per block")
    oc = O.render(twin, fr, counts=True)["counts"]
    for k in ("rays", "node_pops", "box_tests", "prim_tests", "hits", "shadow_rays", "shadow_node_pops"):
        assert st[k] == oc[k], (k, st[k], oc[k])


@pytest.mark.parametrize("merged", [0, 1])
def test_load_round_variants_are_exact(gpu_ctx, twin, merged):
    # A wave's node lanes and leaf lanes waiting for their loads in one merged round or each in its
    # own: the same tests in the same order, so images and every counter match the oracle.
    from specrel_amd.renderer import OPT_MERGED_LOADS, OPT_STATS
    p = S.gen_params("large", n_objects=12, slices=32, n_stars=300)
    big = S.Scene.generate("large", p)
    cam = twin.default_camera()
    cam.tetrad.boost_local((0.2, 0.1, 0.4))
    q = cam.position
    dyn = S.FrameDynamics(flashes=[S.FlashEvent((q[0] - 6.0, q[1], q[2], q[3]), 1.2, 2.0, (1.0, 0.3, 0.2))])
    cases = [(twin, S.build_frame(cam, S.RenderConfig(192, 120), dyn), "twin"),
             (big, S.build_frame(S.default_camera("large", p), S.RenderConfig(256, 160)), "large 12x32")]
    gpu_ctx.set_option(OPT_MERGED_LOADS, merged)
    try:
        for sc, fr, name in cases:
            g, o = render_both(gpu_ctx, sc, fr)
            assert_same(g, o, f"{name}, merged rounds {merged}")
            gpu_ctx.set_option(OPT_STATS, 1)
            gpu_ctx.render_frame(0, fr)
            st = gpu_ctx.stats()
            gpu_ctx.set_option(OPT_STATS, 0)
            oc = O.render(sc, fr, counts=True)["counts"]
            for k in ("rays", "node_pops", "box_tests", "prim_tests", "hits", "shadow_rays", "shadow_node_pops",
                      "shadow_box_tests", "shadow_prim_tests"):
                assert st[k] == oc[k], (name, k, st[k], oc[k])
    finally:
        gpu_ctx.set_option(OPT_STATS, 0)
        gpu_ctx.set_option(OPT_MERGED_LOADS, 0)


@pytest.mark.parametrize("width,top,stack,waves,want", [
    (4, 0, 16, 1, 2), (4, 0, 8, 4, 2), (4, 0, 16, 2, 2),
    (2, 8, 16, 1, 3), (2, 4, 16, 4, 3), (2, 1, 8, 2, 3), (4, 6, 16, 1, 3)])
def test_node_width_and_lds_top_levels_are_exact(gpu_ctx, twin, width, top, stack, waves, want):
    # BVH node fetch width 4 (both children's box pairs fetched with the node's) and the tree's top
    # levels read from a per-block LDS copy visit the reference's sequence: images equal the
    # oracle's, also for rays traced again without postponement past a small stack limit (5), which
    # then start from the LDS copy's root too, and for shadow rays and flashes.
    from specrel_amd.renderer import OPT_LDS_TOP, OPT_LEAF_POSTPONE, OPT_NODE_WIDTH, OPT_STACK, OPT_TRACE_WAVES
    p = S.gen_params("large", n_objects=12, slices=32, n_stars=300)
    big = S.Scene.generate("large", p)
    cam = twin.default_camera()
    cam.tetrad.boost_local((0.2, 0.1, 0.4))
    q = cam.position
    dyn = S.FrameDynamics(flashes=[S.FlashEvent((q[0] - 6.0, q[1], q[2], q[3]), 1.2, 2.0, (1.0, 0.3, 0.2))])
    cases = [(twin, S.build_frame(cam, S.RenderConfig(192, 120), dyn), "twin"),
             (big, S.build_frame(S.default_camera("large", p), S.RenderConfig(256, 160)), "large 12x32")]
    try:
        gpu_ctx.set_option(OPT_NODE_WIDTH, width)
        gpu_ctx.set_option(OPT_LDS_TOP, top)
        gpu_ctx.set_option(OPT_STACK, stack)
        gpu_ctx.set_option(OPT_TRACE_WAVES, waves)
        for sc, fr, name in cases:
            o = O.render(sc, fr)
            gpu_ctx.upload(0, sc)
            for post in (1, 5):
                gpu_ctx.set_option(OPT_LEAF_POSTPONE, post)
                assert_same(gpu_ctx.render_frame(0, fr, want_hits=True), o, f"{name}, width {width}, top {top}, "
                                                                            f"postpone {post}")
                assert gpu_ctx.timings()["trace_variant"] == want
    finally:
        gpu_ctx.set_option(OPT_NODE_WIDTH, 2)
        gpu_ctx.set_option(OPT_LDS_TOP, DEFAULT_LDS_TOP)
        gpu_ctx.set_option(OPT_LEAF_POSTPONE, 1)
        gpu_ctx.set_option(OPT_STACK, 16)
        gpu_ctx.set_option(OPT_TRACE_WAVES, DEFAULT_TRACE_WAVES)


@pytest.mark.parametrize("waves,stack,shape,top", [(1, 16, 0, 0), (2, 16, 1, 0), (4, 16, 2, 0), (1, 8, 0, 0),
                                                   (4, 8, 0, 0), (2, 16, 0, 6), (1, 16, 0, 5)])
def test_fused_shading_equals_the_shading_pass(gpu_ctx, twin, waves, stack, shape, top):
    # The primary-ray kernel shading its own pixels (its tile's star list culled into the LDS its
    # stacks used) and the separate shading pass give the oracle's image: views with flashes and
    # haze, decision-age views, and a scene with stars, for every block size, wave shape and stack.
    from specrel_amd.renderer import (OPT_FUSED_SHADE, OPT_LDS_TOP, OPT_STACK, OPT_TRACE_WAVES, OPT_WAVE_SHAPE)
    p = S.gen_params("large", n_objects=12, slices=32, n_stars=400)
    big = S.Scene.generate("large", p)
    cam = twin.default_camera()
    cam.tetrad.boost_local((0.2, 0.1, 0.4))
    q = cam.position
    dyn = S.FrameDynamics(flashes=[S.FlashEvent((q[0] - 6.0, q[1], q[2], q[3]), 1.2, 2.0, (1.0, 0.3, 0.2))])
    cases = [(twin, S.build_frame(cam, S.RenderConfig(200, 136), dyn), "twin flashes"),
             (twin, S.build_frame(cam, S.RenderConfig(144, 90, S.DecisionAge, S.ShadowOff, S.PastLightCone)),
              "twin decision age"),
             (big, S.build_frame(S.default_camera("large", p), S.RenderConfig(300, 190)), "large 12x32, 400 stars")]
    try:
        gpu_ctx.set_option(OPT_STACK, stack)
        gpu_ctx.set_option(OPT_TRACE_WAVES, waves)
        gpu_ctx.set_option(OPT_WAVE_SHAPE, shape)
        gpu_ctx.set_option(OPT_LDS_TOP, top)
        for sc, fr, name in cases:
            o = O.render(sc, fr)
            gpu_ctx.upload(0, sc)
            for fused in (1, 0):
                gpu_ctx.set_option(OPT_FUSED_SHADE, fused)
                g = gpu_ctx.render_frame(0, fr, want_hits=True)
                assert_same(g, o, f"{name}, fused {fused}, waves {waves}, stack {stack}, shape {shape}, top {top}")
                if fused:  # no shading pass of its own ran
                    assert gpu_ctx.timings()["shade_ms"] == 0.0, name
    finally:
        gpu_ctx.set_option(OPT_FUSED_SHADE, 1)
        gpu_ctx.set_option(OPT_STACK, 16)
        gpu_ctx.set_option(OPT_TRACE_WAVES, DEFAULT_TRACE_WAVES)
        gpu_ctx.set_option(OPT_WAVE_SHAPE, 0)
        gpu_ctx.set_option(OPT_LDS_TOP, DEFAULT_LDS_TOP)


@pytest.mark.parametrize("waves,stars", [(4, 100), (2, 100), (1, 100), (4, 0)])
def test_tile_order_changes_nothing(gpu_ctx, waves, stars):
    # From the second render of the same views on, the primary-ray tiles start slowest first (the
    # costs the previous render measured; 1,350 tiles, more than four per CU of a 256-CU device):
    # every render stays the oracle's, counters included. With stars the star binning pass sorts
    # the order, without them a kernel of its own on the side stream.
    from specrel_amd.renderer import OPT_STATS, OPT_TILE_ORDER, OPT_TRACE_WAVES
    p = S.gen_params("large", n_objects=16, slices=32, n_stars=stars)
    sc = S.Scene.generate("large", p)
    fr = S.build_frame(S.default_camera("large", p), S.RenderConfig(720, 480))
    o = O.render(sc, fr, counts=True)
    gpu_ctx.upload(0, sc)
    gpu_ctx.set_option(OPT_TRACE_WAVES, waves)
    try:
        for order in (1, 0, 1):
            gpu_ctx.set_option(OPT_TILE_ORDER, order)
            for i in range(3):
                assert_same(gpu_ctx.render_frame(0, fr, want_hits=True), o, f"order {order}, render {i}")
            gpu_ctx.set_option(OPT_STATS, 1)
            gpu_ctx.render_frame(0, fr)
            st = gpu_ctx.stats()
            gpu_ctx.set_option(OPT_STATS, 0)
            for k in ("rays", "node_pops", "box_tests", "prim_tests", "hits", "shadow_rays", "shadow_node_pops"):
                assert st[k] == o["counts"][k], (order, k, st[k], o["counts"][k])
    finally:
        gpu_ctx.set_option(OPT_STATS, 0)
        gpu_ctx.set_option(OPT_TILE_ORDER, 1)
        gpu_ctx.set_option(OPT_TRACE_WAVES, DEFAULT_TRACE_WAVES)


def test_renders_enqueued_ahead(gpu_ctx, twin):
    # Renders enqueued without waiting (more of them than the pinned frame-staging ring has slots)
    # each upload their own observer: the image left in HBM is the last frame's, bit-exact.
    rng = np.random.default_rng(3)
    frames = []
    for i in range(7):
        cam = twin.default_camera()
        cam.tetrad.boost_local(rng.uniform(-0.4, 0.4, 3).astype(np.float32))
        cam.tetrad.rotate_local(*[float(x) for x in rng.uniform(-0.2, 0.2, 3)])
        frames.append(S.build_frame(cam, S.RenderConfig(192, 120)))
    gpu_ctx.upload(0, twin)
    for last in (2, 6):
        for f in frames[:last + 1]:
            gpu_ctx.render_resident(0, f, want_hits=True)
        gpu_ctx.sync()
        assert_same(gpu_ctx.download(want_hits=True), O.render(twin, frames[last]), f"last frame {last}")


@pytest.mark.parametrize("stack,waves", [(16, 4), (8, 1), (-8, 4)])
def test_postponed_leaves_are_exact(gpu_ctx, stack, waves):
    # Closest-hit rays that keep a reached leaf pending and traverse on (the default) and rays that
    # test every leaf when reached give the oracle's image; so do rays traced again without
    # postponement once their stack passes a small limit (2 or 5 entries instead of 62, so that
    # many rays take that path). Instrumented renders test every leaf when reached, so their
    # counters stay the oracle's with the option on.
    from specrel_amd.renderer import OPT_LEAF_POSTPONE, OPT_STACK, OPT_STATS, OPT_TRACE_WAVES
    p = S.gen_params("large", n_objects=16, slices=32, n_stars=100)
    sc = S.Scene.generate("large", p)
    fr = S.build_frame(S.default_camera("large", p), S.RenderConfig(480, 320))
    o = O.render(sc, fr, counts=True)
    gpu_ctx.upload(0, sc)
    try:
        gpu_ctx.set_option(OPT_STACK, stack)
        gpu_ctx.set_option(OPT_TRACE_WAVES, waves)
        for post in (1, 0, 2, 5, 1):
            gpu_ctx.set_option(OPT_LEAF_POSTPONE, post)
            assert_same(gpu_ctx.render_frame(0, fr, want_hits=True), o, f"postpone {post}")
        gpu_ctx.set_option(OPT_STATS, 1)
        gpu_ctx.render_frame(0, fr)
        st = gpu_ctx.stats()
        for k in ("rays", "node_pops", "box_tests", "prim_tests", "hits", "shadow_rays", "shadow_node_pops"):
            assert st[k] == o["counts"][k], (k, st[k], o["counts"][k])
    finally:
        gpu_ctx.set_option(OPT_STATS, 0)
        gpu_ctx.set_option(OPT_LEAF_POSTPONE, 1)
        gpu_ctx.set_option(OPT_STACK, 16)
        gpu_ctx.set_option(OPT_TRACE_WAVES, DEFAULT_TRACE_WAVES)
