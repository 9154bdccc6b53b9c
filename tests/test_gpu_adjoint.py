"""GPU adjoint (specrel_adjoint) against the FP64 forward-mode reference (oracle/adjoint_ref.cpp)
and against central differences of the GPU's own f32 loss; incremental scene patching
(specrel_patch_scene) against a full upload. No counterpart in the reference: it has no
gradients, and its update_scene tests are about Dynamic Mode chunks (gpu.rs:704-731)."""
import numpy as np
import pytest

from specrel_amd import adjoint as A
from specrel_amd import scene as S
from specrel_amd.renderer import OPT_ADJOINT_MASK, SpecrelError
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu


def rel_err(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300))


def make_problem(ctx, scene, w, h, seed, obs=None):
    rng = np.random.default_rng(seed)
    if obs is None:
        cam = scene.default_camera()
        obs = A.Observer(np.array(cam.position, np.float64), rng.normal(size=3) * 0.1, rng.normal(size=3) * 0.05)
    tgt = A.Observer(obs.event + rng.normal(size=4) * 0.3, obs.phi + rng.normal(size=3) * 0.1, obs.rot)
    ctx.upload(0, scene)
    ctx.render_resident(0, tgt.frame(w, h))
    target = ctx.download_linear(0)
    ctx.render_resident(0, obs.frame(w, h))
    ctx.sync()
    return obs, target


@pytest.fixture(scope="module")
def twin():
    return S.Scene.generate("twin")


@pytest.mark.parametrize("seed,strict", [(1, 0), (2, 0), (1, 1)])
def test_adjoint_matches_fp64_reference_twin(gpu_ctx, twin, seed, strict):
    w, h = 192, 120
    obs, target = make_problem(gpu_ctx, twin, w, h, seed)
    objs = A.scene_objects(twin)
    gpu_ctx.set_option(OPT_ADJOINT_MASK, strict)
    try:
        g, res = A.gradient(gpu_ctx, obs, target, objs, want_mask=True)
    finally:
        gpu_ctx.set_option(OPT_ADJOINT_MASK, 0)
    ref = O.adjoint_ref(twin, obs.frame(w, h), obs.theta, target, objs, strict=bool(strict))
    for k in ("valid", "miss", "starlight", "border", "silhouette", "edge"):
        assert res[k] == ref["counts"][k], (k, res[k], ref["counts"][k])
    assert np.array_equal(res["mask"], ref["mask"])
    assert res["valid"] > 1000
    hits = w * h - res["miss"] - res["starlight"]
    if not strict:  # silhouettes and the border masked, nothing else of note
        assert res["valid"] >= 0.85 * hits, (res, hits)
    assert abs(res["loss"] - ref["loss"]) <= 1e-5 * ref["loss"]
    err = rel_err(g, ref["grad"])
    print(f"twin seed {seed} strict {strict}: {res['valid']} valid of {hits} hit pixels, gradient relative "
          f"error {err:.2e}")
    assert err <= 1e-4


def test_adjoint_matches_fp64_reference_large_subset(gpu_ctx):
    p = S.gen_params("large", n_objects=32, slices=32, n_stars=300)
    sc = S.Scene.generate("large", p)
    objs = A.scene_objects(sc)
    obs = A.framing_observer(sc, objs, A.busiest_object(objs), 5.0, phi=(0.05, -0.1, 0.02))
    w, h = 512, 320
    obs, target = make_problem(gpu_ctx, sc, w, h, 3, obs=obs)
    g, res = A.gradient(gpu_ctx, obs, target, objs)
    ref = O.adjoint_ref(sc, obs.frame(w, h), obs.theta, target, objs)
    assert res["valid"] == ref["counts"]["valid"] and res["valid"] > 300
    err = rel_err(g, ref["grad"])
    print(f"large subset: {res['valid']} valid pixels, gradient relative error {err:.2e}")
    assert err <= 1e-4


# Central differences with step 1e-3 of the f32 loss against the adjoint on the measured pixel
# set: the agreement reached (see the test's printout), with a margin.
FD_TOL = 1e-4  # measured 5.05e-5 over 16 parameters on 800 pixels


def dilate(m, r):
    out = m.copy()
    for dy in range(-r, r + 1):
        for dx in range(-r, r + 1):
            out |= np.roll(np.roll(m, dy, axis=0), dx, axis=1)
    return out


def test_adjoint_against_gpu_finite_differences(gpu_ctx):
    """Central differences of the renderer's own f32 loss with the measured configuration's step
    1e-3, over a fixed pixel set, on bodies of the large scene seen from 5 units. The set is the
    strict validity rule's (no facet, texture or shadow edge among a pixel's neighbours: a finite
    step must not carry a pixel across one), without clock faces, whose hands and digits move
    with the proper time a pixel sees (a small step can flip a whole digit, a jump no mask in
    space excludes), and without the wireframe blend bands with a pixel around them, which are
    steeper than the step resolves. The FP64 comparisons cover those pixels."""
    p = S.gen_params("large", n_objects=32, slices=32, n_stars=300)
    sc = S.Scene.generate("large", p)
    objs = A.scene_objects(sc)
    k = A.busiest_object(objs)
    obs = A.framing_observer(sc, objs, k, 5.0, phi=(0.05, -0.1, 0.02))
    w, h = 512, 320
    obs, target = make_problem(gpu_ctx, sc, w, h, 3, obs=obs)
    gpu_ctx.set_option(OPT_ADJOINT_MASK, 1)
    try:
        _g, res = A.gradient(gpu_ctx, obs, target, objs, want_mask=True)
    finally:
        gpu_ctx.set_option(OPT_ADJOINT_MASK, 0)
    st = O.pixel_states(sc, obs.frame(w, h))
    wire = (st["flags"] & 1) != 0
    mask = ((res["mask"] > 0) & (st["pattern"] == 0) & ~dilate(wire, 1)).astype(np.uint8)
    assert mask.sum() > 300
    gpu_ctx.render_resident(0, obs.frame(w, h))
    g, _ = A.gradient(gpu_ctx, obs, target, objs, mask_in=mask)
    seen = [j for j in range(len(objs)) if np.abs(g[10 + 3 * j:13 + 3 * j]).sum() > 0][:2]
    assert seen
    params = list(range(10)) + [10 + 3 * j + a for j in seen for a in range(3)]
    fd = A.fd_gradient(gpu_ctx, sc, obs, w, h, target, objs, mask, params, step=1e-3)
    gs = np.array([g[q] for q in params])
    fs = np.array([fd[q] for q in params])
    err = rel_err(gs, fs)
    print(f"FD check (step 1e-3) over {len(params)} parameters on {int(mask.sum())} pixels: relative "
          f"difference {err:.2e}")
    for q in params:
        print(f"   param {q}: adjoint {g[q]:.6g}, difference {fd[q]:.6g}")
    assert err <= FD_TOL
    # the scene is back as it was: the same loss as before the object differences
    gpu_ctx.render_resident(0, obs.frame(w, h))
    again = gpu_ctx.adjoint(target, objs, mask_in=mask)["loss"]
    gpu_ctx.render_resident(0, obs.frame(w, h))
    assert gpu_ctx.adjoint(target, objs, mask_in=mask)["loss"] == again


def test_patch_scene_matches_a_full_upload(gpu_ctx, twin):
    fr = S.build_frame(twin.default_camera(), S.RenderConfig(160, 100))
    gpu_ctx.upload(0, twin)
    base = gpu_ctx.render_frame(0, fr)
    booster = A.ObjectBooster(twin)
    objs = A.scene_objects(twin)
    seen = [((base["hit"] >= f) & (base["hit"] < f + c)).sum() for f, c, _ in objs]
    first, count, center = objs[int(np.argmax(seen))]  # the object covering most pixels
    nr, pr = booster.boost(first, count, center, np.array([0.02, 0.05, -0.03]))
    assert gpu_ctx.patch(0, twin, nr, pr)
    patched = gpu_ctx.render_frame(0, fr)
    gpu_ctx.upload(0, twin)
    full = gpu_ctx.render_frame(0, fr)
    for k in ("rgb", "hit", "lam", "tau"):
        assert np.array_equal(patched[k], full[k]), k
    assert not np.array_equal(patched["rgb"], base["rgb"])
    nr, pr = booster.undo()
    assert gpu_ctx.patch(0, twin, nr, pr)
    for k, v in gpu_ctx.render_frame(0, fr).items():
        assert np.array_equal(v, base[k]), k


def test_patch_scene_falls_back_on_topology_change(gpu_ctx):
    a = S.Scene.generate("cube")
    b = S.Scene.generate("twin")
    gpu_ctx.upload(0, a)
    assert not gpu_ctx.patch(0, b, [(0, 1)], [])  # other counts: uploaded whole
    fr = S.build_frame(b.default_camera(), S.RenderConfig(96, 60))
    got = gpu_ctx.render_frame(0, fr)
    o = O.render(b, fr)
    assert np.array_equal(got["rgb"], o["rgb"])


def test_adjoint_argument_errors(gpu_ctx, twin):
    gpu_ctx.upload(0, twin)
    w, h = 64, 40
    tgt = np.zeros((h, w, 3), np.float32)
    gpu_ctx.render_resident(0, S.build_frame(twin.default_camera(),
                                             S.RenderConfig(w, h, S.LookbackIsochrones, S.ShadowOff)))
    with pytest.raises(SpecrelError):
        gpu_ctx.adjoint(tgt)
    obs = A.Observer(np.array(twin.default_camera().position, np.float64), np.zeros(3))
    gpu_ctx.render_resident(0, obs.frame(w, h))
    objs = A.scene_objects(twin)
    with pytest.raises(SpecrelError):
        gpu_ctx.adjoint(tgt, list(reversed(objs)))
    res = gpu_ctx.adjoint(tgt, objs)
    assert res["valid"] + res["miss"] + res["starlight"] + res["border"] + res["silhouette"] + res["edge"] == w * h


def test_adjoint_full_workload_against_fp64(gpu_ctx):
    """The measured configuration: 32-object subset of the large scene (3,932,160 tets), 512x320,
    observer at rest framing the busiest object, target at a hidden observer (rapidity 0.8, event
    offset N(0,1)^4), 106 parameters."""
    p = S.gen_params("large", n_objects=32, slices=128, n_stars=10000)
    sc = S.Scene.generate("large", p)
    assert sc.num_prims == 32 * 320 * 3 * 128
    obs, hidden, objs = A.adjoint_problem(sc, seed=7)
    w, h = 512, 320
    gpu_ctx.upload(0, sc)
    gpu_ctx.render_resident(0, hidden.frame(w, h))
    target = gpu_ctx.download_linear(0)
    gpu_ctx.render_resident(0, obs.frame(w, h))
    assert len(objs) == 32
    g, res = A.gradient(gpu_ctx, obs, target, objs)
    ref = O.adjoint_ref(sc, obs.frame(w, h), obs.theta, target, objs)
    assert res["valid"] == ref["counts"]["valid"] and g.size == 106
    err = rel_err(g, ref["grad"])
    cam_err = rel_err(g[:10], ref["grad"][:10])
    print(f"full workload: {res['valid']} valid of {w * h} pixels (silhouette {res['silhouette']}, "
          f"edge {res['edge']}, miss {res['miss']}), gradient relative error {err:.2e} "
          f"(observer part {cam_err:.2e})")
    assert err <= 1e-4


def test_batched_adjoint_matches_single_views_and_fp64(gpu_ctx):
    """Batched adjoint over accelerated bodies and light clocks: every view's gradient is
    bit-identical to the single-view adjoint of that view, and matches the FP64 reference on a
    1% sample of its valid pixels."""
    p = S.gen_params("accel_clocks", n_objects=12, n_clocks=4, slices=48)
    sc = S.Scene.generate("accel_clocks", p)
    gpu_ctx.upload(0, sc)
    obs, hidden, sizes, objs = A.batch_problem(sc, 5, seed=21, w_min=96, w_max=320, distance=(4.0, 8.0))
    targets = A.batch_targets(gpu_ctx, 0, hidden, sizes)
    grads, res = A.batch_gradients(gpu_ctx, 0, obs, sizes, targets, objs, want_mask=True)
    assert sum(r["valid"] for r in res) > 500
    rng = np.random.default_rng(5)
    sample_masks = []
    for v, (o, (w, h)) in enumerate(zip(obs, sizes)):
        gpu_ctx.render_resident(0, o.frame(w, h))
        g1, r1 = A.gradient(gpu_ctx, o, targets[v], objs, want_mask=True)
        assert np.array_equal(g1, grads[v]) and r1["loss"] == res[v]["loss"], f"view {v}"
        assert np.array_equal(r1["mask"], res[v]["mask"])
        m = (res[v]["mask"] > 0) & (rng.random(res[v]["mask"].shape) < 0.01)
        if m.sum() == 0 and res[v]["valid"] > 0:  # keep one pixel of a small view
            idx = np.argwhere(res[v]["mask"] > 0)[0]
            m[tuple(idx)] = True
        sample_masks.append(m.astype(np.uint8))
    grads_s, res_s = A.batch_gradients(gpu_ctx, 0, obs, sizes, targets, objs, masks_in=sample_masks)
    errs = []
    for v, (o, (w, h)) in enumerate(zip(obs, sizes)):
        ref = O.adjoint_ref(sc, o.frame(w, h), o.theta, targets[v], objs, mask_in=sample_masks[v])
        assert res_s[v]["valid"] == int(sample_masks[v].sum())
        if np.linalg.norm(ref["grad"]) > 0:
            errs.append(rel_err(grads_s[v], ref["grad"]))
    print(f"batched adjoint: {len(obs)} views, {sum(r['valid'] for r in res)} valid pixels; FP64 on a 1% sample: "
          f"max relative error {max(errs):.2e}")
    assert errs and max(errs) <= 1e-4
