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
