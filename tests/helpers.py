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
