"""Writes tests/golden/oracle_golden.json and demo_golden.json: SHA-256 digests of the oracle's
outputs (RGB8, hit index, lambda, tau) for fixed scenes and views (the second: the five demo
scenes under the reference's 9-configuration harness). They freeze the oracle and the scene producer
against unintended change; the reference itself cannot be run here (no Rust toolchain), so these
are regression vectors of the restatement, not reference outputs."""
import hashlib
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

from oracle import pyoracle as O  # noqa: E402
from specrel_amd import scene as S  # noqa: E402

CASES = {
    "cube_320x200_shaded": dict(scene="cube", w=320, h=200, mode=0, shadow=1, view=0),
    "twin_192x120_shaded": dict(scene="twin", w=192, h=120, mode=0, shadow=1, view=0),
    "twin_192x120_isochrones_frontier": dict(scene="twin", w=192, h=120, mode=1, shadow=0, view=2),
    "decision_age_192x120_prophecy": dict(scene="decision_age", w=192, h=120, mode=3, shadow=0, view=3),
}


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def demo_harness():
    """Digests of the five demo scenes under the reference's 9-configuration harness at 192x120
    (gpu.rs:1256-1311), written to demo_golden.json."""
    from tests.helpers import DEMO_SCENES, HARNESS_CONFIGS, harness_frame
    out = {}
    for name in DEMO_SCENES:
        sc = S.Scene.generate(name)
        for c in range(len(HARNESS_CONFIGS)):
            r = O.render(sc, harness_frame(name, c))
            out[f"{name}_{c}"] = {"scene": name, "config": c, **{k: digest(r[k]) for k in ("rgb", "hit", "lam", "tau")}}
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "demo_golden.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print("wrote", path)


def main():
    demo_harness()
    out = {}
    for key, spec in CASES.items():
        sc = S.Scene.generate(spec["scene"])
        fr = S.build_frame(S.default_camera(spec["scene"]),
                           S.RenderConfig(spec["w"], spec["h"], spec["mode"], spec["shadow"], spec["view"]))
        r = O.render(sc, fr)
        out[key] = dict(spec, **{k: digest(r[k]) for k in ("rgb", "hit", "lam", "tau")})
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "oracle_golden.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print("wrote", path)


if __name__ == "__main__":
    main()
