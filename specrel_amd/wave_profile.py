"""Per-wave timeline of the tiled primary-ray kernel on the large scene's headline frame (the
128-object, 128-slice scene with 10,000 stars at 1280x800, observer at the origin event with
rapidity 1, shaded with null shadows), from the waves' own clock records (SPECREL_OPT_WAVE_CLOCK):
when the waves start and end, how many are resident over the kernel's life, the share of its time
spent with few waves left (the tail), how the work spreads over the XCDs, and which tiles end last.

    python -m specrel_amd.wave_profile [--width 1280] [--height 800] [--frames 5] [--objects 128]

Prints a markdown summary. Needs a gfx950 device.
"""
import argparse
import sys
import time

import numpy as np

from . import scene as S
from .renderer import OPT_WAVE_CLOCK, Context


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def resident_curve(start, end):
    """Times (sorted) at which the resident wave count changes, and the count after each."""
    t = np.concatenate([start, end])
    d = np.concatenate([np.ones_like(start, np.int64), -np.ones_like(end, np.int64)])
    o = np.lexsort((-d, t))  # at equal times, starts before ends
    return t[o], np.cumsum(d[o])


def time_below(t, count, n, t0, t1):
    """Share of [t0, t1] with fewer than n waves resident."""
    span = np.diff(np.append(t, t1))
    return float(span[count < n].sum() / max(t1 - t0, 1))


def summarize(rec, khz, trace_ms):
    us = 1000.0 / khz  # microseconds per tick
    start = rec["start"].astype(np.int64)
    end = rec["end"].astype(np.int64)
    t0, t1 = int(start.min()), int(end.max())
    dur = (end - start) * us
    t, count = resident_curve(start, end)
    done = np.sort(end - t0) * us
    n = len(rec)
    xcc = rec["xcc"]
    per_xcc = {int(x): (int((xcc == x).sum()), float((end[xcc == x].max() - t0) * us)) for x in np.unique(xcc)}
    cu = (rec["hw_id"] >> 8) & 0xF
    se = (rec["hw_id"] >> 13) & 0x7
    sh = (rec["hw_id"] >> 12) & 0x1
    units = xcc.astype(np.int64) * 256 + se * 32 + sh * 16 + cu
    last = np.argsort(end)[-8:][::-1]
    return {
        "waves": n,
        "span_us": (t1 - t0) * us,
        "trace_ms_events": trace_ms,
        "all_started_us": float((start.max() - t0) * us),
        "duration_us": {q: float(np.percentile(dur, p)) for q, p in
                        (("p50", 50), ("p90", 90), ("p99", 99), ("max", 100))},
        "finished_us": {q: float(done[min(n - 1, int(p * n))]) for q, p in
                        (("50%", 0.5), ("90%", 0.9), ("99%", 0.99), ("99.9%", 0.999))},
        "share_below": {k: time_below(t, count, k, t0, t1) for k in (64, 256, 1024, 4096)},
        "peak_resident": int(count.max()),
        "compute_units_used": int(len(np.unique(units))),
        "per_xcc": per_xcc,
        "last_waves": [(int(rec["tile"][i]), int(rec["quarter"][i]), float((start[i] - t0) * us),
                        float(dur[i])) for i in last],
    }


def report(runs, args, scene):
    r = runs[-1]
    lines = [
        f"# k_trace per-wave timeline, large scene {args.width}x{args.height}",
        "",
        f"Command: `python -m specrel_amd.wave_profile --width {args.width} --height {args.height} "
        f"--frames {args.frames} --objects {args.objects}"
        + "".join(f" --option {o}" for o in args.option) + "`",
        "",
        f"Scene: {scene.num_prims} tets, {scene.num_nodes} nodes, {scene.num_stars} stars. "
        f"{r['waves']} waves (one per 8x8 quarter of a 16x16 tile), {r['peak_resident']} resident at the peak on "
        f"{r['compute_units_used']} compute units. Frames after warm-up (tile order converged): {len(runs)}.",
        "",
        "| frame | span µs (wave clocks) | trace ms (HIP events) | all started µs | 50% / 90% / 99% / 99.9% done µs "
        "| time with < 64 / 256 / 1024 / 4096 waves resident |",
        "|---|---|---|---|---|---|",
    ]
    for i, x in enumerate(runs):
        f, b = x["finished_us"], x["share_below"]
        lines.append(f"| {i} | {x['span_us']:.1f} | {x['trace_ms_events']:.4f} | {x['all_started_us']:.1f} | "
                     f"{f['50%']:.1f} / {f['90%']:.1f} / {f['99%']:.1f} / {f['99.9%']:.1f} | "
                     f"{100 * b[64]:.1f}% / {100 * b[256]:.1f}% / {100 * b[1024]:.1f}% / {100 * b[4096]:.1f}% |")
    d = r["duration_us"]
    lines += [
        "",
        f"Wave durations (last frame): median {d['p50']:.1f} µs, p90 {d['p90']:.1f}, p99 {d['p99']:.1f}, "
        f"max {d['max']:.1f}.",
        "",
        "Per XCD (last frame): waves, and when its last wave ended (µs after the first start): "
        + ", ".join(f"{k}: {v[0]} / {v[1]:.1f}" for k, v in sorted(r["per_xcc"].items())) + ".",
        "",
        "Last waves to end (tile, quarter, start µs, duration µs): "
        + "; ".join(f"({a}, {b}, {c:.1f}, {e:.1f})" for a, b, c, e in r["last_waves"]) + ".",
    ]
    return "\n".join(lines)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--width", type=int, default=1280)
    ap.add_argument("--height", type=int, default=800)
    ap.add_argument("--frames", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--objects", type=int, default=128)
    ap.add_argument("--slices", type=int, default=128)
    ap.add_argument("--stars", type=int, default=10000)
    ap.add_argument("--option", action="append", default=[], metavar="N=V",
                    help="specrel_set_option(N, V) before the renders (repeatable)")
    args = ap.parse_args(argv)
    p = S.gen_params("large", n_objects=args.objects, slices=args.slices, n_stars=args.stars)
    t = time.perf_counter()
    log("building the large scene")
    scene = S.Scene.generate("large", p)
    log(f"scene: {scene.num_prims} tets in {time.perf_counter() - t:.1f} s")
    frame = S.build_frame(S.default_camera("large", p),
                          S.RenderConfig(args.width, args.height, S.RelativisticShaded, S.ShadowLorentzInvariant,
                                         S.PastLightCone))
    ctx = Context(0)
    for o in args.option:
        k, v = o.split("=")
        ctx.set_option(int(k), int(v))
    ctx.upload(0, scene)
    for _ in range(args.warmup):
        ctx.render_resident(0, frame)
        ctx.sync()
    ctx.set_option(OPT_WAVE_CLOCK, 1)
    runs = []
    for i in range(args.frames):
        ctx.render_resident(0, frame)
        ctx.sync()
        rec, khz = ctx.wave_clock()
        if not len(rec) or not khz:
            raise RuntimeError("no wave clock records (the tiled primary-ray kernel did not run)")
        runs.append(summarize(rec, khz, ctx.timings()["trace_ms"]))
        log(f"frame {i}: span {runs[-1]['span_us']:.1f} us")
    ctx.set_option(OPT_WAVE_CLOCK, 0)
    print(report(runs, args, scene))
    ctx.close()


if __name__ == "__main__":
    main()
