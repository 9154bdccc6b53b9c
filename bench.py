#!/usr/bin/env python3
"""Benchmark of the SpecRel per-pixel hot path on one MI355X.

Workload (the headline single-GPU configuration): the large scene -- 128 seeded bodies
(320-triangle icospheres, radius U[0.5,3], centres U[-50,50]^3, |v| U[0,0.9c), a quarter with
constant proper acceleration), 128 time slices each = 15,728,640 tetrahedra, plus 10,000 stars
on a sphere of radius 500 (T log-uniform 3000-30000 K) -- seen by an observer at the origin
event boosted to rapidity 1.0, 1280x800 past-light-cone rays, null shadows, f32.

A step is one full frame: star pass and star binning, primary-ray trace, shading (with shadow
rays and the binned stars), bloom pyramid, tone map to RGB8, with every input already resident
in HBM and the image left there (no readback).
value = primary rays per second of the whole job = W*H*steps / wall time, in Mrays/s.

Prints ONE JSON line. Extra legs: `roofline` for the dominant kernel (k_trace): the compulsory
bytes of the frame (each distinct record touched once, counted by the instrumented render) over
its HIP-event duration against the HBM peak, with the measured traffic and the per-visit work
beside it; `cpu_baseline` from the CPU oracle (a restatement of the reference renderer) timed on
this host's cores on a bounded row sample; `parity`: those oracle rows against the same rows of
the GPU frame, bit for bit (the run exits 1 on a mismatch); `e2e_ms_with_readback`: the median of
50 render calls through the host boundary with the RGB8 readback into the caller's buffer;
`batched`: 64 observers of uneven resolution (160x100 .. 1920x1200) over the same scene in one
batched render (autotuned launch configuration, graph replay) against a loop of single-view
renders of the same frames, in views/s; `adjoint`: forward + adjoint times of the image gradient
on the 32-object subset at 512x320; `batched_adjoint`: the same for 32 observers of uneven
resolution over the accelerated-bodies scene in one batch.
"""
import argparse
import json
import os
import sys
import time
import types

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mrays/s at 1280x800 (15.7M tets) vs CPU ref; batched views/s; fwd+adjoint ms"
HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
CUS = 256  # MI355X compute units
SIMDS_PER_CU = 4
ENGINE_CLOCK_HZ = 2.4e9  # MI355X peak engine clock


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def work_bytes(st, w, h, n_stars, in_trace, fused=False):
    """Bytes the traversal and shading request from the memory hierarchy per launch of the two
    per-pixel kernels, every visit priced with the reference's record sizes: 48 B per node pop
    (BvhNode4D / GpuNode), 32 B per child box tested, 80 B per hot record tested (SimplexGeom4D),
    192 + 48 B per hit (cold attributes + material). The rays of a wave visit nearly the same
    records, which the L1 serves, so this is work, not HBM traffic.
      k_trace: the primary traversal, plus 4 B per pixel for the hit slot written; with the shadow
               queries traced there (in_trace), also the shadow traversal, the hit's 192 + 48 B
               and 4 B per pixel of occlusion mask written.
      k_shade: 4 B per pixel of hit slot read, per hit the 80 B record intersected again plus
               192 + 48 B, the shadow queries priced like the primary ones (or 4 B per pixel of
               occlusion mask read), 64 B per star of the view, and 16 B per pixel of HDR written.
    With the shading fused into k_trace (fused), k_trace does both but reads neither the slot nor
    the occlusion bits back."""
    shadow = 48 * st["shadow_node_pops"] + 32 * st["shadow_box_tests"] + 80 * st["shadow_prim_tests"]
    trace = 48 * st["node_pops"] + 32 * st["box_tests"] + 80 * st["prim_tests"] + 4 * w * h
    shade = 4 * w * h + (80 + 240) * st["hits"] + 64 * n_stars + 16 * w * h
    if in_trace:
        trace += shadow + 240 * st["hits"] + 4 * w * h
        shade += 4 * w * h
    else:
        shade += shadow
    if fused:
        return {"k_trace": trace + shade - (8 if in_trace else 4) * w * h}
    return {"k_trace": trace, "k_shade": shade}


def compulsory_bytes(st, w, h, n_stars, in_trace, fused=False):
    """Algorithmic HBM bytes per launch: every distinct record the frame touches, read once, and
    every per-pixel result written once -- what the kernels must move between HBM and the chip
    however well the caches serve repeated visits. Records in the device layout: 72 B per
    interior node whose child boxes some ray tested (both children's boxes, 64 B, plus the two
    child entries, 8 B), 80 B per primitive record some ray tested, 192 B of attributes + 48 B of
    material per primitive some hit read (the distinct counts come from the instrumented render).
      k_trace: those records (primary and shadow rays), 4 B hit slot per pixel written, and with
               the shadow queries traced there 4 B of occlusion bits per pixel written.
      k_shade: 4 B slot (+ 4 B occlusion bits) per pixel read, per distinct hit primitive its
               80 + 192 + 48 B, 64 B per star, 16 B of HDR per pixel written.
      fused:   k_trace with the shading in it: the records once (a hit primitive's record and
               attributes are the traversal's), 64 B per star, slot, occlusion bits and HDR
               written (24 B per pixel with shadows traced, else 20)."""
    records = 72 * st["distinct_nodes"] + 80 * st["distinct_prims"] + 240 * st["distinct_hits"]
    if fused:
        return {"k_trace": records + 64 * n_stars + ((8 if in_trace else 4) + 16) * w * h}
    trace = records + 4 * w * h + (4 * w * h if in_trace else 0)
    shade = (8 if in_trace else 4) * w * h + (80 + 240) * st["distinct_hits"] + 64 * n_stars + 16 * w * h
    return {"k_trace": trace, "k_shade": shade}


def cpu_baseline(scene, frame, w, h, rows, runs=5):
    """The oracle (a C restatement of the reference's render_frame) on this host's cores: one
    warm-up and `runs` timed renders of the same row sample (every (h // rows)-th row, the full
    10,000-star loop per pixel as the reference does), median, extrapolated to the frame, plus
    bloom + tone map of a full frame. The warm-up and the first timed render also return their
    rows (pre-bloom HDR, hit index, lambda, tau) for the parity check against the GPU frame."""
    from oracle import pyoracle as O  # the CPU restatement of the reference renderer
    import numpy as np

    cores = O.threads()
    step = max(1, h // rows)
    n = h // step
    checked = []
    hdr, _, hits = O.render_rows(scene, frame, 0, step, n, nthreads=cores, hits=True)  # warm-up
    checked.append((0, step, hdr, hits))
    times = []
    for r in range(runs):
        t0 = time.perf_counter()
        out = O.render_rows(scene, frame, step // 2, step, n, nthreads=cores, hits=(r == 0))
        times.append(time.perf_counter() - t0)
        if r == 0:
            checked.append((step // 2, step, out[0], out[2]))
    t_rows = float(np.median(times))
    # The same rows without the star loop: the split between traversal + shading and stars.
    no_stars = types.SimpleNamespace(arrays=type(scene.arrays).from_buffer_copy(scene.arrays))
    no_stars.arrays.num_stars = 0
    tt = []
    for _ in range(runs):
        t0 = time.perf_counter()
        O.render_rows(no_stars, frame, step // 2, step, n, nthreads=cores)
        tt.append(time.perf_counter() - t0)
    t_trace = float(np.median(tt))
    hdr = np.full((h, w, 3), 0.25, np.float32)
    tp = []
    for _ in range(runs):
        t0 = time.perf_counter()
        O.pack(O.bloom(hdr))
        tp.append(time.perf_counter() - t0)
    t_post = float(np.median(tp))
    frame_s = t_rows * (h / n) + t_post
    res = {
        "value": round(w * h / frame_s / 1e6, 5),
        "unit": "Mrays/s",
        "cores": cores,
        "kind": "port",
        "sample": f"{n} of {h} rows (every {step}th, {n * w} rays incl. the full star loop and shadow rays), "
                  f"median of {runs} timed runs after a warm-up: {t_rows:.2f} s, plus bloom+tone map of a full "
                  f"{w}x{h} frame in {t_post:.3f} s; extrapolated frame time {frame_s:.2f} s",
        "split_s_per_frame": {"traversal_shading_shadows": round(t_trace * h / n, 3),
                              "stars": round((t_rows - t_trace) * h / n, 3), "bloom_tonemap": round(t_post, 3)},
    }
    return res, checked


def parity_rows(ctx, frame, checked):
    """The oracle's rows (from the CPU baseline) against the same rows of the GPU frame: pre-bloom
    linear image, hit index, lambda and tau bit for bit."""
    import numpy as np

    g = ctx.render_frame(0, frame, want_hits=True)
    lin = ctx.download_linear(0)
    bad = {"linear": 0, "hit": 0, "lam": 0, "tau": 0}
    rows = pixels = 0
    for row0, step, hdr, hits in checked:
        ys = np.arange(row0, row0 + step * len(hdr), step)
        rows += len(ys)
        pixels += hdr.shape[0] * hdr.shape[1]
        bad["linear"] += int((lin[ys].view(np.uint32) != hdr.view(np.uint32)).any(axis=2).sum())
        for k in ("hit", "lam", "tau"):
            bad[k] += int((g[k][ys].view(np.uint32) != hits[k].view(np.uint32)).sum())
    return {"rows": rows, "pixels": pixels, "mismatched_pixels": bad, "ok": not any(bad.values()),
            "compared": "pre-bloom linear RGB, hit index, lambda, tau bit for bit against the oracle's rows"}


def bench_batch(ctx, scene, n_views, steps, seed):
    """Views/s of one batched render of `n_views` observers against a loop of single-view
    renders of the same frames (each enqueue + wait, outputs left resident either way)."""
    from specrel_amd import scene as S
    from specrel_amd.renderer import OPT_GRAPH

    obs = S.batch_observers(S.scene_box(scene), n_views, seed=seed)
    frames = [S.build_frame(c, r) for c, r in obs]
    rays = sum(f.size[0] * f.size[1] for f in frames)
    arr = ctx._frames(frames)
    log(f"batch: {n_views} views, {rays} rays; autotuning")
    tuning = ctx.autotune(0, arr, iters=5)
    log(f"autotune: {tuning}")
    ctx.render_batch_resident(0, arr)
    ctx.sync()
    t0 = time.perf_counter()
    kern = 0.0
    for _ in range(steps):
        ctx.render_batch_resident(0, arr)
        ctx.sync()
        kern += ctx.timings()["kernel_ms"]
    batch_s = (time.perf_counter() - t0) / steps
    # The loop a caller of the single-view interface writes: one render per view, launched
    # directly (no graph), with the same launch options otherwise.
    ctx.set_option(OPT_GRAPH, 0)
    for f in frames[:4]:
        ctx.render_resident(0, f)
        ctx.sync()
    loops = max(1, steps // 2)
    t0 = time.perf_counter()
    for _ in range(loops):
        for f in frames:
            ctx.render_resident(0, f)
            ctx.sync()
    loop_s = (time.perf_counter() - t0) / loops
    ctx.set_option(OPT_GRAPH, tuning["graph"])
    return {
        "views": n_views,
        "rays": rays,
        "views_per_s": round(n_views / batch_s, 2),
        "Mrays_per_s": round(rays / batch_s / 1e6, 2),
        "ms_per_batch": round(batch_s * 1e3, 3),
        "kernel_ms_per_batch": round(kern / steps, 3),
        "loop_views_per_s": round(n_views / loop_s, 2),
        "loop_ms": round(loop_s * 1e3, 3),
        "speedup_vs_loop": round(loop_s / batch_s, 3),
        "tuning": tuning,
        "workload": f"{n_views} observers, width U[160,1920] x height width*5/8, events U(scene box), "
                    f"|rapidity| U[0,2], random rotate_local, past-cone/simultaneity mix, seed {seed}",
    }


def bench_adjoint(ctx, steps, slot=1):
    """Forward render + adjoint on the 32-object subset of the large scene (3,932,160 tets at
    128 slices) at 512x320, seen by an observer at rest 8 units behind the object with the most
    neighbours: L2 loss on the linear image against a target rendered at a hidden observer
    (rapidity 0.8 along a seeded direction, event offset N(0,1)^4), gradient with respect to
    event (4), rapidity (3), rotation (3) and 32 object boosts (3 each)."""
    import numpy as np
    from specrel_amd import adjoint as A
    from specrel_amd import scene as S

    p = S.gen_params("large", n_objects=32, slices=128, n_stars=10000)
    t0 = time.perf_counter()
    sc = S.Scene.generate("large", p)
    log(f"adjoint scene: {sc.num_prims} tets in {time.perf_counter() - t0:.1f} s")
    obs, hidden, objs = A.adjoint_problem(sc, seed=7)
    w, h = 512, 320
    ctx.upload(slot, sc)
    ctx.render_resident(slot, hidden.frame(w, h))
    target = ctx.download_linear(0)
    frame = obs.frame(w, h)
    for _ in range(2):
        ctx.render_resident(slot, frame)
        ctx.sync()
        A.gradient(ctx, obs, target, objs)
    fwd, adj, wall = [], [], []
    for _ in range(steps):
        t0 = time.perf_counter()
        ctx.render_resident(slot, frame)
        ctx.sync()
        fwd.append(ctx.timings()["kernel_ms"])
        g, res = A.gradient(ctx, obs, target, objs)
        wall.append((time.perf_counter() - t0) * 1e3)
        adj.append(ctx.timings()["adjoint_ms"])
    f, a = float(np.median(fwd)), float(np.median(adj))
    return {
        "tets": sc.num_prims,
        "pixels": w * h,
        "params": int(g.size),
        "fwd_ms": round(f, 4),
        "adjoint_ms": round(a, 4),
        "fwd_plus_adjoint_ms": round(f + a, 4),
        "ratio_to_fwd": round((f + a) / f, 3),
        "wall_ms_per_gradient": round(float(np.median(wall)), 3),
        "valid_pixels": int(res["valid"]),
        "masked": {k: int(res[k]) for k in ("miss", "starlight", "border", "silhouette", "edge")},
        "loss": res["loss"],
        "workload": "32-object subset of the large scene (128 slices), 512x320, observer at rest 8 units behind the "
                    "busiest object; L2 on the linear pre-bloom image against a hidden observer (rapidity 0.8, event "
                    "offset N(0,1)^4, seed 7); 106 parameters",
    }


def bench_extrude(ctx, p, steps):
    """Worldtube extrusion of the large scene's tubes on the GPU (specrel_extrude_tubes) against
    the host extrusion of the same descriptors on all cores; records compared byte for byte."""
    import numpy as np
    from specrel_amd import scene as S

    d = S.DeferredScene("large", p)
    tubes = d.tubes()
    _, _, _, counts = ctx.extrude_tubes(tubes, download=False)  # warm-up (allocations)
    ms = []
    for _ in range(steps):
        ctx.extrude_tubes(tubes, download=False)
        ms.append(ctx.timings()["extrude_ms"])
    ms.sort()
    kernel_ms = ms[len(ms) // 2]
    t0 = time.perf_counter()
    got = ctx.extrude_tubes(tubes)
    e2e_s = time.perf_counter() - t0
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    t0 = time.perf_counter()
    ref = S.extrude_tubes_host(tubes)
    host_s = time.perf_counter() - t0
    same = all(np.array_equal(x.view(np.uint32), y.view(np.uint32)) for x, y in zip(got[:3], ref[:3])) and \
        np.array_equal(got[3], ref[3])
    n = int(counts.sum())
    out_bytes = n * (80 + 192 + 32)
    del got
    # BVH over the extruded records: on the device (the records never leave it) and on the host.
    ctx.extrude_tubes(tubes, download=False)
    t0 = time.perf_counter()
    dn, do, dst = ctx.build_bvh(None, n)
    dev_build_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    hn, ho = S.build_bvh_host(ref[2])
    host_build_s = time.perf_counter() - t0
    tree_same = np.array_equal(dn.view(np.uint32), hn.view(np.uint32)) and np.array_equal(do, ho)
    bvh = {"device_ms": round(dev_build_s * 1e3, 1), "device_stats": dst, "host_ms": round(host_build_s * 1e3, 1),
           "host_cores": cores, "speedup": round(host_build_s / dev_build_s, 1), "nodes": int(len(dn)),
           "bit_identical_to_host": bool(tree_same)}
    return {"tubes": len(tubes), "tets": n, "kernel_ms": round(kernel_ms, 3),
            "Mtets_per_s": round(n / kernel_ms / 1e3, 1),
            "achieved_GBs": round(out_bytes / (kernel_ms * 1e-3) / 1e9, 1),
            "algorithmic_bytes": out_bytes,
            "e2e_ms_with_download": round(e2e_s * 1e3, 1),
            "host_ms": round(host_s * 1e3, 1), "host_cores": cores,
            "speedup_kernel_vs_host": round(host_s * 1e3 / kernel_ms, 1),
            "bit_identical_to_host": bool(same), "bvh_build": bvh,
            "workload": "the large scene's worldtubes (icospheres, 320 triangles, inertial and accelerated), "
                        "every prism of every slice; outputs GpuGeom + GpuAttr + padded bounds"}


def bench_batch_adjoint(ctx, steps, slot=2, n_views=32):
    """Batched forward + adjoint: 32 observers of uneven resolution (128x80 .. 1024x640), each with
    its own parameters and gradient, over the accelerated-bodies scene (64 Born-rigid bodies with
    proper acceleration U[0.05, 0.5], 16 light clocks, 256 slices), one launch per pass."""
    import numpy as np
    from specrel_amd import adjoint as A
    from specrel_amd import scene as S

    t0 = time.perf_counter()
    sc = S.Scene.generate("accel_clocks", S.gen_params("accel_clocks"))
    log(f"accelerated scene: {sc.num_prims} tets, {len(sc.objects())} objects in {time.perf_counter() - t0:.1f} s")
    obs, hidden, sizes, objs = A.batch_problem(sc, n_views, seed=11)
    ctx.upload(slot, sc)
    targets = A.batch_targets(ctx, slot, hidden, sizes)
    frames = ctx._frames([o.frame(w, h) for o, (w, h) in zip(obs, sizes)])
    for _ in range(2):
        ctx.render_batch_resident(slot, frames)
        ctx.sync()
        ctx.adjoint_batch(targets, objs)
    fwd, adj = [], []
    for _ in range(steps):
        ctx.render_batch_resident(slot, frames)
        ctx.sync()
        fwd.append(ctx.timings()["kernel_ms"])
        res = ctx.adjoint_batch(targets, objs)
        adj.append(ctx.timings()["adjoint_ms"])
    f, a = float(np.median(fwd)), float(np.median(adj))
    return {
        "tets": sc.num_prims,
        "views": n_views,
        "rays": int(sum(w * h for w, h in sizes)),
        "params_per_view": 10 + 3 * len(objs),
        "fwd_ms": round(f, 4),
        "adjoint_ms": round(a, 4),
        "fwd_plus_adjoint_ms": round(f + a, 4),
        "ratio_to_fwd": round((f + a) / f, 3),
        "valid_pixels": int(sum(r["valid"] for r in res)),
        "workload": f"{n_views} observers framing the bodies (width U[128,1024], height width*5/8), each with "
                    f"its own target (hidden event offset N(0,0.5)^4, rapidity offset N(0,0.1)^3), seed 11; "
                    f"64 accelerated bodies + 16 light clocks, 256 slices",
    }


def copy_probe():
    """A plain copy kernel of known size: 2^28 floats (1 GiB read + 1 GiB written per launch).
    Its FETCH_SIZE / WRITE_SIZE readings scale the counters of the renderer kernels to bytes."""
    import torch

    x = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(4):
        y.copy_(x)
    torch.cuda.synchronize()
    log("copy probe: 4 launches of a 1 GiB device-to-device copy kernel")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--width", type=int, default=1280)
    ap.add_argument("--height", type=int, default=800)
    ap.add_argument("--objects", type=int, default=128)
    ap.add_argument("--slices", type=int, default=128)
    ap.add_argument("--stars", type=int, default=10000)
    ap.add_argument("--cpu-rows", type=int, default=80, help="rows of the CPU baseline's sample")
    ap.add_argument("--e2e-calls", type=int, default=50, help="timed render + readback calls (median)")
    ap.add_argument("--views", type=int, default=64, help="observers of the batched leg (0: skip it)")
    ap.add_argument("--batch-steps", type=int, default=5)
    ap.add_argument("--adjoint-steps", type=int, default=10, help="forward+adjoint repetitions (0: skip the leg)")
    ap.add_argument("--extrude-steps", type=int, default=5, help="GPU extrusion repetitions (0: skip the leg)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--stack", type=int, default=0,
                    help="traversal stack: K>0 positions in LDS, -N top entries in registers (0: library default)")
    ap.add_argument("--persistent", type=int, default=-1, help="1/0: persistent primary-ray kernel (-1: default)")
    ap.add_argument("--shadow-queue", type=int, default=-1, help="1/0: queued shadow rays (-1: default)")
    ap.add_argument("--merged", type=int, default=-1,
                    help="1/0: merged load rounds of node and leaf lanes in the primary trace (-1: default)")
    ap.add_argument("--star-overlap", type=int, default=-1, help="1/0: star pass beside the primary rays (-1: default)")
    ap.add_argument("--wave-shape", type=int, default=-1,
                    help="pixels a primary-ray wave takes: 0 8x8, 1 16x4, 2 4x16 (-1: default)")
    ap.add_argument("--trace-waves", type=int, default=0, help="1, 2, 4: waves per primary-ray block (0: default)")
    ap.add_argument("--shadow-trace", type=int, default=-1,
                    help="1/0: shadow queries traced by the primary-ray kernel (-1: default)")
    ap.add_argument("--node-width", type=int, default=0,
                    help="2 or 4: BVH node fetch width of the primary trace (0: default)")
    ap.add_argument("--lds-top", type=int, default=-1,
                    help="0..8: top tree levels each primary-ray block reads from LDS (-1: default)")
    ap.add_argument("--fused-shade", type=int, default=-1,
                    help="1/0: the primary-ray kernel shades its own pixels (-1: default)")
    ap.add_argument("--stage-events", type=int, default=1,
                    help="device events per timed frame: 1 only around the primary-ray kernel (its "
                         "duration for the roofline), 2 at every stage boundary, 0 none; the stage "
                         "breakdown comes from --breakdown-steps further frames with every event")
    ap.add_argument("--trace-events-every", type=int, default=8,
                    help="with --stage-events 1, the timed frames that record the pair of events around "
                         "the primary-ray kernel: every Nth (the others record none)")
    ap.add_argument("--breakdown-steps", type=int, default=20,
                    help="frames rendered after the timed ones with an event at every stage boundary")
    ap.add_argument("--copy-probe", action="store_true",
                    help="run a plain 1 GiB device copy (the HBM-counter calibration) and exit")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe()
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the path shards by image tiles or "
              f"views, see DESIGN.md", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    from specrel_amd import scene as S
    from specrel_amd.renderer import (OPT_PERSISTENT, OPT_MERGED_LOADS, OPT_SHADOW_IN_TRACE, OPT_SHADOW_QUEUE,
                                      OPT_STACK, OPT_STAR_OVERLAP, OPT_STATS, OPT_TRACE_WAVES, OPT_WAVE_SHAPE,
                                      OPT_NODE_WIDTH, OPT_LDS_TOP, OPT_FUSED_SHADE, OPT_STAGE_EVENTS,
                                      SHADOWS_IN_TRACE,
                                      Context)

    w, h = args.width, args.height
    p = S.gen_params("large", n_objects=args.objects, slices=args.slices, n_stars=args.stars)
    t0 = time.perf_counter()
    log("building the large scene (meshes, worldlines, extrusion, SAH BVH)")
    scene = S.Scene.generate("large", p)
    log(f"scene: {scene.num_prims} tets, {scene.num_nodes} nodes, {scene.num_stars} stars "
        f"in {time.perf_counter() - t0:.1f} s")
    cam = S.default_camera("large", p)
    frame = S.build_frame(cam, S.RenderConfig(w, h, S.RelativisticShaded, S.ShadowLorentzInvariant,
                                               S.PastLightCone))
    ctx = Context(0)
    if args.stack:
        ctx.set_option(OPT_STACK, args.stack)
    if args.persistent >= 0:
        ctx.set_option(OPT_PERSISTENT, args.persistent)
    if args.shadow_queue >= 0:
        ctx.set_option(OPT_SHADOW_QUEUE, args.shadow_queue)
    if args.merged >= 0:
        ctx.set_option(OPT_MERGED_LOADS, args.merged)
    if args.shadow_trace >= 0:
        ctx.set_option(OPT_SHADOW_IN_TRACE, args.shadow_trace)
    if args.trace_waves:
        ctx.set_option(OPT_TRACE_WAVES, args.trace_waves)
    if args.star_overlap >= 0:
        ctx.set_option(OPT_STAR_OVERLAP, args.star_overlap)
    if args.wave_shape >= 0:
        ctx.set_option(OPT_WAVE_SHAPE, args.wave_shape)
    if args.node_width:
        ctx.set_option(OPT_NODE_WIDTH, args.node_width)
    if args.lds_top >= 0:
        ctx.set_option(OPT_LDS_TOP, args.lds_top)
    if args.fused_shade >= 0:
        ctx.set_option(OPT_FUSED_SHADE, args.fused_shade)
    t0 = time.perf_counter()
    ctx.upload(0, scene)
    log(f"uploaded to {ctx.device_name()} in {time.perf_counter() - t0:.2f} s")

    # Instrumented pass (untimed): traversal counters for the algorithmic bytes.
    ctx.set_option(OPT_STATS, 1)
    ctx.render_resident(0, frame, True)
    ctx.sync()
    st = ctx.stats()
    in_trace = ctx.timings()["shadow_path"] == SHADOWS_IN_TRACE
    ctx.set_option(OPT_STATS, 0)
    log(f"counters: {st}")

    # Timed frames: device events only where the roofline needs them (around the primary-ray
    # kernel, on every Nth frame), since each event record idles the device for a few
    # microseconds between the kernels it separates.
    every = max(args.trace_events_every, 1)
    ctx.set_option(OPT_STAGE_EVENTS, args.stage_events)
    for _ in range(args.warmup):
        ctx.render_resident(0, frame)
        ctx.sync()
    stage_keys = ("star_prep_ms", "trace_ms", "shade_ms", "bloom_ms", "composite_ms", "kernel_ms")
    trace_sum, trace_n, level = 0.0, 0, args.stage_events
    t0 = time.perf_counter()
    for i in range(args.steps):
        want = args.stage_events if args.stage_events != 1 or i % every == 0 else 0
        if want != level:
            ctx.set_option(OPT_STAGE_EVENTS, want)
            level = want
        ctx.render_resident(0, frame)
        ctx.sync()
        if level >= 1:
            trace_sum += ctx.timings()["trace_ms"]
            trace_n += 1
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1e3 / args.steps
    value = w * h * args.steps / wall / 1e6
    trace_timed = trace_sum / max(trace_n, 1)
    log(f"{args.steps} steps in {wall:.3f} s: {ms_per_step:.3f} ms/frame, {value:.1f} Mrays/s; "
        f"primary-ray kernel {trace_timed:.4f} ms over {trace_n} timed launches")
    # Stage breakdown (not timed): the same frames with an event at every stage boundary.
    ctx.set_option(OPT_STAGE_EVENTS, 2)
    parts = {k: 0.0 for k in stage_keys}
    n_break = max(args.breakdown_steps, 1)
    t0 = time.perf_counter()
    for i in range(n_break):
        ctx.render_resident(0, frame)
        ctx.sync()
        t = ctx.timings()
        for k in parts:
            parts[k] += t[k]
    ms_break = (time.perf_counter() - t0) * 1e3 / n_break
    parts = {k: v / n_break for k, v in parts.items()}
    log(f"stage breakdown over {n_break} frames ({ms_break:.3f} ms/frame with every stage event): {parts}")
    if args.stage_events == 0:
        trace_timed = parts["trace_ms"]
    ctx.set_option(OPT_STAGE_EVENTS, args.stage_events)

    # End to end through the host boundary, as GpuRenderer::render (the reference's render()
    # includes the readback): frame upload, render, RGB8 into the caller's framebuffer. Median.
    fb = np.zeros(w * h, np.uint32)
    e2e = []
    for i in range(args.e2e_calls + 3):
        t0 = time.perf_counter()
        ctx.render_frame(0, frame, want_hits=False, rgb=fb)
        if i >= 3:
            e2e.append((time.perf_counter() - t0) * 1e3)
    e2e_ms = float(np.median(e2e))
    ctx.set_option(OPT_STAGE_EVENTS, 2)  # the other legs read every stage's time

    fused = parts["shade_ms"] == 0.0  # the primary-ray kernel shaded its own pixels (breakdown frames)
    alg = compulsory_bytes(st, w, h, scene.num_stars, in_trace, fused)
    work = work_bytes(st, w, h, scene.num_stars, in_trace, fused)
    times = {"k_trace": trace_timed} if fused else {"k_trace": trace_timed, "k_shade": parts["shade_ms"]}
    dom = max(times, key=times.get)
    achieved = alg[dom] / (times[dom] * 1e-3) / 1e9
    traffic = None
    l1 = None
    valu = None
    tf = os.path.join(ROOT, "profiles", f"{dom}_large_1280x800_traffic.txt")
    if os.path.exists(tf) and (w, h) == (1280, 800):
        with open(tf) as f:
            prof = json.load(f)
        traffic = prof.get("hbm_bytes_per_launch")
        if prof.get("l1_accesses_per_launch"):
            cycles = times[dom] * 1e-3 * CUS * ENGINE_CLOCK_HZ
            l1 = {"accesses_per_launch": prof["l1_accesses_per_launch"],
                  "accesses_per_cu_cycle": round(prof["l1_accesses_per_launch"] / cycles, 4),
                  "hit_rate": prof.get("l1_hit_rate"),
                  "note": "vector L1 tag lookups per CU per cycle at the 2.4 GHz peak engine clock "
                          "(TCP_TOTAL_CACHE_ACCESSES of the profile over this run's kernel time)"}
        if prof.get("valu_insts_per_launch"):
            # a wave64 VALU instruction occupies its SIMD (16 lanes wide) for 4 cycles
            simd_cycles = times[dom] * 1e-3 * CUS * SIMDS_PER_CU * ENGINE_CLOCK_HZ
            valu = {"insts_per_launch": prof["valu_insts_per_launch"],
                    "issue_frac": round(prof["valu_insts_per_launch"] * 4 / simd_cycles, 4),
                    "note": "SIMD cycles spent issuing VALU instructions (SQ_INSTS_VALU of the profile, 4 cycles "
                            "each) over the SIMD cycles of this run's kernel time at the 2.4 GHz peak clock"}
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                "kernel": dom, "kernel_ms": round(times[dom], 4),
                "algorithmic_bytes_per_launch": int(alg[dom]),
                "algorithmic": "compulsory: each distinct record touched read once, each per-pixel result written "
                               "once (instrumented distinct counts)",
                "traffic_over_algorithmic": round(traffic / alg[dom], 3) if traffic else None,
                "work": {"bytes_per_launch": int(work[dom]),
                         "GBs": round(work[dom] / (times[dom] * 1e-3) / 1e9, 1),
                         "note": "every node / record visit priced with the reference's record sizes; served "
                                 "mostly by the L1, so not bounded by HBM"},
                "l1": l1,
                "valu": valu,
                "other_kernels": {k: {"ms": round(times[k], 4), "algorithmic_bytes": int(alg[k]),
                                      "achieved_GBs": round(alg[k] / (times[k] * 1e-3) / 1e9, 2)}
                                  for k in times if k != dom}}
    assert roofline["frac"] <= 1.0, roofline

    batched = None
    if args.views > 0:
        batched = bench_batch(ctx, scene, args.views, args.batch_steps, seed=2)
        log(f"batched: {batched}")

    adjoint = None
    if args.adjoint_steps > 0:
        adjoint = bench_adjoint(ctx, args.adjoint_steps)
        log(f"adjoint: {adjoint}")

    batch_adjoint = None
    if args.adjoint_steps > 0:
        batch_adjoint = bench_batch_adjoint(ctx, args.adjoint_steps)
        log(f"batched adjoint: {batch_adjoint}")

    extrude = None
    if args.extrude_steps > 0:
        extrude = bench_extrude(ctx, p, args.extrude_steps)
        log(f"extrusion: {extrude}")

    cpu = None
    parity = None
    if not args.no_cpu:
        log("CPU baseline: oracle on a row sample")
        cpu, checked = cpu_baseline(scene, frame, w, h, args.cpu_rows)
        log(f"cpu baseline {cpu}")
        parity = parity_rows(ctx, frame, checked)
        log(f"parity against the oracle's rows: {parity}")

    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"large single-view forward, {w}x{h}, {scene.num_prims} tets "
                               f"({args.objects} icospheres x {args.slices} slices), {scene.num_stars} stars, "
                               f"rapidity-1 observer, null shadows",
                   "width": w, "height": h, "tets": scene.num_prims, "nodes": scene.num_nodes,
                   "stars": scene.num_stars},
        "kernel_ms_per_step": {k: round(v, 4) for k, v in parts.items()},
        "stage_events": {"timed_frames": args.stage_events, "trace_timed_launches": trace_n,
                         "breakdown_frames": n_break, "breakdown_ms_per_step": round(ms_break, 4),
                         "note": "value and roofline.kernel_ms come from the timed frames (at 1: a pair of events "
                                 "around the primary-ray kernel on every Nth frame, none on the others); "
                                 "kernel_ms_per_step from the breakdown frames after them, with an event at "
                                 "every stage boundary"},
        "shadow_rays_traced_in": "k_trace" if in_trace else "k_shade",
        "pixels_shaded_in": "k_trace" if fused else "k_shade",
        "e2e_ms_with_readback": round(e2e_ms, 3),
        "e2e_calls": len(e2e),
        "parity": parity,
        "counters": st,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "batched": batched,
        "adjoint": adjoint,
        "batched_adjoint": batch_adjoint,
        "extrude": extrude,
    }
    if cpu:
        out["speedup_vs_cpu"] = round(value / cpu["value"], 1)
    ctx.close()
    print(json.dumps(out), flush=True)
    if parity is not None and not parity["ok"]:
        log("PARITY FAILURE: the GPU frame differs from the oracle's rows")
        sys.exit(1)


if __name__ == "__main__":
    main()
