"""Benchmark: batched stroke and batched fill tessellation on one MI355X.

Headline workload (batched stroke): 10,000 open paths with log-uniform segment counts in
[8, 256] (25% lines / 25% quadratics / 50% cubics), points uniform in [0, 2048]^2,
tolerance 0.05, per-path width U[0.5, 8], join uniform over {Miter, Round, Bevel}, caps
uniform over {Butt, Square, Round}; output f32 positions + u32 indices.  A step = one
batched pass over the 10,000 paths (staging-room kernel and scan, one emit kernel into
per-path staging, offset scan, compaction copy into the final arrays); ``value`` = output
triangles per second over the timed steps.  Inputs are synthetic (seeded
generator) and resident in HBM before the timed region.

The same JSON line carries:
  * plain_per_path: the plain per-path drop-in (host buffers in/out, one call per path) on
    a sample of the paths, and the batched/plain ratio;
  * flatten: lyon_geom flattening of the stroke workload's segments with device-resident
    inputs and outputs (points/s and the emit kernel's roofline);
  * validators: lyon_algorithms' approximate_signed_area over the stroke workload's paths
    and hit-test winding numbers of 100,000 query points, device memory in and out;
  * ingest: the stroke workload written as SVG path data and parsed on the device
    (lyon_extra's PathParser), text and paths in device memory;
  * fill: the batched fill workload (4,096 closed paths, half self-intersecting star
    polygons with 5-512 points, half loops of 4-256 cubics, EvenOdd/NonZero 50/50,
    tolerance 0.05, coordinates in [0, 4096]^2) timed the same way, with its plain
    per-path sample;
  * roofline: algorithmic bytes of the stroke emit kernel (1 B/verb + 8 B/point +
    8 B/vertex + 4 B/index) over its average duration (HIP events on the library's
    stream), as a fraction of the spec peak and of this device's measured copy rate (a
    1 GiB device copy, read + write); traffic from rocprofv3 counters when given
    (--traffic-json);
  * serving: the serving loop (continuous batching, 4 in-flight batches) under seeded
    Poisson request streams, with p50/p99/max request latency: short strokes, a 4-256
    segment mix, and the serving configuration (100k req/s offered, log-uniform 4-65536
    segments, 60% strokes / 40% random-walk fills, 32 GiB pool);
  * stroke_variable_width: the stroke workload's paths through the variable-width stroker
    (a width factor per point), and stroke_graph_replay / fill.graph: HIP-graph replays;
  * cpu_baseline: the oracle (C++ restatement of lyon's tessellators) on the box's host
    cores over a bounded sample of the same workloads, single thread and all cores.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def kernel_times(L, ctx):
    a = C.c_float(0)
    b = C.c_float(0)
    L.lyg_ctx_kernel_times(ctx, C.byref(a), C.byref(b))
    return a.value, b.value


def plain_rate(fn, ctx, paths, copts, vcap, icap):
    """Triangles/s of the plain per-path drop-in: one C ABI call per path (lyg_stroke_path /
    lyg_fill_path), host buffers in and out, preallocated once (as a caller reusing its
    VertexBuffers would)."""
    import numpy as np
    from lyon_amd import _lib as B
    from lyon_amd.tessellation import _path_view
    v = np.empty((vcap, 2), np.float32)
    ix = np.empty(icap, np.uint32)
    views = [_path_view(p) for p in paths]

    def one(k):
        out = B.OutputC(v.ctypes.data_as(B.F32P), ix.ctypes.data_as(B.U32P), vcap, icap, 0, 0, 0)
        st = fn(ctx, C.byref(views[k]), C.byref(copts[k]), C.byref(out))
        if st != 0:
            raise RuntimeError(f"plain call: status {st}")
        return out.n_indices // 3

    for k in range(min(5, len(paths))):
        one(k)
    tris = 0
    t0 = time.perf_counter()
    for k in range(len(paths)):
        tris += one(k)
        if k % 64 == 63:
            log(f"  plain call {k + 1}/{len(paths)}")
    return tris / (time.perf_counter() - t0)


def median_time(fn, reps=5):
    """Median wall time of reps calls after one warm-up call; returns (seconds, result)."""
    r = fn()
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        r = fn()
        ts.append(time.perf_counter() - t0)
    ts.sort()
    return ts[len(ts) // 2], r


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--paths", type=int, default=10_000)
    ap.add_argument("--fill-paths", type=int, default=4096)
    ap.add_argument("--fill-steps", type=int, default=5)
    ap.add_argument("--plain-sample", type=int, default=300)
    ap.add_argument("--fill-plain-sample", type=int, default=48)
    ap.add_argument("--cpu-sample", type=int, default=400)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fill", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-flatten", action="store_true")
    ap.add_argument("--no-validate", action="store_true")
    ap.add_argument("--no-ingest", action="store_true")
    ap.add_argument("--no-large", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=4000,
                    help="requests of the stream at the serving configuration")
    ap.add_argument("--serve-split", type=int, default=4,
                    help="windows per part of split strokes in the serving streams' second run")
    ap.add_argument("--large-vertices", type=int, default=1 << 20)
    ap.add_argument("--large-steps", type=int, default=3)
    ap.add_argument("--traffic-json",
                    default=os.path.join(ROOT, "profiles", "stroke_emit_batched_stroke_traffic.json"),
                    help="HBM traffic per emit launch measured with rocprofv3 counters")
    ap.add_argument("--profile", action="store_true",
                    help="only run the batched stroke steps and a 1 GiB device copy (the "
                         "reference kernel for the traffic counters), for rocprofv3")
    ap.add_argument("--profile-large", action="store_true",
                    help="only run the single large fill (banded) steps, for rocprofv3")
    ap.add_argument("--profile-vw", action="store_true",
                    help="only run the variable-width stroke steps, for rocprofv3")
    ap.add_argument("--profile-fill", action="store_true",
                    help="only run the batched fill steps and a 1 GiB device copy, for rocprofv3")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the sharding of this "
              "path across GPUs is described in DESIGN.md and not run.", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    from lyon_amd import (BuffersBuilder, FillOptions, FillRule, FillTessellator, StrokeOptions,
                          StrokeTessellator, VertexBuffers, workloads)
    from lyon_amd import _lib as B
    from lyon_amd.batch import DeviceBatch, FillBatcher, StrokeBatcher, synchronize
    from lyon_amd.tessellation import _Ctx

    L = B.lib()
    ctx = _Ctx.get()
    L.lyg_ctx_enable_timing(ctx, 1)

    if args.profile_vw:
        from lyon_amd.batch import StrokeVWBatcher
        batch = workloads.stroke_batch(args.paths, 8, 256, 2048.0, seed=0)
        opts = workloads.stroke_options(args.paths, tolerance=0.05, w_min=0.5, w_max=8.0)
        db = DeviceBatch(batch, opts)
        widths = np.random.default_rng(12345).uniform(0.5, 2.0, len(batch.points))
        vwb = StrokeVWBatcher(widths.astype(np.float32))
        for k in range(args.steps):
            vwb.run(db, check=False)
            synchronize()
            log(f"variable-width step {k} done")
        return

    if args.profile:
        batch = workloads.stroke_batch(args.paths, 8, 256, 2048.0, seed=0)
        opts = workloads.stroke_options(args.paths, tolerance=0.05, w_min=0.5, w_max=8.0)
        db = DeviceBatch(batch, opts)
        batcher = StrokeBatcher()
        from lyon_amd.batch import DeviceBuffer
        src = DeviceBuffer(1 << 30)
        dst = DeviceBuffer(1 << 30)
        for k in range(args.steps):
            batcher.run(db)
            L.lyg_device_copy(ctx, dst.ptr, src.ptr, 1 << 30)
            synchronize()
            log(f"profile step {k} done")
        print(json.dumps({"profile": "stroke batch + 1 GiB copy", "steps": args.steps}))
        return

    if args.profile_large:
        from lyon_amd.path import PathBatch
        walk = workloads.random_walk_path(args.large_vertices, seed=0)
        wdb = DeviceBatch(PathBatch.from_paths([walk]), workloads.one_fill_options(0.1, 1))
        lb = FillBatcher(banded=True)
        for k in range(args.steps):
            lb.run(wdb)
            synchronize()
            log(f"profile large fill step {k} done")
        print(json.dumps({"profile": "single large fill (banded, NonZero)", "steps": args.steps}))
        return

    if args.profile_fill:
        fbatch = workloads.fill_batch(args.fill_paths, seed=0)
        fopts = workloads.fill_options(args.fill_paths, tolerance=0.05)
        fdb = DeviceBatch(fbatch, fopts)
        fb = FillBatcher()
        from lyon_amd.batch import DeviceBuffer
        src = DeviceBuffer(1 << 30)
        dst = DeviceBuffer(1 << 30)
        for k in range(args.steps):
            fb.run(fdb, check=False)
            L.lyg_device_copy(ctx, dst.ptr, src.ptr, 1 << 30)
            synchronize()
            log(f"profile fill step {k} done")
        print(json.dumps({"profile": "fill batch + 1 GiB copy", "steps": args.steps}))
        return

    # ---- batched stroke (headline) ----------------------------------------------------
    log("generating stroke workload")
    batch = workloads.stroke_batch(args.paths, 8, 256, 2048.0, seed=0)
    opts = workloads.stroke_options(args.paths, tolerance=0.05, w_min=0.5, w_max=8.0)
    db = DeviceBatch(batch, opts)
    batcher = StrokeBatcher()
    log("stroke warmup")
    res = None
    for _ in range(max(args.warmup, 1)):
        res = batcher.run(db)
    synchronize()
    tris_per_step = res.n_indices // 3
    n_vertices, n_indices = res.n_vertices, res.n_indices
    log(f"stroke: timing {args.steps} steps")
    emit_ms, count_ms = [], []
    synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        batcher.run(db)
        a, b = kernel_times(L, ctx)
        count_ms.append(a)
        emit_ms.append(b)
    synchronize()
    ms_per_step = (time.perf_counter() - t0) * 1e3 / args.steps
    value = tris_per_step / (ms_per_step / 1e3)
    alg_bytes = len(batch.verbs) + 8 * len(batch.points) + 8 * n_vertices + 4 * n_indices
    emit_avg = float(np.mean(emit_ms))
    achieved = alg_bytes / (emit_avg / 1e3) / 1e9

    # the same step with long paths split into parts walked by different waves
    from lyon_amd.batch import set_stroke_split
    split_runs = []
    for w in (0, 4, 16):
        set_stroke_split(w)
        batcher.run(db)
        synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            batcher.run(db)
        synchronize()
        sms = (time.perf_counter() - t0) * 1e3 / args.steps
        split_runs.append({"windows_per_part": w, "ms_per_step": sms,
                           "value": tris_per_step / (sms / 1e3)})
        log(f"  stroke split into parts of {w} windows (0: whole paths): {sms:.2f} ms")
    set_stroke_split(64)

    # the same step captured once into a HIP graph (the staged pipeline on fixed buffers)
    # and replayed with no host synchronisation inside
    from lyon_amd.batch import StrokeGraph
    graph = None
    try:
        g = StrokeGraph(db, n_vertices, n_indices)
        g.launch()
        gres = g.result()
        ok = gres.n_indices == n_indices and gres.n_vertices == n_vertices
        t0 = time.perf_counter()
        for _ in range(args.steps):
            g.launch()
        g.result()
        gms = (time.perf_counter() - t0) * 1e3 / args.steps
        graph = {"ms_per_step": gms, "value": tris_per_step / (gms / 1e3),
                 "same_counts_as_staged": bool(ok),
                 "note": "lyg_stroke_batch_graph: the default staged pipeline (plan, scans, "
                         "cut, heads, emit of paths and parts into staging, check, scans, "
                         "compaction, second walk of any path that needs it) captured on fixed "
                         "output buffers, replayed back to back"}
        log(f"  stroke graph replay: {gms:.2f} ms")
        del g
    except Exception as e:  # reported, not fatal: the headline does not depend on it
        graph = {"error": str(e)}

    # the same paths and options through the variable-width stroker, every endpoint with
    # its own width factor (seeded U[0.5, 2] per point)
    vw_run = None
    try:
        from lyon_amd.batch import StrokeVWBatcher
        wrng = np.random.default_rng(12345)
        widths = wrng.uniform(0.5, 2.0, len(batch.points)).astype(np.float32)
        vwb = StrokeVWBatcher(widths)
        vres = vwb.run(db, check=False)
        synchronize()
        vts = []
        for _ in range(max(3, min(args.steps, 10))):
            t0 = time.perf_counter()
            vwb.run(db, check=False)
            synchronize()
            vts.append((time.perf_counter() - t0) * 1e3)
        vms = float(np.median(vts))
        vtris = vres.n_indices // 3
        vw_run = {"ms_per_step": vms, "triangles_per_step": vtris,
                  "value": vtris / (vms / 1e3), "unit": "triangles/s",
                  "failed_paths": int(np.count_nonzero(vres.status)),
                  "note": "lyg_stroke_batch_vw over the stroke workload's paths and options, "
                          "width factor per point U[0.5, 2]: one GPU lane per path (count "
                          "pass, scans, emit pass); median of steps"}
        log(f"  variable-width stroke: {vms:.2f} ms, {vtris} triangles")
        del vwb
    except Exception as e:  # reported, not fatal: the headline does not depend on it
        vw_run = {"error": str(e)}

    # measured copy rate of this device (1 GiB device-to-device copy kernel, read + write)
    from lyon_amd.batch import DeviceBuffer as _DB
    _src, _dst = _DB(1 << 30), _DB(1 << 30)
    L.lyg_device_copy(ctx, _dst.ptr, _src.ptr, 1 << 30)
    synchronize()
    t0 = time.perf_counter()
    for _ in range(5):
        L.lyg_device_copy(ctx, _dst.ptr, _src.ptr, 1 << 30)
    synchronize()
    copy_gbs = 2.0 * (1 << 30) * 5 / (time.perf_counter() - t0) / 1e9
    del _src, _dst

    log("stroke: plain per-path drop-in")
    ns = min(args.plain_sample, args.paths)
    sos = [StrokeOptions(line_width=float(o["line_width"]), miter_limit=float(o["miter_limit"]),
                         tolerance=float(o["tolerance"]), start_cap=int(o["start_cap"]),
                         end_cap=int(o["end_cap"]), line_join=int(o["line_join"])).to_c()
           for o in opts[:ns]]
    vmax = int(np.max(np.diff(res.vertex_offsets.astype(np.int64))[:ns]))
    imax = int(np.max(np.diff(res.index_offsets.astype(np.int64))[:ns]))
    plain_value = plain_rate(L.lyg_stroke_path, ctx, [batch.path(i) for i in range(ns)], sos,
                             vmax, imax)

    # ---- batched flattening (lyon_geom segments) -----------------------------------------
    flat = None
    if not args.no_flatten:
        from lyon_amd.batch import DeviceBuffer, flatten_segments_device
        log("flatten: the segments of the stroke workload")
        kinds, ctrl = workloads.segments_of(batch)
        nseg = len(kinds)
        dk = DeviceBuffer.from_numpy(kinds)
        dc = DeviceBuffer.from_numpy(ctrl.reshape(-1))
        doff = DeviceBuffer(4 * (nseg + 1))
        try:
            flatten_segments_device(dk, dc, nseg, 0.05, None, doff)
            fl_total = 0
        except Exception as e:
            fl_total = getattr(e, "needed", 0)
        dout = DeviceBuffer(8 * max(fl_total, 1))
        for _ in range(max(args.warmup, 1)):
            flatten_segments_device(dk, dc, nseg, 0.05, dout, doff)
        fl_emit, fl_count = [], []
        t0 = time.perf_counter()
        for _ in range(args.steps):
            flatten_segments_device(dk, dc, nseg, 0.05, dout, doff)
            a, b = kernel_times(L, ctx)
            fl_count.append(a)
            fl_emit.append(b)
        fl_ms = (time.perf_counter() - t0) * 1e3 / args.steps
        fl_bytes = 40 * nseg + 8 * fl_total  # kind + 8 floats + offset per segment, 8 B/point
        fl_emit_ms = float(np.mean(fl_emit))
        fl_ach = fl_bytes / (fl_emit_ms / 1e3) / 1e9
        flat = {"workload": f"the {nseg} line/quadratic/cubic segments of the stroke workload, "
                            "tol 0.05 (for_each_flattened_with_t, device memory in and out)",
                "segments": nseg, "points_per_step": fl_total,
                "value": fl_total / (fl_ms / 1e3), "unit": "points/s", "ms_per_step": fl_ms,
                "kernel_ms": {"count_scan_avg": float(np.mean(fl_count)), "emit_avg": fl_emit_ms},
                "roofline": {"bound": "hbm", "achieved": fl_ach, "peak": HBM_PEAK_GBS,
                             "unit": "GB/s", "frac": fl_ach / HBM_PEAK_GBS,
                             "kernel": "flatten_runs_table_kernel + flatten_emit_table_kernel",
                             "algorithmic_bytes": fl_bytes}}
        del dk, dc, doff, dout

    # ---- path validators (lyon_algorithms hit_test / area) -----------------------------
    validators = None
    if not args.no_validate:
        from lyon_amd import algorithms as A
        from lyon_amd.batch import DeviceBuffer
        log("validators: signed areas and winding numbers over the stroke workload")
        rng = np.random.default_rng(0)
        nq = 100_000
        qp = rng.integers(0, args.paths, nq).astype(np.uint32)
        xy = (rng.random((nq, 2)) * 2048.0).astype(np.float32)
        dq, dp = DeviceBuffer.from_numpy(xy.reshape(-1)), DeviceBuffer.from_numpy(qp)
        dw, da = DeviceBuffer(4 * nq), DeviceBuffer(4 * args.paths)
        A.signed_areas_device(db, 0.05, da)
        A.winding_numbers_device(db, dq, dp, nq, 0.05, FillRule.NonZero, dw)
        reps = 5
        t0 = time.perf_counter()
        for _ in range(reps):
            A.signed_areas_device(db, 0.05, da)
        area_ms = (time.perf_counter() - t0) * 1e3 / reps
        t0 = time.perf_counter()
        for _ in range(reps):
            A.winding_numbers_device(db, dq, dp, nq, 0.05, FillRule.NonZero, dw)
        wind_ms = (time.perf_counter() - t0) * 1e3 / reps
        validators = {
            "workload": f"approximate_signed_area of the {args.paths} stroke-workload paths; "
                        f"{nq} winding-number queries (uniform points in [0,2048]^2, each "
                        "against a random path of the batch), tol 0.05, device memory in/out",
            "signed_areas": {"ms": area_ms, "paths_per_s": args.paths / (area_ms / 1e3)},
            "winding_numbers": {"ms": wind_ms, "queries_per_s": nq / (wind_ms / 1e3)}}
        del dq, dp, dw, da

    # ---- SVG path data ingest (lyon_extra PathParser) ---------------------------------
    ingest = None
    if not args.no_ingest:
        from lyon_amd import svg as S
        from lyon_amd.batch import DeviceBuffer
        log("ingest: the stroke workload as SVG path data")
        pts9 = np.char.mod("%.9g", batch.points.astype(np.float64)).reshape(-1, 2)
        pairs = np.char.add(np.char.add(pts9[:, 0], " "), pts9[:, 1])
        words = {0: "L ", 1: "Q ", 2: "C ", 3: "M "}
        strings = []
        for i in range(args.paths):
            v0, v1 = int(batch.verb_offsets[i]), int(batch.verb_offsets[i + 1])
            pi, parts = int(batch.point_offsets[i]), []
            for v in batch.verbs[v0:v1]:
                k = {0: 1, 1: 2, 2: 3, 3: 1}.get(int(v), 0)
                if k:
                    parts.append(words[int(v)] + " ".join(pairs[pi:pi + k]))
                pi += k + (1 if int(v) == 4 else 0)
            strings.append(" ".join(parts))
        text, offs = S.pack_strings(strings)
        td, od = DeviceBuffer.from_numpy(text), DeviceBuffer.from_numpy(offs)
        r = S.parse_paths_device(td, od, args.paths)
        reps = 5
        t0 = time.perf_counter()
        for _ in range(reps):
            r = S.parse_paths_device(td, od, args.paths, out=r)
        ing_ms = (time.perf_counter() - t0) * 1e3 / reps
        ok = (r.n_verbs == len(batch.verbs) and r.n_points == len(batch.points) and
              r.n_errors == 0)
        ingest = {"workload": f"the {args.paths} stroke-workload paths as absolute SVG path "
                              "data (9 significant digits), parsed on the device (one parse "
                              "pass into per-string staging, offset scans, compaction) into the "
                              "previous call's buffers, text and paths in device memory",
                  "text_bytes": int(len(text)), "ms": ing_ms,
                  "paths_per_s": args.paths / (ing_ms / 1e3),
                  "text_gb_per_s": len(text) / (ing_ms / 1e3) / 1e9, "round_trip_counts_ok": ok}
        del td, od, r

    # ---- batched fill -------------------------------------------------------------------
    fill = None
    fbatch = fopts = None
    if not args.no_fill:
        log("generating fill workload")
        fbatch = workloads.fill_batch(args.fill_paths, seed=0)
        fopts = workloads.fill_options(args.fill_paths, tolerance=0.05)
        fdb = DeviceBatch(fbatch, fopts)
        fb = FillBatcher()
        log("fill warmup")
        fres = fb.run(fdb, check=False)
        synchronize()
        ftris = fres.n_indices // 3
        fstatus = fres.status
        log(f"fill: timing {args.fill_steps} steps")
        sweep_ms, copy_ms = [], []
        t0 = time.perf_counter()
        for k in range(args.fill_steps):
            fb.run(fdb, check=False)
            a, b = kernel_times(L, ctx)
            copy_ms.append(a)
            sweep_ms.append(b)
            log(f"  fill step {k} done")
        synchronize()
        fms = (time.perf_counter() - t0) * 1e3 / args.fill_steps
        fns = min(args.fill_plain_sample, args.fill_paths)
        fos = [FillOptions(tolerance=0.05, fill_rule=FillRule(int(o["fill_rule"]))).to_c()
               for o in fopts[:fns]]
        log("fill: plain per-path drop-in")
        fvmax = int(np.max(np.diff(fres.vertex_offsets.astype(np.int64))[:fns]))
        fimax = int(np.max(np.diff(fres.index_offsets.astype(np.int64))[:fns]))
        fplain = plain_rate(L.lyg_fill_path, ctx, [fbatch.path(i) for i in range(fns)], fos,
                            fvmax, fimax)
        fill = {"workload": f"batched fill: {args.fill_paths} closed paths, half "
                            "self-intersecting star polygons {n/k} (5-512 points), half loops "
                            "of 4-256 cubics, EvenOdd/NonZero 50/50, tol 0.05, coords in "
                            "[0,4096]^2",
                "value": ftris / (fms / 1e3), "unit": "triangles/s", "ms_per_step": fms,
                "paths_per_s": args.fill_paths / (fms / 1e3),
                "triangles_per_step": ftris, "vertices_per_step": fres.n_vertices,
                "failed_paths": int(np.count_nonzero(fstatus)),
                "kernel_ms": {"sweep_round0_avg": float(np.mean(sweep_ms)),
                              "copy_avg": float(np.mean(copy_ms))},
                "plain_per_path": {"value": fplain, "unit": "triangles/s",
                                   "sample_paths": fns},
                "batched_over_plain": (ftris / (fms / 1e3)) / fplain}
        fbytes = len(fbatch.verbs) + 8 * len(fbatch.points) + 8 * fres.n_vertices + \
            4 * fres.n_indices
        fach = fbytes / (fms / 1e3) / 1e9
        fill["roofline"] = {"bound": "hbm", "achieved": fach, "peak": HBM_PEAK_GBS,
                            "unit": "GB/s", "frac": fach / HBM_PEAK_GBS,
                            "scope": "whole batched step (sweep launches, scans, copy)",
                            "algorithmic_bytes": fbytes}
        # the same step's first (and, for this workload, only) sweep round captured into
        # HIP graphs and replayed with no host synchronisation inside
        from lyon_amd.batch import FillGraph
        try:
            fg = FillGraph(fdb, fres.n_vertices, fres.n_indices)
            fg.launch()
            gres = fg.result(check=False)
            same_counts = gres.n_indices == fres.n_indices and \
                gres.n_vertices == fres.n_vertices
            # replays and eager steps alternate, so that both see the same clocks
            gts, ets = [], []
            for _ in range(args.fill_steps):
                t0 = time.perf_counter()
                fg.launch()
                fg.result(check=False)
                gts.append((time.perf_counter() - t0) * 1e3)
                t0 = time.perf_counter()
                fb.run(fdb, check=False)
                synchronize()
                ets.append((time.perf_counter() - t0) * 1e3)
            fgms = float(np.median(gts))
            fill["graph"] = {"ms_per_step": fgms, "value": ftris / (fgms / 1e3),
                             "eager_ms_same_window": float(np.median(ets)),
                             "same_counts_as_eager": bool(same_counts),
                             "note": "lyg_fill_batch_graph: the sweep round planned at capture "
                                     "(one graph per class stream, launched on the context's "
                                     "side streams), offset scans and copy, on fixed output "
                                     "buffers; median of replays (launch + result), each "
                                     "followed by an eager step, whose median is beside it"}
            log(f"  fill graph replay: {fgms:.2f} ms (eager in the same window "
                f"{np.median(ets):.2f} ms)")
            del fg
        except Exception as e:  # reported, not fatal: the headline does not depend on it
            fill["graph"] = {"error": str(e)}
        # the same workload through the banded fill (lyon's coverage, its own triangles),
        # at each band size; the fastest is reported
        from lyon_amd.batch import set_band_size
        bb = FillBatcher(banded=True)
        banded_runs = []
        for epb in (32, 64, 128, 256):
            set_band_size(epb, 16384)
            bres = bb.run(fdb, check=False)
            synchronize()
            ts = []
            for _ in range(args.fill_steps):
                t0 = time.perf_counter()
                bb.run(fdb, check=False)
                synchronize()
                ts.append((time.perf_counter() - t0) * 1e3)
            bms = float(np.median(ts))
            btris = bres.n_indices // 3
            bbytes = len(fbatch.verbs) + 8 * len(fbatch.points) + 8 * bres.n_vertices + \
                4 * bres.n_indices
            banded_runs.append({"edges_per_band": epb, "ms_per_step": bms,
                                "triangles_per_step": btris, "value": btris / (bms / 1e3),
                                "paths_per_s": args.fill_paths / (bms / 1e3),
                                "failed_paths": int(np.count_nonzero(bres.status)),
                                "algorithmic_bytes": bbytes})
            log(f"  banded fill, {epb} edges per band: {bms:.2f} ms")
        set_band_size(256, 16384)
        best = min(banded_runs, key=lambda r: r["ms_per_step"])
        bach = best["algorithmic_bytes"] / (best["ms_per_step"] / 1e3) / 1e9
        fill["banded"] = {
            "note": "the same paths and options through lyg_fill_batch_banded: the covered "
                    "region of lyon's fill, trapezoid triangles (not lyon's); median of "
                    f"{args.fill_steps} steps per band size",
            "best": best, "by_band_size": banded_runs,
            "paths_per_s": best["paths_per_s"],
            "over_lyon_triangles_form": fms / best["ms_per_step"],
            "roofline": {"bound": "hbm", "achieved": bach, "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": bach / HBM_PEAK_GBS,
                         "scope": "whole banded step", "algorithmic_bytes":
                             best["algorithmic_bytes"]}}

    # ---- single large fill (banded) -----------------------------------------------------------
    large = None
    if not args.no_large:
        from lyon_amd.path import PathBatch
        log("single large fill: generating the random walk")
        walk = workloads.random_walk_path(args.large_vertices, seed=0)
        wbatch = PathBatch.from_paths([walk])
        large = {"workload": f"single large fill: one closed random-walk polyline of "
                             f"{args.large_vertices} vertices (N(0,1) steps, f32), rescaled to "
                             "[0,8192]^2, tolerance 0.1, banded fill (coverage of lyon's fill)",
                 "vertices": args.large_vertices}
        wdb = DeviceBatch(wbatch, workloads.one_fill_options(0.1, 1))
        lb = FillBatcher(banded=True)
        for rule, rname in ((1, "nonzero"), (0, "evenodd")):
            wdb = DeviceBatch(wbatch, workloads.one_fill_options(0.1, rule))
            wres = lb.run(wdb)
            synchronize()
            prep, sweep = [], []
            t0 = time.perf_counter()
            for _ in range(args.large_steps):
                lb.run(wdb)
                a, b = kernel_times(L, ctx)
                prep.append(a)
                sweep.append(b)
            synchronize()
            lms = (time.perf_counter() - t0) * 1e3 / args.large_steps
            ltris = wres.n_indices // 3
            lbytes = len(wbatch.verbs) + 8 * len(wbatch.points) + 8 * wres.n_vertices + \
                4 * wres.n_indices
            lach = lbytes / (lms / 1e3) / 1e9
            large[rname] = {"ms_per_step": lms, "triangles": ltris,
                            "vertices": wres.n_vertices, "value": ltris / (lms / 1e3),
                            "unit": "triangles/s",
                            "kernel_ms": {"band_prep_avg": float(np.mean(prep)),
                                          "sweep_avg": float(np.mean(sweep))},
                            "roofline": {"bound": "hbm", "achieved": lach, "peak": HBM_PEAK_GBS,
                                         "unit": "GB/s", "frac": lach / HBM_PEAK_GBS,
                                         "scope": "whole banded step (all kernels and host "
                                                  "syncs)", "algorithmic_bytes": lbytes}}
            log(f"  {rname}: {lms:.2f} ms, {ltris} triangles")
        if not args.no_cpu:
            from oracle import pyoracle as O
            sample_n = 1 << 16
            sw = workloads.random_walk_path(sample_n, seed=0)
            secs, (st, rv, ri) = median_time(lambda: O.fill(sw, 0.1, 1), reps=3)
            large["cpu_baseline"] = {
                "value": (len(ri) // 3) / secs, "unit": "triangles/s", "cores": 1,
                "kind": "port", "sample": f"the restated lyon sweep on a {sample_n}-vertex walk "
                "of the same generator, NonZero (median of 3 after a warm-up; the full walk "
                "takes ~65 s per rule)", "seconds": secs}

    # ---- plain workload (one closed 64-segment path) ------------------------------------------
    plain = None
    if not args.no_cpu:
        from oracle import pyoracle as O
        log("plain workload: GPU drop-in and oracle")
        pp = workloads.plain_path(0)
        sopt = StrokeOptions(line_width=4.0, miter_limit=4.0, line_join=0, start_cap=0,
                             end_cap=0).to_c()
        fopt = FillOptions(tolerance=0.1).to_c()
        from lyon_amd.tessellation import _path_view
        view = _path_view(pp)
        pvb = np.empty((1 << 16, 2), np.float32)
        pib = np.empty(1 << 18, np.uint32)

        def one_call(fn, copt):
            out = B.OutputC(pvb.ctypes.data_as(B.F32P), pib.ctypes.data_as(B.U32P), 1 << 16,
                            1 << 18, 0, 0, 0)
            st = fn(ctx, C.byref(view), C.byref(copt), C.byref(out))
            if st != 0:
                raise RuntimeError(f"plain workload: status {st}")
            return out.n_indices // 3

        gs, _ = median_time(lambda: one_call(L.lyg_stroke_path, sopt), reps=9)
        gf, _ = median_time(lambda: one_call(L.lyg_fill_path, fopt), reps=9)
        cs, (_, _, sri) = median_time(lambda: O.stroke(pp, 4.0, 4.0, 0.1, 0, 0, 0), reps=9)
        cf, (_, _, fri) = median_time(lambda: O.fill(pp, 0.1, 0), reps=9)
        plain = {"workload": "one closed path of 64 segments (25% line / 25% quadratic / 50% "
                             "cubic, points U[0,1024]^2, seed 0): stroke width 4, Miter limit 4, "
                             "Butt, tolerance 0.1; fill EvenOdd, tolerance 0.1",
                 "stroke": {"triangles": len(sri) // 3, "gpu_ms": gs * 1e3, "cpu_ms": cs * 1e3},
                 "fill": {"triangles": len(fri) // 3, "gpu_ms": gf * 1e3, "cpu_ms": cf * 1e3},
                 "note": "GPU: one lyg_*_path call with host buffers (PCIe round trip); CPU: "
                         "the oracle on one core; medians of 9 after a warm-up"}

    # ---- serving loop: Poisson request streams ---------------------------------------------
    serving = None
    if not args.no_serve:
        from lyon_amd.serve import Server
        serving = []
        # (name, offered req/s, requests, stroke fraction, segments min/max, pool GiB,
        #  fill kind (0 the fill workload's shapes, 1 random paths as the strokes, 2 random
        #  walks), windows per part of split strokes (64: the library's default))
        W = args.serve_split
        streams = [
            ("stroke, 4-64 segments", 100000.0, 100000, 1.0, 4, 64, 8, 0, 64),
            ("stroke, 4-64 segments", 100000.0, 100000, 1.0, 4, 64, 8, 0, W),
            ("stroke, 4-256 segments", 100000.0, 100000, 1.0, 4, 256, 8, 0, 64),
            ("stroke, 4-256 segments", 100000.0, 100000, 1.0, 4, 256, 8, 0, W),
            ("60% stroke / 40% fill, 4-256 segments", 5000.0, 5000, 0.6, 4, 256, 8, 0, W),
            ("serving config: 60% stroke / 40% fill, log-uniform 4-65536 segments "
             "(lines/quads/cubics 25/25/50, points uniform in [0,4096]^2, fills closed, "
             "EvenOdd/NonZero), 32 GiB pool", 100000.0, args.serve_requests, 0.6, 4, 65536, 32,
             1, W),
            ("60% stroke / 40% fill (fills: closed random walks), log-uniform 4-65536 "
             "segments, 32 GiB pool", 100000.0, args.serve_requests, 0.6, 4, 65536, 32, 2, W),
        ]
        os.environ["LYG_STREAM_PROGRESS"] = "1"
        for name, rate, nreq, sfrac, smin, smax, pool_gib, fill_kind, split in streams:
            log(f"serving stream: {name} at {rate:.0f} req/s, stroke split {split}")
            # fills whose estimated crossings exceed 4e6 are refused at admission: their
            # output grows with the square of their segments (random segments across the
            # whole box) and one of them would hold the device for seconds
            with Server(pool_bytes=pool_gib << 30, max_batch_paths=1024, max_inflight=4,
                        fill_banded_min_edges=2048 if fill_kind else 0,
                        stroke_split_windows=split,
                        max_fill_crossings=4_000_000 if fill_kind else 0) as srv:
                # warm-up stream (first launches on the workers' streams, their buffers
                # growing to size), then the measured one on the same server
                srv.run_stream(n_requests=max(nreq // 4, 100), rate_per_s=rate,
                               stroke_frac=sfrac, min_segs=smin, max_segs=smax,
                               tolerance=0.05, n_distinct=256, seed=2, fill_kind=fill_kind)
                r = srv.run_stream(n_requests=nreq, rate_per_s=rate, stroke_frac=sfrac,
                                   min_segs=smin, max_segs=smax, tolerance=0.05,
                                   n_distinct=256, seed=1, fill_kind=fill_kind)
            serving.append({"stream": name, "requests": nreq, "warmup_requests": max(nreq // 4, 100),
                            "pool_gib": pool_gib, "stroke_split_windows": split,
                            "offered_rate": rate, "completed_rate": r["completed_rate"],
                            "failed": r["failed"], "rejected_at_admission": r["rejected"],
                            "p50_us": r["p50_us"], "p99_us": r["p99_us"], "max_us": r["max_us"],
                            "triangles_per_s": r["triangles_per_s"],
                            "mean_batch": r["mean_batch"], "seconds": r["seconds"]})
            log(f"  completed {r['completed_rate']:.0f} req/s, p50 {r['p50_us'] / 1e3:.2f} ms, "
                f"p99 {r['p99_us'] / 1e3:.2f} ms, failed {r['failed']} (rejected {r['rejected']})")

    # ---- CPU baseline ---------------------------------------------------------------------
    cpu = None
    if not args.no_cpu:
        log("cpu baseline (oracle)")
        from oracle import pyoracle as O
        from lyon_amd.path import PathBatch
        ncores = len(os.sched_getaffinity(0))
        m = min(args.cpu_sample, args.paths)
        sub = PathBatch.from_paths([batch.path(i) for i in range(m)])
        so = np.zeros((m, 6), np.float32)
        for j, k in enumerate(("line_width", "miter_limit", "tolerance", "start_cap", "end_cap",
                               "line_join")):
            so[:, j] = opts[k][:m]
        s1, ctris = median_time(lambda: O.batch_triangles(sub, so, False, 1))
        sa, atris = median_time(lambda: O.batch_triangles(sub, so, False, ncores))
        cpu = {"value": ctris / s1, "unit": "triangles/s", "cores": 1, "kind": "port",
               "sample": f"first {m} of the {args.paths} batched-stroke paths, single thread "
                         "(median of 5 after a warm-up)",
               "all_cores_value": atris / sa, "all_cores": ncores, "nproc": os.cpu_count(),
               "cpu_model": cpu_model()}
        if fbatch is not None:
            fm = min(args.cpu_sample, args.fill_paths)
            fsub = PathBatch.from_paths([fbatch.path(i) for i in range(fm)])
            fo = np.zeros((fm, 2), np.float32)  # tolerance, fill rule per path
            fo[:, 0] = 0.05
            fo[:, 1] = fopts["fill_rule"][:fm]
            f1s, fct = median_time(lambda: O.batch_triangles(fsub, fo, True, 1))
            fas, fat = median_time(lambda: O.batch_triangles(fsub, fo, True, ncores))
            cpu["fill"] = {"value": fct / f1s, "unit": "triangles/s", "cores": 1,
                           "sample": f"first {fm} of the {args.fill_paths} batched-fill paths "
                                     "(median of 5 after a warm-up)",
                           "all_cores_value": fat / fas, "all_cores": ncores}

    traffic = None
    if args.traffic_json and os.path.exists(args.traffic_json):
        traffic = json.load(open(args.traffic_json)).get("bytes_per_launch")

    out = {
        "metric": "output triangles/s batched fill+stroke vs plain per-path; p50/p99 req latency",
        "value": value,
        "unit": "triangles/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded SplitMix64/PCG32 generator)",
        "config": {"workload": f"batched stroke: {args.paths} open paths, log-uniform 8-256 "
                               "segments (25% line/25% quad/50% cubic), coords U[0,2048]^2, "
                               "tol 0.05, width U[0.5,8], join {Miter,Round,Bevel}, caps "
                               "{Butt,Square,Round}",
                   "paths": args.paths, "triangles_per_step": tris_per_step,
                   "vertices_per_step": n_vertices, "paths_per_s": args.paths / (ms_per_step / 1e3),
                   "parallelism": "single GPU"},
        "plain_per_path": {"value": plain_value, "unit": "triangles/s",
                           "sample_paths": ns, "batched_over_plain": value / plain_value},
        "kernel_ms": {"sizing_avg": float(np.mean(count_ms)), "emit_avg": emit_avg},
        "stroke_split": split_runs,
        "stroke_graph_replay": graph,
        "stroke_variable_width": vw_run,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "measured_copy_gbs": copy_gbs, "frac_of_measured_copy": achieved / copy_gbs,
                     "kernel": "stroke_kernel<true, true> (emit over paths and parts of long paths)",
                     "algorithmic_bytes": alg_bytes},
        "flatten": flat,
        "validators": validators,
        "ingest": ingest,
        "fill": fill,
        "single_large_fill": large,
        "plain_workload": plain,
        "serving": serving,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
