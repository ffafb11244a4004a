"""Benchmark: batched stroke tessellation on one MI355X.

Workload (batched stroke): 10,000 open paths with log-uniform segment counts in
[8, 256] (25% lines / 25% quadratics / 50% cubics), points uniform in [0, 2048]^2,
tolerance 0.05, per-path width U[0.5, 8], join uniform over {Miter, Round, Bevel}, caps
uniform over {Butt, Square, Round}; output f32 positions + u32 indices.  Data is
synthetic (seeded generator), resident in HBM before the timed region.

A step = one batched pass over the 10,000 paths: count kernel, offset scan, emit kernel.
``value`` = output triangles per second over the timed steps.  The same JSON line carries:
  * plain: the plain per-path drop-in (host buffers in/out, one call per path) on a
    sample of the paths, as triangles/s, and the batched/plain ratio;
  * roofline: algorithmic bytes of the emit kernel (1 B/verb + 8 B/point + 8 B/vertex +
    4 B/index) over its average duration (HIP events on the library's stream);
  * cpu_baseline: the oracle (C++ restatement of lyon's StrokeTessellator) on the box's
    host cores over a bounded sample of the same workload.
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--paths", type=int, default=10_000)
    ap.add_argument("--plain-sample", type=int, default=300)
    ap.add_argument("--cpu-sample", type=int, default=400)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the multi-GPU "
              "sharding of this path is described in DESIGN.md but not run.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from lyon_amd import (BuffersBuilder, StrokeOptions, StrokeTessellator, VertexBuffers,
                          workloads)
    from lyon_amd import _lib as B
    from lyon_amd.batch import DeviceBatch, StrokeBatcher
    from lyon_amd.tessellation import _Ctx

    log("generating workload")
    batch = workloads.stroke_batch(args.paths, 8, 256, 2048.0, seed=0)
    opts = workloads.stroke_options(args.paths, tolerance=0.05, w_min=0.5, w_max=8.0)
    db = DeviceBatch(batch, opts)
    batcher = StrokeBatcher()
    L = B.lib()
    ctx = _Ctx.get()
    L.lyg_ctx_enable_timing(ctx, 1)

    log("warmup")
    res = None
    for _ in range(max(args.warmup, 1)):
        res = batcher.run(db)
    torch.cuda.synchronize()
    tris_per_step = res.n_indices // 3
    n_vertices = res.n_vertices
    n_indices = res.n_indices

    log(f"timing {args.steps} steps")
    emit_ms = []
    count_ms = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        batcher.run(db)
        a = C.c_float(0)
        b = C.c_float(0)
        L.lyg_ctx_kernel_times(ctx, C.byref(a), C.byref(b))
        count_ms.append(a.value)
        emit_ms.append(b.value)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    value = tris_per_step / (ms_per_step / 1e3)

    # roofline of the emit kernel: algorithmic bytes per launch / average duration
    alg_bytes = (len(batch.verbs) * 1 + len(batch.points) * 8 + n_vertices * 8 + n_indices * 4)
    emit_avg = float(np.mean(emit_ms))
    achieved = alg_bytes / (emit_avg / 1e3) / 1e9

    # plain per-path drop-in on a sample (host in/out, one call per path)
    log("plain per-path drop-in")
    tess = StrokeTessellator()
    ns = min(args.plain_sample, args.paths)
    sos = [StrokeOptions(line_width=float(o["line_width"]), miter_limit=float(o["miter_limit"]),
                         tolerance=float(o["tolerance"]), start_cap=int(o["start_cap"]),
                         end_cap=int(o["end_cap"]), line_join=int(o["line_join"]))
           for o in opts[:ns]]
    paths = [batch.path(i) for i in range(ns)]
    for i in range(min(20, ns)):
        tess.tessellate_path(paths[i], sos[i], BuffersBuilder(VertexBuffers()))
    plain_tris = 0
    tp0 = time.perf_counter()
    for i in range(ns):
        vb = VertexBuffers()
        tess.tessellate_path(paths[i], sos[i], BuffersBuilder(vb))
        plain_tris += len(vb.indices) // 3
    tp1 = time.perf_counter()
    plain_value = plain_tris / (tp1 - tp0)

    cpu = None
    if not args.no_cpu:
        log("cpu baseline (oracle)")
        from oracle import pyoracle as O
        from lyon_amd.path import PathBatch
        m = min(args.cpu_sample, args.paths)
        sub = PathBatch.from_paths([batch.path(i) for i in range(m)])
        so = np.zeros((m, 6), np.float32)
        so[:, 0] = opts["line_width"][:m]
        so[:, 1] = opts["miter_limit"][:m]
        so[:, 2] = opts["tolerance"][:m]
        so[:, 3] = opts["start_cap"][:m]
        so[:, 4] = opts["end_cap"][:m]
        so[:, 5] = opts["line_join"][:m]
        O.batch_triangles(sub, so, False, 1)
        c0 = time.perf_counter()
        ctris = O.batch_triangles(sub, so, False, 1)
        c1 = time.perf_counter()
        ncores = len(os.sched_getaffinity(0))
        a0 = time.perf_counter()
        atris = O.batch_triangles(sub, so, False, ncores)
        a1 = time.perf_counter()
        cpu = {"value": ctris / (c1 - c0), "unit": "triangles/s", "cores": 1, "kind": "port",
               "sample": f"first {m} of the {args.paths} batched-stroke paths, single thread",
               "all_cores_value": atris / (a1 - a0), "all_cores": ncores}

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
                   "vertices_per_step": n_vertices, "parallelism": "single GPU"},
        "plain_per_path": {"value": plain_value, "unit": "triangles/s",
                           "sample_paths": ns, "batched_over_plain": value / plain_value},
        "kernel_ms": {"count_avg": float(np.mean(count_ms)), "emit_avg": emit_avg},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "stroke_kernel<emit>", "algorithmic_bytes": alg_bytes},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
