"""The serving loop: results equal to the plain calls, the same alone and in a mix of
concurrent requests, admission within a small pool, and a Poisson request stream."""
import threading

import numpy as np
import pytest

from helpers import assert_same_geometry, bbox_diag
from lyon_amd import (FillOptions, FillRule, LineCap, LineJoin, StrokeOptions, workloads)
from lyon_amd.serve import Server
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu


def canon(idx):
    t = np.asarray(idx, np.int64).reshape(-1, 3)
    if len(t) == 0:
        return t
    k = np.argmin(t, axis=1)
    r = np.arange(len(t))
    t = np.stack([t[r, k], t[r, (k + 1) % 3], t[r, (k + 2) % 3]], 1)
    return t[np.lexsort(t.T[::-1])]


def requests(n_each=10, seed=11):
    sb = workloads.stroke_batch(n_each, 4, 400, 4096.0, seed=seed)
    so = workloads.stroke_options(n_each, tolerance=0.05, seed=seed)
    fb = workloads.fill_batch(n_each, seed=seed)
    fo = workloads.fill_options(n_each, tolerance=0.05, seed=seed)
    out = []
    for i in range(n_each):
        o = so[i]
        out.append(("stroke", sb.path(i),
                    StrokeOptions(line_width=float(o["line_width"]),
                                  miter_limit=float(o["miter_limit"]),
                                  tolerance=float(o["tolerance"]), start_cap=int(o["start_cap"]),
                                  end_cap=int(o["end_cap"]), line_join=int(o["line_join"]))))
        out.append(("fill", fb.path(i),
                    FillOptions(tolerance=0.05, fill_rule=FillRule(int(fo[i]["fill_rule"])))))
    return out


def submit(s, kind, path, opts):
    return s.submit_stroke(path, opts) if kind == "stroke" else s.submit_fill(path, opts)


def same(a, b):
    return (a[0] == b[0] and (a[0] != 0 or (
        np.array_equal(a[1].view(np.uint32), b[1].view(np.uint32)) and np.array_equal(a[2], b[2]))))


@pytest.mark.parametrize("stroke_sizing", [0, 1])
def test_results_match_oracle(stroke_sizing):
    # stroke_sizing 0: one staged emit pass per batch; 1: count pass then emit
    with Server(pool_bytes=1 << 30, max_batch_paths=64, max_inflight=2,
                stroke_sizing=stroke_sizing) as s:
        for k, (kind, p, o) in enumerate(requests(6, seed=3)):
            t = submit(s, kind, p, o)
            st, v, i = s.result(t)
            s.release(t)
            print(f"{kind} request {k}: status {st}", flush=True)
            if kind == "stroke":
                rst, rv, ri = O.stroke(p, o.line_width, o.miter_limit, o.tolerance,
                                       int(o.start_cap), int(o.end_cap), int(o.line_join))
                assert st == rst
                assert_same_geometry((v, i), (rv, ri), max(bbox_diag(p.points), o.line_width),
                                     what=f"served stroke {k}")
            else:
                rst, rv, ri = O.fill(p, o.tolerance, int(o.fill_rule))
                assert st == rst
                if st == 0:
                    assert np.array_equal(v.view(np.uint32), rv.view(np.uint32))
                    assert np.array_equal(canon(i), canon(ri))


def test_same_alone_and_in_the_mix():
    targets = requests(10, seed=11)
    noise = requests(40, seed=99)
    with Server(pool_bytes=2 << 30, max_batch_paths=32, max_inflight=3) as s:
        alone = []
        for kind, p, o in targets:
            t = submit(s, kind, p, o)
            alone.append(s.result(t))
            s.release(t)
        # the same requests again, interleaved with a stream of others from two threads
        stop = threading.Event()
        noise_tickets = [[], []]

        def spam(j):
            k = 0
            while not stop.is_set() and k < 400:
                kind, p, o = noise[(7 * k + j) % len(noise)]
                noise_tickets[j].append(submit(s, kind, p, o))
                k += 1

        th = [threading.Thread(target=spam, args=(j,)) for j in range(2)]
        for t_ in th:
            t_.start()
        mixed_t = [submit(s, kind, p, o) for kind, p, o in targets]
        mixed = [s.result(t) for t in mixed_t]
        stop.set()
        for t_ in th:
            t_.join()
        for t in mixed_t:
            s.release(t)
        for lst in noise_tickets:
            for t in lst:
                s.wait(t)
                s.release(t)
        st = s.stats()
        print("stats", st, flush=True)
        assert st["mean_batch"] > 1.0, "requests were never batched together"
        for k, (a, b) in enumerate(zip(alone, mixed)):
            assert same(a, b), f"request {k} ({targets[k][0]}) differs alone vs in the mix"


def test_admission_in_a_small_pool():
    pool = 24 << 20
    reqs = requests(12, seed=5)
    with Server(pool_bytes=pool, max_batch_paths=16, max_inflight=2) as s:
        ts = [submit(s, kind, p, o) for kind, p, o in reqs]
        res = []
        for t in ts:  # a client that releases each result once it has read it
            res.append(s.result(t))
            s.release(t)
        st = s.stats()
        assert st["completed"] == len(reqs)
        assert st["pool_high_water"] <= pool
        assert all(r[0] in (0, 5) for r in res)


def test_poisson_stream():
    with Server(pool_bytes=4 << 30, max_batch_paths=256, max_inflight=4) as s:
        r = s.run_stream(n_requests=4000, rate_per_s=20000, stroke_frac=0.6, min_segs=4,
                         max_segs=512, tolerance=0.05, n_distinct=64, seed=1)
        print("stream", r, flush=True)
        assert r["requests"] == 4000
        assert r["triangles"] > 0
        assert 0 < r["p50_us"] <= r["p99_us"] <= r["max_us"]


def c5_fill_sizes(n, seed):
    """Edge counts drawn log-uniformly from [4, 65536], as the serving stream's fills."""
    rng = np.random.default_rng(seed)
    return np.round(np.exp(rng.uniform(np.log(4), np.log(65536), n))).astype(int)


def test_c5_requests_match_oracle():
    """Requests drawn as the 4-65536-segment serving stream draws them, both fill rules:
    strokes equal to the oracle's; fills below the banded threshold equal to lyon's
    (vertices bit for bit, the same triangles); banded fills with coverage equal to the
    winding number of the flattened edges and the exact area (relative 1e-6)."""
    sb = workloads.stroke_batch(6, 4, 65536, 4096.0, seed=21)
    so = workloads.stroke_options(6, tolerance=0.05, seed=21)
    sizes = list(c5_fill_sizes(6, 5)) + [12, 1500, 40000]
    with Server(pool_bytes=8 << 30, max_batch_paths=64, max_inflight=2,
                fill_banded_min_edges=2048) as s:
        reqs = []
        for i in range(6):
            o = so[i]
            opts = StrokeOptions(line_width=float(o["line_width"]),
                                 miter_limit=float(o["miter_limit"]), tolerance=0.05,
                                 start_cap=int(o["start_cap"]), end_cap=int(o["end_cap"]),
                                 line_join=int(o["line_join"]))
            reqs.append(("stroke", sb.path(i), opts))
        for k, n in enumerate(sizes):
            p = workloads.random_walk_path(int(max(n, 3)), seed=100 + k, extent=4096.0)
            reqs.append(("fill", p, FillOptions(tolerance=0.05, fill_rule=FillRule(k & 1))))
        # fills as the serving configuration draws them (closed random curves), small
        # enough for the oracle's sweep
        fb = workloads.stroke_batch(6, 4, 160, 4096.0, seed=9, closed=True)
        for k in range(6):
            reqs.append(("fill", fb.path(k), FillOptions(tolerance=0.05,
                                                         fill_rule=FillRule(k & 1))))
        tickets = [submit(s, kind, p, o) for kind, p, o in reqs]
        n_banded = 0
        for k, ((kind, p, o), t) in enumerate(zip(reqs, tickets)):
            st, v, i = s.result(t)
            was_banded = s.banded(t)
            s.release(t)
            n_banded += was_banded
            print(f"c5 {kind} request {k}: {len(p.verbs)} verbs, status {st}, "
                  f"banded {was_banded}", flush=True)
            if kind == "stroke":
                rst, rv, ri = O.stroke(p, o.line_width, o.miter_limit, o.tolerance,
                                       int(o.start_cap), int(o.end_cap), int(o.line_join))
                assert st == rst
                assert_same_geometry((v, i), (rv, ri), max(bbox_diag(p.points), o.line_width),
                                     what=f"served c5 stroke {k}")
            elif not was_banded:
                rst, rv, ri = O.fill(p, o.tolerance, int(o.fill_rule))
                assert st == rst
                if st == 0:
                    assert np.array_equal(v.view(np.uint32), rv.view(np.uint32))
                    assert np.array_equal(canon(i), canon(ri))
            else:
                assert st == 0
                rule = int(o.fill_rule)
                r = O.winding_compare(p, 0.05, rule, v, i, n=1024,
                                      eps=float(np.spacing(np.float32(4096.0))))
                assert r["mismatch"] == 0, (k, r)
                exact, _ = O.fill_area(p, 0.05, rule)
                t3 = np.asarray(i, np.int64).reshape(-1, 3)
                vv = np.asarray(v, np.float64)
                a, b, c = vv[t3[:, 0]], vv[t3[:, 1]], vv[t3[:, 2]]
                area = float(np.abs((b - a)[:, 0] * (c - a)[:, 1]
                                    - (b - a)[:, 1] * (c - a)[:, 0]).sum() / 2)
                assert abs(area - exact) <= 1e-6 * exact, (k, area, exact)
        assert n_banded >= 3


def test_c5_stream_random_walk_fills():
    """A short stream at the serving configuration's sizes (log-uniform 4-65536 segments,
    60/40 stroke/fill, 100k req/s offered, 32 GiB pool) with random-walk fills: every
    request completes without error."""
    with Server(pool_bytes=32 << 30, max_batch_paths=256, max_inflight=4,
                fill_banded_min_edges=2048) as s:
        r = s.run_stream(n_requests=2000, rate_per_s=100000, stroke_frac=0.6, min_segs=4,
                         max_segs=65536, tolerance=0.05, n_distinct=128, seed=3, fill_kind=2)
        print("c5 stream (walk fills)", r, flush=True)
        assert r["requests"] == 2000 and r["failed"] == 0
        assert r["triangles"] > 0
        assert 0 < r["p50_us"] <= r["p99_us"] <= r["max_us"]


def test_admission_rejects_beyond_the_worst_case_limit():
    """A fill whose worst-case output (Bezout crossing bound) exceeds the request limit
    completes at once with NO_SPACE and is counted; a small one beside it runs."""
    big = workloads.stroke_batch(1, 60000, 65536, 4096.0, seed=4, closed=True).path(0)
    small = workloads.fill_batch(2, seed=4).path(0)
    with Server(pool_bytes=1 << 30, max_batch_paths=16, max_inflight=1,
                max_request_bytes=64 << 20) as s:
        tb = s.submit_fill(big, FillOptions(tolerance=0.05))
        ts = s.submit_fill(small, FillOptions(tolerance=0.05))
        assert s.result(tb)[0] == 20  # LYG_ERR_NO_SPACE
        st, v, i = s.result(ts)
        assert st == 0 and len(i) > 0
        s.release(tb)
        s.release(ts)
        assert s.stats()["rejected"] == 1
