"""Full-size parity of the batched workloads against the oracle run on all host cores.

* Batched stroke, all 10,000 paths of the stroke workload (log-uniform 8-256 segments,
  tolerance 0.05, per-path width/join/cap): per path the index topology is bit-exact and
  every position within 1e-5 of the path's bounding-box diagonal.
* Batched fill, all 4,096 paths of the fill workload (stars and cubic loops, EvenOdd /
  NonZero per path): per path the status, the vertices (bit-exact, in lyon's order) and the
  set of triangles are the oracle's (lyon buffers a span's triangles before writing them, so
  only their order in the index buffer differs).
"""
import numpy as np
import pytest

from lyon_amd import workloads
from lyon_amd.batch import DeviceBatch, FillBatcher, StrokeBatcher
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu


def _path_diags(batch):
    p = np.asarray(batch.points, np.float64).reshape(-1, 2)
    po = np.asarray(batch.point_offsets, np.int64)
    lo = np.minimum.reduceat(p, po[:-1], axis=0)
    hi = np.maximum.reduceat(p, po[:-1], axis=0)
    d = hi - lo
    return np.maximum(np.hypot(d[:, 0], d[:, 1]), 1e-30)


def test_stroke_workload_all_paths():
    n = 10_000
    batch = workloads.stroke_batch(n, 8, 256, 2048.0)
    opts = workloads.stroke_options(n, 0.05)
    print("full-size stroke: GPU", flush=True)
    res = StrokeBatcher().run(DeviceBatch(batch, opts))
    gvo = res.vertex_offsets.astype(np.int64)
    gio = res.index_offsets.astype(np.int64)
    assert np.all(res.status == 0)
    so = np.zeros((n, 6), np.float32)
    for j, k in enumerate(("line_width", "miter_limit", "tolerance", "start_cap", "end_cap",
                           "line_join")):
        so[:, j] = opts[k]
    print("full-size stroke: oracle", flush=True)
    st, ovo, oio, ov, oi = O.batch_tessellate(batch, so, False)
    assert np.all(st == 0)
    assert np.array_equal(np.diff(gvo), np.diff(ovo)), "per-path vertex counts differ"
    assert np.array_equal(np.diff(gio), np.diff(oio)), "per-path index counts differ"
    diag = _path_diags(batch)
    gv = res.all_vertices()
    gi = res.all_indices()
    # compare in chunks of paths (the arrays hold ~0.5e9 indices)
    worst = 0.0
    for a in range(0, n, 500):
        b = min(n, a + 500)
        i0, i1 = int(gio[a]), int(gio[b])
        cnt = np.diff(gio[a:b + 1])
        base = np.repeat(gvo[a:b], cnt)
        local = gi[i0:i1].astype(np.int64) - base
        bad = np.nonzero(local != oi[i0:i1])[0]
        assert len(bad) == 0, f"paths {a}-{b}: {len(bad)} indices differ, first at {bad[0] + i0}"
        v0, v1 = int(gvo[a]), int(gvo[b])
        err = np.abs(gv[v0:v1].astype(np.float64) - ov[v0:v1]).max(axis=1)
        lim = np.repeat(diag[a:b], np.diff(gvo[a:b + 1])) * 1e-5
        assert np.all(err <= lim), f"paths {a}-{b}: position error {float((err / lim).max())}x limit"
        worst = max(worst, float((err / lim).max()) if len(err) else 0.0)
    print(f"full-size stroke: {n} paths, {len(gi) // 3} triangles, worst position error "
          f"{worst:.3g} of the limit", flush=True)


def _canonical(tri):
    """Triangles rotated (orientation kept) so the smallest id comes first."""
    t = tri.reshape(-1, 3)
    k = np.argmin(t, axis=1)
    r = np.arange(len(t))
    return np.stack([t[r, k], t[r, (k + 1) % 3], t[r, (k + 2) % 3]], 1)


def test_fill_workload_all_paths():
    n = 4096
    batch = workloads.fill_batch(n)
    opts = workloads.fill_options(n, 0.05)
    print("full-size fill: GPU", flush=True)
    res = FillBatcher().run(DeviceBatch(batch, opts), check=False)
    gst = res.status
    gvo = res.vertex_offsets.astype(np.int64)
    gio = res.index_offsets.astype(np.int64)
    fo = np.zeros((n, 2), np.float32)
    fo[:, 0] = opts["tolerance"]
    fo[:, 1] = opts["fill_rule"]
    print("full-size fill: oracle", flush=True)
    st, ovo, oio, ov, oi = O.batch_tessellate(batch, fo, True)
    assert np.array_equal(gst, st), "statuses differ"
    assert np.array_equal(np.diff(gvo), np.diff(ovo)), "per-path vertex counts differ"
    assert np.array_equal(np.diff(gio), np.diff(oio)), "per-path index counts differ"
    gv = res.all_vertices()
    assert np.array_equal(gv.view(np.uint32), ov.view(np.uint32)), "vertices differ"
    gi = res.all_indices().astype(np.int64)
    tcnt = np.diff(gio) // 3
    pid = np.repeat(np.arange(n), tcnt)
    gl = _canonical(gi - np.repeat(gvo[:-1], np.diff(gio)))
    ol = _canonical(oi.astype(np.int64))
    og = np.lexsort((gl[:, 2], gl[:, 1], gl[:, 0], pid))
    oo = np.lexsort((ol[:, 2], ol[:, 1], ol[:, 0], pid))
    assert np.array_equal(gl[og], ol[oo]), "triangle sets differ"
    print(f"full-size fill: {n} paths, {len(gl)} triangles, {int((gst != 0).sum())} failed "
          f"(as the oracle)", flush=True)
