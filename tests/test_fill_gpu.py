"""GPU parity of the fill tessellator (plain and batched) against the oracle.

Bar, per path:
  * the same status (where lyon panics — a Rust index out of bounds or unreachable!() in
    its sweep, reachable only without intersection handling — the GPU returns
    InternalError);
  * vertices bit-exact and in the same order;
  * the same set of triangles (each rotated to start at its smallest id; lyon writes a
    span's triangles when the span ends, the GPU as they form, so only the order in the
    index buffer differs);
  * per-sample coverage identical on a 4096 x 4096 grid over the path (samples within 1e-4
    of a triangle edge excluded) and total area within 1e-6 relative.
"""
import numpy as np
import pytest

from helpers import FILL_KATS, bbox_diag, fill_corpus
from lyon_amd import BuffersBuilder, FillOptions, FillRule, FillTessellator, VertexBuffers
from lyon_amd import workloads
from lyon_amd.path import Path, PathBuilder, polygon
from lyon_amd.tessellation import Orientation, TessellationError
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu

PANIC = 6      # oracle status: the reference panics
INTERNAL = 5   # TessellationError::Internal


def gpu_fill(path, opts, vertex_offset=0):
    buf = VertexBuffers()
    try:
        FillTessellator().tessellate_path(path, opts,
                                          BuffersBuilder(buf).with_vertex_offset(vertex_offset))
    except TessellationError as e:
        return e.code, None, None
    return 0, buf.vertices, buf.indices


def ref_fill(path, opts, vertex_offset=0):
    return O.fill(path, opts.tolerance, int(opts.fill_rule), int(opts.sweep_orientation),
                  opts.handle_intersections, vertex_offset)


def canon(idx):
    t = np.asarray(idx, np.int64).reshape(-1, 3)
    if len(t) == 0:
        return t
    k = np.argmin(t, axis=1)
    r = np.arange(len(t))
    t = np.stack([t[r, k], t[r, (k + 1) % 3], t[r, (k + 2) % 3]], 1)
    return t[np.lexsort(t.T[::-1])]


def check(path, opts, what, coverage=True, vertex_offset=0):
    gst, gv, gi = gpu_fill(path, opts, vertex_offset)
    rst, rv, ri = ref_fill(path, opts, vertex_offset)
    if rst == PANIC:
        assert gst == INTERNAL, f"{what}: reference panics, gpu status {gst}"
        return
    assert gst == rst, f"{what}: status gpu {gst} != reference {rst}"
    if rst != 0:
        return
    assert len(gv) == len(rv), f"{what}: {len(gv)} vertices != {len(rv)}"
    assert np.array_equal(np.asarray(gv, np.float32).view(np.uint32),
                          np.asarray(rv, np.float32).view(np.uint32)), f"{what}: vertices differ"
    assert len(gi) == len(ri), f"{what}: {len(gi) // 3} triangles != {len(ri) // 3}"
    assert np.array_equal(canon(gi), canon(ri)), f"{what}: triangle sets differ"
    if coverage and len(ri):
        lv = np.asarray(gv, np.float32)
        c = O.coverage_compare(rv, np.asarray(ri, np.int64) - vertex_offset, lv,
                               np.asarray(gi, np.int64) - vertex_offset, n=4096, eps=1e-4)
        assert c["mismatch"] == 0, f"{what}: {c['mismatch']} samples differ"
        rel = abs(c["area_a"] - c["area_b"]) / max(c["area_a"], 1e-300)
        assert rel <= 1e-6, f"{what}: area differs by {rel}"


OPTION_SETS = {
    "evenodd": FillOptions(tolerance=0.05, fill_rule=FillRule.EvenOdd),
    "nonzero": FillOptions(tolerance=0.05, fill_rule=FillRule.NonZero),
    "horizontal": FillOptions(tolerance=0.1, fill_rule=FillRule.EvenOdd,
                              sweep_orientation=Orientation.Horizontal),
    "no_intersections": FillOptions(tolerance=0.05, fill_rule=FillRule.NonZero,
                                    handle_intersections=False),
}

CORPUS = fill_corpus()


@pytest.mark.parametrize("opt_name", list(OPTION_SETS))
def test_corpus(opt_name):
    """KAT polygons, lyon's fuzz and earcut fixtures, the plain workload path, stars."""
    opts = OPTION_SETS[opt_name]
    for k, (name, p) in enumerate(CORPUS):
        print(f"{opt_name} {name}", flush=True)
        check(p, opts, f"{opt_name}/{name}", coverage=(opt_name == "evenodd"),
              vertex_offset=7 * k)


@pytest.mark.parametrize("name,pts,step,expected", FILL_KATS, ids=[k[0] for k in FILL_KATS])
def test_kat_rotations(name, pts, step, expected):
    """fill_tests.rs triangle counts under rotation (every 0.05 rad)."""
    base = polygon(pts, True)
    for a in np.arange(0.0, 2 * np.pi, 0.05, dtype=np.float32):
        c, s = np.cos(a), np.sin(a)
        pp = base.points @ np.array([[c, s], [-s, c]], np.float32)
        path = Path(base.verbs, pp.astype(np.float32))
        st, v, i = gpu_fill(path, FillOptions(tolerance=0.05))
        assert st == 0
        rst, rv, ri = ref_fill(path, FillOptions(tolerance=0.05))
        assert len(i) // 3 == len(ri) // 3, f"{name} at {a}"
        check(path, FillOptions(tolerance=0.05, fill_rule=FillRule.NonZero), f"{name}@{a}",
              coverage=False)


def test_errors_and_empty():
    p = polygon([(0, 0), (1, 1), (0, 1)], True)
    for t in (0.0, -1.0, float("nan")):
        st, _, _ = gpu_fill(p, FillOptions(tolerance=t))
        assert st == 4  # UnsupportedParamater::ToleranceIsNaN
    empty = Path(np.zeros(0, np.uint8), np.zeros((0, 2), np.float32))
    st, v, i = gpu_fill(empty, FillOptions())
    assert st == 0 and len(v) == 0 and len(i) == 0
    b = PathBuilder(); b.begin((0, 0)); b.line_to((float("nan"), 1)); b.line_to((1, 1)); b.end(True)
    st, _, _ = gpu_fill(b.build(), FillOptions())
    assert st == 3  # UnsupportedParamater::PositionIsNaN


def test_curves_and_vertex_offset():
    """Curved paths (the plain workload and cubic loops) at several tolerances."""
    batch = workloads.fill_batch(16, seed=3)
    for i in range(len(batch)):
        for tol in (0.5, 0.05):
            check(batch.path(i), FillOptions(tolerance=tol, fill_rule=FillRule(i % 2)),
                  f"fill_batch[{i}] tol {tol}", coverage=(i < 4), vertex_offset=1000 * i)


def test_batch_matches_plain_calls():
    """The batched form equals tessellate_path per path into one shared VertexBuffers."""
    from lyon_amd.batch import DeviceBatch, FillBatcher
    n = 192
    batch = workloads.fill_batch(n, seed=0)
    opts = workloads.fill_options(n, tolerance=0.05)
    res = FillBatcher().run(DeviceBatch(batch, opts), vertex_offset=5)
    vo = res.vertex_offsets
    status = res.status
    allv = res.all_vertices()
    alli = res.all_indices()
    io = res.index_offsets
    for i in range(n):
        o = FillOptions(tolerance=0.05, fill_rule=FillRule(int(opts[i]["fill_rule"])))
        rst, rv, ri = ref_fill(batch.path(i), o, vertex_offset=5 + int(vo[i]))
        assert status[i] == rst, f"path {i}: status {status[i]} != {rst}"
        gv = allv[vo[i]:vo[i + 1]]
        gi = alli[io[i]:io[i + 1]]
        assert np.array_equal(gv.view(np.uint32), np.asarray(rv, np.float32).view(np.uint32)), \
            f"batch path {i}: vertices differ"
        assert np.array_equal(canon(gi), canon(ri)), f"batch path {i}: triangle sets differ"


def test_lds_resident_form_and_fallback():
    """The batched fill gives the same bytes whether its first round keeps each path's hot
    arrays in LDS (the default), starts them in LDS and moves them to global memory as
    they grow, or keeps them in LDS at an eighth of the estimated capacities (so that many
    paths outgrow them and are swept again in the global form); paths of every size of the
    fill workload, the 513-point stars included."""
    from lyon_amd.batch import DeviceBatch, FillBatcher, set_fill_lds_resident, synchronize
    n = 512
    batch = workloads.fill_batch(n, seed=0)
    opts = workloads.fill_options(n, tolerance=0.05)
    db = DeviceBatch(batch, opts)
    outs = []
    try:
        for mode in (1, 0, 8):
            set_fill_lds_resident(mode)
            res = FillBatcher().run(db)
            synchronize()
            outs.append((mode, res.status.copy(), res.all_vertices().copy(),
                         res.all_indices().copy()))
    finally:
        set_fill_lds_resident(True)
    m0, st0, v0, i0 = outs[0]
    assert np.all(st0 == 0)
    for m, st, v, i in outs[1:]:
        assert np.array_equal(st, st0), f"mode {m}: statuses differ"
        assert np.array_equal(v.view(np.uint32), v0.view(np.uint32)), f"mode {m}: vertices differ"
        assert np.array_equal(i, i0), f"mode {m}: indices differ"


def test_many_launches_in_one_round():
    """Paths in many size classes and placement forms (fill-workload stars and cubic loops,
    closed random walks of 2,000 to 60,000 vertices) give a round more launches than there
    used to be work counters: every path is swept (status 0) and the bytes equal those of
    the migrating form, launched the old way."""
    from lyon_amd.batch import DeviceBatch, FillBatcher, set_fill_lds_resident, synchronize
    from lyon_amd.path import PathBatch
    base = workloads.fill_batch(1024, seed=2)
    paths = [base.path(i) for i in range(len(base))]
    for k, nv in enumerate((2000, 7000, 20000, 60000)):
        paths.append(workloads.random_walk_path(nv, seed=k + 1, extent=4096.0))
    batch = PathBatch.from_paths(paths)
    opts = workloads.fill_options(len(paths), tolerance=0.05, seed=3)
    db = DeviceBatch(batch, opts)
    outs = []
    try:
        for mode in (1, 0):
            set_fill_lds_resident(mode)
            res = FillBatcher().run(db)
            synchronize()
            outs.append((res.status.copy(), res.all_vertices().copy(), res.all_indices().copy()))
    finally:
        set_fill_lds_resident(True)
    (s1, v1, i1), (s0, v0, i0) = outs
    assert np.all(s1 == 0) and np.array_equal(s1, s0)
    assert np.array_equal(v1.view(np.uint32), v0.view(np.uint32))
    assert np.array_equal(i1, i0)


def test_very_large_path():
    """fill_tests.rs:2140-2166: a path of 1,000,000 endpoints (two wiggling sides of a long
    band) fills without error; the same vertices and triangle count as the oracle and the
    same area."""
    from lyon_amd import PathBuilder
    n = 1_000_000
    d = np.float32(0.0)
    step = np.float32(0.1)
    pts = np.zeros((n, 2), np.float32)
    for k in range(n // 2):
        pts[k] = (np.cos(d, dtype=np.float32), d)
        d = np.float32(d + step)
    for k in range(n // 2, n):
        pts[k] = (np.float32(np.cos(d, dtype=np.float32) + np.float32(30.0)), d)
        d = np.float32(d - step)
    b = PathBuilder()
    b.begin((0.0, 0.0))
    for p in pts:
        b.line_to(p)
    b.end(True)
    path = b.build()
    vb = VertexBuffers()
    FillTessellator().tessellate_path(path, FillOptions(), BuffersBuilder(vb))
    st, rv, ri = O.fill(path, 0.1, 0)
    assert st == 0
    assert len(vb.vertices) == len(rv) and len(vb.indices) == len(ri)
    c = O.coverage_compare(rv, ri, vb.vertices, vb.indices, n=256)
    assert c["mismatch"] == 0 and abs(c["area_a"] - c["area_b"]) <= 1e-6 * c["area_b"]
