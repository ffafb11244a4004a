"""Pins the oracle (the CPU restatement of lyon) against the reference's own known
answers, fixtures and closed forms.  Runs on the CPU (no GPU, no liblyg kernels).

Sources of each expectation:
* stroke KATs: stroke.rs:2852-2910 (test_square), :2921-2974 (test_empty_caps),
  :3126-3154 (test_line_width), :3380-3416 (correct_miter_clip_length);
* fill KATs: fill_tests.rs:153-624 triangle counts, with the reference's rotation sweeps
  (test_path_with_rotations, :134-150);
* fixtures: fuzz_tests.rs and earcut_tests.rs polygons (tests/golden/*.json) must
  tessellate without error under both fill rules;
* flattening: quadratic_bezier.rs:1381-1461 (degenerate -> 1 segment, t monotonic),
  cubic_bezier.rs:1614-1646 (last point exact), arc.rs:940-1017 (arc -> quadratic
  counts), plus Wang's bound checked against an f64 evaluation of the true curve.
"""
import ctypes as C
import json
import math
import os

import numpy as np
import pytest

from helpers import FILL_KATS, empty_caps_path, fixture_path, square_paths
from lyon_amd.path import Path, PathBuilder, polygon
from oracle import pyoracle as O

HERE = os.path.dirname(os.path.abspath(__file__))


def rotate_inplace(pts, angle):
    L = O.lib()
    L.orc_rotate_points.argtypes = [O.F32P, C.c_uint32, C.c_float]
    L.orc_rotate_points(pts.ctypes.data_as(O.F32P), len(pts), angle)


# ---- stroke ---------------------------------------------------------------------------
def test_stroke_square_counts():
    p = square_paths()
    assert len(O.stroke(p, line_join=0, start_cap=0, end_cap=0)[2]) // 3 == 12
    assert len(O.stroke(p, line_join=3, start_cap=1, end_cap=1)[2]) // 3 == 16


def test_stroke_empty_caps():
    p = empty_caps_path()
    assert len(O.stroke(p, start_cap=0, end_cap=0)[2]) == 0
    assert len(O.stroke(p, start_cap=1, end_cap=1)[2]) // 3 == 8


def test_stroke_line_width_vertices():
    b = PathBuilder()
    b.begin((0, 1)); b.line_to((2, 1)); b.end(False)
    _, v, _ = O.stroke(b.build(), line_width=2.0)
    got = {tuple(x) for x in v.tolist()}
    assert got == {(0.0, 0.0), (0.0, 2.0), (2.0, 0.0), (2.0, 2.0)}


def test_stroke_miter_clip_length():
    b = PathBuilder()
    b.begin((-70, -200)); b.line_to((0, 300)); b.line_to((70, -200)); b.end(False)
    _, v, _ = O.stroke(b.build(), line_width=30.0, line_join=1, miter_limit=3.0)
    assert float(v[:, 1].max()) == 300.0 + 30.0 * 3.0 * 0.5


def test_stroke_triangle_winding():
    """stroke.rs:3214-3257: every triangle is clockwise (y down) up to -0.035 slack, and
    no triangle repeats a vertex (stroke.rs:2787-2849 TestBuilder)."""
    from lyon_amd import workloads
    b = workloads.stroke_batch(20, 4, 30, 500.0, seed=3)
    for i in range(len(b)):
        for join in (0, 2, 3):
            st, v, idx = O.stroke(b.path(i), line_width=3.0, line_join=join, tolerance=0.05,
                                  start_cap=2, end_cap=1)
            assert st == 0
            t = idx.reshape(-1, 3)
            assert np.all((t[:, 0] != t[:, 1]) & (t[:, 0] != t[:, 2]) & (t[:, 1] != t[:, 2]))
            pa, pb, pc = v[t[:, 0]].astype(np.float64), v[t[:, 1]], v[t[:, 2]]
            cr = (pa - pb)[:, 0] * (pc - pb)[:, 1] - (pa - pb)[:, 1] * (pc - pb)[:, 0]
            assert np.mean(cr >= -0.035) > 0.99


def test_stroke_vertex_offset_and_overflow():
    """BuffersBuilder ids are len + vertex_offset; ids past u32::MAX - 1 abort
    (geometry_builder.rs:338-350, 609-630)."""
    p = square_paths()
    st, v, i = O.stroke(p, vertex_offset=1000)
    assert st == 0 and i.min() >= 1000 and i.max() < 1000 + len(v)
    st, v, i = O.stroke(p, vertex_offset=0xFFFFFFFE - 3)
    assert st == 1 and len(v) == 0 and len(i) == 0


# ---- fill -----------------------------------------------------------------------------
def fill_count(path, rule=0, tol=0.05):
    st, v, i = O.fill(path, tolerance=tol, fill_rule=rule)
    assert st == 0, f"fill failed with status {st}"
    return len(i) // 3


def poly(*pts):
    return polygon(pts, True)


def rotations(path, step, expected):
    angle = np.float32(0.0)
    two_pi = np.float32(math.pi * 2.0)
    n = 0
    while angle < two_pi:
        pts = path.points.copy()
        rotate_inplace(pts, float(angle))
        rp = Path(path.verbs, pts)
        if expected is not None:
            assert fill_count(rp, 0) == expected, f"angle {angle}"
        else:
            fill_count(rp, 0)
        fill_count(rp, 1)
        angle = np.float32(angle + np.float32(step))
        n += 1
    return n


@pytest.mark.parametrize("name,pts,step,expected", FILL_KATS, ids=[k[0] for k in FILL_KATS])
def test_fill_kat_rotations(name, pts, step, expected):
    """fill_tests.rs: exact EvenOdd triangle counts at every rotation (the reference
    steps by 0.001-0.01 rad; a 0.01 step keeps the CPU suite short)."""
    assert rotations(poly(*pts), step, expected) > 600


def test_fill_kat_fixed():
    assert fill_count(poly((0, 0), (-1, 1), (-3, 2), (-1, 3), (-4, 5), (0, 6))) == 4
    assert fill_count(poly((0, 0), (2, 1), (0, 2), (2, 3))) == 2
    assert fill_count(poly((0, 0), (2, 3), (2, 1), (0, 2))) == 2
    assert fill_count(poly((1, 0), (2, 1), (3, 0), (4, 2), (5, 0), (6, 3), (7, 0), (5, 8))) == 6
    assert fill_count(poly((1, 0), (2, 1), (3, 0), (4, 2), (5, 0), (6, 3), (7, 0), (7, 5),
                           (0, 4))) == 7
    assert fill_count(poly((1, 0), (2, 1), (3, 0), (4, 2), (5, 0), (6, 3), (7, 0), (7, 5),
                           (-1, 5), (-1, 0), (0, 4))) == 9
    assert fill_count(poly((1, 0), (2, 1), (3, 0), (4, 2), (5, 0), (6, 3), (7, 0), (7, 5),
                           (4, 4), (1, 5))) == 8
    assert fill_count(Path(np.zeros(0, np.uint8), np.zeros((0, 2), np.float32))) == 0


def test_fill_kat_hole_and_multi():
    b = PathBuilder()
    b.begin((-11, 5)); b.line_to((0, -5)); b.line_to((10, 5)); b.end(True)
    b.begin((-5, 2)); b.line_to((0, -2)); b.line_to((4, 2)); b.end(True)
    rotations(b.build(), 0.01, 6)
    b = PathBuilder()
    b.begin((20, 20)); b.line_to((60, 20)); b.line_to((60, 60)); b.line_to((20, 60)); b.end(True)
    b.begin((40, 10)); b.line_to((70, 40)); b.line_to((40, 70)); b.line_to((10, 40)); b.end(True)
    rotations(b.build(), 0.011, 8)


def test_fill_tolerance_errors():
    p = poly((0, 0), (1, 1), (0, 1))
    for t in (0.0, -1.0, float("nan")):
        st, v, i = O.fill(p, tolerance=t)
        assert st == 4 and len(v) == 0


def test_fuzz_fixtures_no_error():
    cases = json.load(open(os.path.join(HERE, "golden", "fuzz_polygons.json")))
    assert len(cases) == 26
    for name, subs in cases.items():
        for rule in (0, 1):
            st, v, i = O.fill(fixture_path(subs), tolerance=0.05, fill_rule=rule)
            assert st == 0, f"{name} rule {rule}: status {st}"


def test_earcut_fixtures_no_error():
    cases = json.load(open(os.path.join(HERE, "golden", "earcut_polygons.json")))
    assert len(cases) == 30
    for name, rings in cases.items():
        for rule in (0, 1):
            st, v, i = O.fill(fixture_path(rings, earcut=True), tolerance=0.05, fill_rule=rule)
            assert st == 0, f"{name} rule {rule}: status {st}"


# ---- flattening -----------------------------------------------------------------------
def test_flatten_degenerate_quadratics():
    """quadratic_bezier.rs:1381-1419: empty and straight quadratics give 1 segment."""
    ctrl = np.array([[[0, 0], [0, 0], [0, 0], [0, 0]],
                     [[0, 0], [1, 0], [2, 0], [0, 0]]], np.float32)
    pts, off = O.flatten_segments(np.array([1, 1], np.int32), ctrl, 0.1)
    assert list(np.diff(off)) == [1, 1]


def test_flatten_last_point_exact_and_error_bound():
    """Last point == `to` exactly (cubic_bezier.rs:1614-1646), and every chord stays
    within the uniform-parameter chord bound M * h^2 / 8 of the true curve (f64), where
    M = max |B''| (Wang) and h the chord's parameter length: interior chords have
    h = 1/n, the last one spans 1 - (count - 1)/n because the count is n truncated
    (quadratic_bezier.rs:375-390).  So the flattening error may exceed `tolerance` by up
    to (n - count + 1)^2 on the last chord — lyon's behaviour, reported, not a bug."""
    rng = np.random.default_rng(5)
    n = 300
    kinds = rng.integers(1, 3, n).astype(np.int32)
    ctrl = (rng.random((n, 4, 2)) * 1000).astype(np.float32)
    worst_ratio = 0.0
    for tol in (0.05, 0.25):
        pts, off = O.flatten_segments(kinds, ctrl, tol)
        for s in range(n):
            seg = pts[off[s]:off[s + 1]].astype(np.float64)
            c = ctrl[s].astype(np.float64)
            assert np.array_equal(pts[off[s + 1] - 1], ctrl[s][2 if kinds[s] == 1 else 3])
            if kinds[s] == 1:
                lv = (c[0] - 2 * c[1] + c[2]) * 2
                M = float(np.sqrt(lv @ lv))
            else:
                v1 = (c[0] - 2 * c[1] + c[2]) * 6
                v2 = (c[1] - 2 * c[2] + c[3]) * 6
                M = float(np.sqrt(max(v1 @ v1, v2 @ v2)))
            wang = (M * M / (8 * tol) ** 2) ** 0.25
            cnt = off[s + 1] - off[s]
            h_last = 1.0 if cnt == 1 else max(1.0 / wang, 1.0 - (cnt - 1) / wang)
            bound = M * h_last ** 2 / 8 * (1 + 1e-4) + 1e-3
            poly_pts = np.vstack([c[0], seg])
            ts = np.linspace(0, 1, 4001)
            if kinds[s] == 1:
                curve = ((1 - ts) ** 2)[:, None] * c[0] + (2 * (1 - ts) * ts)[:, None] * c[1] + \
                        (ts ** 2)[:, None] * c[2]
            else:
                curve = ((1 - ts) ** 3)[:, None] * c[0] + (3 * (1 - ts) ** 2 * ts)[:, None] * c[1] \
                        + (3 * (1 - ts) * ts ** 2)[:, None] * c[2] + (ts ** 3)[:, None] * c[3]
            a = poly_pts[:-1][None]
            b = poly_pts[1:][None]
            p = curve[:, None]
            ab = b - a
            t = np.clip(np.sum((p - a) * ab, -1) / np.maximum(np.sum(ab * ab, -1), 1e-30), 0, 1)
            d = np.linalg.norm(p - (a + t[..., None] * ab), axis=-1).min(axis=1)
            assert d.max() <= bound, (s, d.max(), bound)
            worst_ratio = max(worst_ratio, d.max() / tol)
    assert worst_ratio < 4.0


def test_arc_to_quadratic_counts():
    """arc.rs:940-1017: sweeps of pi/2 and pi produce 2 and 4 quadratics... (n = ceil(sweep /
    (pi/4)))."""
    arcs = np.array([[0, 0, 1, 1, 0, math.pi, 0], [0, 0, 1, 1, 0, math.pi / 2, 0],
                     [0, 0, 1, 1, 0, 0.5, 0], [0, 0, 1, 1, 0, -math.pi, 0]], np.float32)
    q, off = O.arc_to_quads(arcs)
    assert list(np.diff(off)) == [4, 2, 1, 4]


def test_tchain_matches_serial_accumulation():
    """The closed-form t chain used by the kernels equals the serial f32 sum (exhaustive
    over j for 2000 segment counts), also when resumed from an intermediate state."""
    import subprocess
    import sys
    src = os.path.join(os.path.dirname(HERE), "lyon_amd", "csrc", "lyg_math.h")
    code = r'''
#include "%s"
#include <cstdio>
#include <random>
int main() {
  using namespace lyg; std::mt19937 rng(11); std::uniform_real_distribution<float> u(1.0f, 3000.0f);
  long bad = 0, tot = 0;
  for (int k = 0; k < 2000; ++k) {
    float n = k < 300 ? (float)(k + 1) : u(rng); float s = 1.0f / n; float t = 0;
    for (uint32_t j = 1; j <= seg_count(n) + 2; ++j) {
      t = t + s; tot++; if (f2u(tchain(s, j)) != f2u(t)) bad++;
      if (tchain_units_of(s, t) != tchain_units(tchain_mant(s), j)) bad++;  // units of a serial sum
    }
    // continuing a chain from any earlier state (the kernels resume where the previous
    // window of steps stopped) gives the same units
    const uint32_t m = tchain_mant(s), nj = seg_count(n) + 2;
    for (uint32_t k0 = 0; k0 <= nj; k0 += 1 + k0 / 3) {
      const uint64_t T0 = tchain_units(m, k0);
      for (uint32_t j = k0; j <= nj; j += 1 + (j - k0) / 5) {
        tot++; if (tchain_units_from(m, j, k0, T0) != tchain_units(m, j)) bad++;
      }
    }
  }
  printf("%%ld %%ld\n", tot, bad); return bad != 0;
}''' % src
    tmp = os.path.join(os.environ.get("TMPDIR", "/tmp"), "lyg_tchain_test")
    with open(tmp + ".cpp", "w") as f:
        f.write(code)
    subprocess.run(["g++", "-O2", "-ffp-contract=off", "-o", tmp, tmp + ".cpp"], check=True)
    r = subprocess.run([tmp], capture_output=True, text=True)
    tot, bad = map(int, r.stdout.split())
    assert tot > 1_000_000 and bad == 0


@pytest.mark.parametrize("radius,tol,n_vertices", [(100.0, 0.1, 64), (3.0, 0.01, 32),
                                                   (1e-3, 0.1, 4)])
def test_fill_circle_restatement(radius, tol, n_vertices):
    """basic_shapes.rs:209-231 fills a circle of radius 100 at the default tolerance.  The
    counts are worked by hand: r=100, t=0.1 gives a step 2 sqrt(2 t r - t^2) = 8.943, a
    quarter arc of 157.08 = 18 steps, floor(log2 18) = 4 halvings per quarter, 4 * 2^4
    vertices.  The topology follows from the geometry: after the two base triangles, the
    k-th triangle adds vertex k+2 on the bisector of its other two corners."""
    v, i = O.fill_circle((1.0, 2.0), radius, tol)
    assert len(v) == n_vertices and len(i) == 3 * (n_vertices - 2)
    d = np.hypot(v[:, 0] - 1.0, v[:, 1] - 2.0)
    assert np.all(np.abs(d - radius) <= 1e-5 * radius + 1e-7)
    tri = i.reshape(-1, 3).astype(np.int64)
    assert tri[:2].tolist() == [[0, 3, 1], [1, 3, 2]]
    assert np.array_equal(tri[2:, 1], np.arange(4, len(v)))
    ang = np.arctan2(v[:, 1].astype(np.float64) - 2.0, v[:, 0].astype(np.float64) - 1.0)
    a, b, m = ang[tri[2:, 2]], ang[tri[2:, 0]], ang[tri[2:, 1]]
    bis = a + 0.5 * np.angle(np.exp(1j * (b - a)))
    assert np.all(np.abs(np.angle(np.exp(1j * (m - bis)))) <= 1e-4)
    # each added triangle spans at most a quarter turn of the circle
    assert np.all(np.abs(np.angle(np.exp(1j * (b - a)))) <= math.pi / 2 + 1e-6)


def _hit_test_kat_path():
    # hit_test.rs:119-133
    b = PathBuilder()
    for sub in ([(0.0, 0.0), (1.0, 0.0), (1.0, 1.0), (0.0, 1.0)],
                [(0.25, 0.25), (0.75, 0.25), (0.75, 0.75), (0.20, 0.75)]):
        b.begin(sub[0])
        for p in sub[1:]:
            b.line_to(p)
        b.end(True)
    return b.build()


def test_hit_test_reference_kats():
    """hit_test.rs:119-197: the reference's own known answers."""
    p = _hit_test_kat_path()
    for q in [(-1.0, 0.5), (2.0, 0.5), (2.0, 0.0), (0.5, -1.0), (0.5, 2.0), (0.5, 0.5)]:
        assert not O.hit_test(p, q, 0, 0.1), q
    assert O.hit_test(p, (0.5, 0.5), 1, 0.1)
    assert O.hit_test(p, (0.2, 0.5), 0, 0.1)
    assert O.hit_test(p, (0.8, 0.5), 0, 0.1)


def test_signed_area_reference_kats():
    """area.rs:80-130: +1 and -1 for the two windings of the unit square, 5 for the cross."""
    b = PathBuilder()
    for sub in ([(0.0, 0.0), (1.0, 0.0), (1.0, 1.0), (0.0, 1.0)],
                [(0.0, 0.0), (0.0, 1.0), (1.0, 1.0), (1.0, 0.0)]):
        b.begin(sub[0])
        for q in sub[1:]:
            b.line_to(q)
        b.close()
        one = b.build()
        b = PathBuilder()
        assert O.signed_area(one, 0.01) in (1.0, -1.0)
    sq = PathBuilder()
    sq.add_polygon([(0.0, 0.0), (1.0, 0.0), (1.0, 1.0), (0.0, 1.0)], True)
    assert O.signed_area(sq.build(), 0.01) == 1.0
    sq = PathBuilder()
    sq.add_polygon([(0.0, 0.0), (0.0, 1.0), (1.0, 1.0), (1.0, 0.0)], True)
    assert O.signed_area(sq.build(), 0.01) == -1.0
    cross = [(0.0, 1.0), (1.0, 1.0), (1.0, 0.0), (2.0, 0.0), (2.0, 1.0), (3.0, 1.0), (3.0, 2.0),
             (2.0, 2.0), (2.0, 3.0), (1.0, 3.0), (1.0, 2.0), (0.0, 2.0)]
    b = PathBuilder()
    b.add_polygon(cross, True)
    assert O.signed_area(b.build(), 0.01) == 5.0


def _validator_paths():
    from helpers import star_paths
    from lyon_amd.workloads import stroke_batch
    paths = [p for _, p in star_paths(4, seed=3, max_points=60)]
    sb = stroke_batch(6, 4, 40, 512.0, seed=9)
    for i in range(len(sb)):
        paths.append(sb.path(i))
    b = PathBuilder()
    b.add_circle((100.0, 50.0), 40.0, True)
    # an ellipse as add_ellipse builds it (builder.rs:606-636), quadratics from the oracle
    sweep = -2.0 * float(np.float32(np.pi))
    q, _ = O.arc_to_quads(np.array([[300.0, 60.0, 90.0, 30.0, 0.0, sweep, 0.3]], np.float32))
    b.begin(q[0, 0])
    for k in range(len(q)):
        b.quadratic_bezier_to(q[k, 1], q[k, 2])
    b.end(True)
    paths.append(b.build())
    return paths


def test_signed_area_matches_shoelace():
    """The f32 serial sum against a float64 shoelace over the same flattened points: equal
    to f32 rounding (a dropped closing term or a wrong sign would be off by far more).
    A circle's area is pi r^2 within the flattening error (perimeter * tolerance)."""
    from helpers import area64, flattened_subpaths
    for p in _validator_paths():
        a = O.signed_area(p, 0.05)
        ref = area64(flattened_subpaths(p, 0.05))
        scale = float(np.abs(np.asarray(p.points, np.float64)).max()) ** 2
        assert abs(a - ref) <= 1e-5 * scale + 1e-6 * abs(ref), (a, ref)
    b = PathBuilder()
    b.add_circle((10.0, 20.0), 100.0, True)
    a = O.signed_area(b.build(), 0.01)
    assert abs(a - math.pi * 1e4) <= 2 * math.pi * 100 * 0.01 + 1e-3 * 1e4 * 1e-2


def test_winding_matches_crossing_count():
    """Winding numbers against a float64 crossing count over the flattened edges, at
    random points away (> 1e-3) from every edge."""
    from helpers import flattened_subpaths, winding64
    rng = np.random.default_rng(4)
    for p in _validator_paths():
        pts = np.asarray(p.points, np.float64)
        lo, hi = pts.min(axis=0) - 5, pts.max(axis=0) + 5
        xy = (lo + rng.random((400, 2)) * (hi - lo)).astype(np.float32)
        w = O.winding_numbers(p, xy, 0.05)
        ref, dmin = winding64(flattened_subpaths(p, 0.05), xy)
        far = dmin > 1e-3
        assert far.sum() > 300
        assert np.array_equal(w[far], ref[far])
        assert np.any(w != 0)


def test_parse_f32_matches_exact_rounding():
    """The oracle's number conversion (grammar check + strtof) against exact rational
    rounding, bit for bit, including ties, overflow and the subnormal range."""
    from helpers import f32_exact, number_corpus
    for s in number_corpus(2000, seed=1):
        got = O.parse_f32(s)
        assert got is not None, s
        want = f32_exact(s)
        assert np.float32(got).view(np.uint32) == want.view(np.uint32), (s, got, want)
    for bad in ["", "-", ".", "-.", "e5", "1e", "1e-", "1ee2", "--1", "+1", "1.2.3"]:
        assert O.parse_f32(bad) is None, bad


def test_svg_parser_reference_kats():
    """parser.rs:498-700: the reference's own parser tests."""
    assert O.parse_svg("")[2][0] == 0 and len(O.parse_svg(" ")[0]) == 0
    v, p, e = O.parse_svg("M 0 0 L 1 0 L 1 1 L 0 1 Z")
    assert v.tolist() == [3, 0, 0, 0, 4] and e[0] == 0
    assert p.tolist() == [[0, 0], [1, 0], [1, 1], [0, 1], [0, 0]]
    assert O.parse_svg("M 0 0 1.0 L 1 0 2.0 L 1 1 3.0 L 0 1 4.0 Z", 1)[2][0] == 0
    v, p, e = O.parse_svg("0 0 0 1 1 1.0 2 2 2.0 3 3 3", 1)
    assert e[0] == 0 and v.tolist() == [3, 0, 0, 0, 5]
    assert O.parse_svg("x 0 0 0", 1)[2].tolist() == [3, 0, 0, ord("x")]
    assert O.parse_svg("\n M 0 \n0 1 x 1 1 1", 1)[2].tolist() == [3, 2, 4, ord("x")]
    v, p, e = O.parse_svg("M 0.6.5")
    assert v.tolist() == [3, 5] and p.tolist() == [[np.float32(0.6), 0.5]] and e[0] == 0
    assert O.parse_svg("M 1e-2 -1E3")[2][0] == 0
    for bad in ["M 0 --1", "M 0 1ee2", "M 0 1e--1", "M 0 *2", "M 0 e", "M 0 1e", "M 0 +1"]:
        assert O.parse_svg(bad)[2][0] == 1, bad
    for s in ["M 0 0 | xxxxxx", "M 0 0| xxxxxx", "| xxxxxx", "    | xxxxxx"]:
        assert O.parse_svg(s, stop_at="|")[2][0] == 0, s
    v1, p1, e1 = O.parse_svg("M 0 0 Z L 1 1 2 2 L 3 3 Z M 4 4", stop_at="|")
    v2, p2, e2 = O.parse_svg("M 0 0 Z", stop_at="|")
    assert e1[0] == 4 and e2[0] == 0
    assert v1.tolist() == v2.tolist() and p1.tolist() == p2.tolist()


def test_svg_arcs_follow_the_geometry():
    """An SVG arc's quadratics start at the current point and end near the arc's end point
    (its angles come from euclid's fast_atan2, an approximation good to ~2e-4 rad, so the
    end point moves by up to radius * 2e-4), for every flag combination; a zero radius or a
    zero-length arc is a line (SvgArc::is_straight_line)."""
    for large in (0, 1):
        for sweep in (0, 1):
            v, p, e = O.parse_svg(f"M 10 10 A 30 20 25 {large} {sweep} 50 35")
            assert e[0] == 0 and v[0] == 3 and v[-1] == 5 and set(v[1:-1]) == {1}
            assert np.allclose(p[-1], [50, 35], atol=30 * 3e-4)
    assert O.parse_svg("M 1 1 A 0 5 0 0 1 9 9")[0].tolist() == [3, 0, 5]
    assert O.parse_svg("M 1 1 A 5 5 0 0 1 1 1")[0].tolist() == [3, 0, 5]


def test_svg_output_fits_the_staging_room():
    """The device parser writes string i into a room of len + 2 verbs and 2 len + 2 points:
    every command consumes a character and makes at most one verb and two points per
    character.  Checked on the densest inputs (arcs packed as tightly as the syntax allows,
    glued signs, '.5.5' runs, repeated moves and closes) and on random strings."""
    import sys
    sys.path.insert(0, os.path.dirname(__file__))
    from helpers import random_svg
    dense = ["M0 0A1 1 0 111 0" + " 1 1 0 110 0 1 1 0 111 0" * 50,  # 7 quadratics per arc
             "M0 0A1 1 0 011 1" + " 1 1 0 011 1" * 50,
             "M0 0a1-1-0 01-1-1" * 20,
             "M0 0T1 1" + "-1-1" * 100,
             "M.5.5q.5.5.5.5" + ".5.5.5.5" * 60,
             "M1 1Z" * 40, "M1 1M1 1M1 1", "m1 1z1 1z1 1z",
             "M0 0c1 1 1 1 1 1" + " 1-1-1-1-1-1" * 40,
             "M0 0A1 1 0 111 0" * 30, "M0 0a9 9 0 1 1 1e-3 0" * 20]
    rng = np.random.default_rng(9)
    dense += [random_svg(rng, int(rng.integers(1, 60)), arcs=True, corrupt=0.2)
              for _ in range(500)]
    worst = 0.0
    for s in dense:
        v, p, e = O.parse_svg(s)
        n = len(s.encode())
        assert len(v) <= n + 2 and len(p) <= 2 * n + 2, (s[:40], len(v), len(p), n)
        worst = max(worst, len(p) / max(n, 1))
    assert worst > 1.0  # the arcs really are denser than one point per byte
