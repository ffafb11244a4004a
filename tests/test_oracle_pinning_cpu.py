"""Pins the oracle's flattening counts, round join / cap subdivisions, the quadratic
sharp-turn split and arc flattening against what the geometry fixes, computed here in f64
from the curves themselves (no lyon formula is restated): a count or position produced by a
dropped term, a wrong rounding or a transposed operand in the oracle fails one of these.

* Flattened segment counts (quadratic_bezier.rs:368-391, 943-974; cubic_bezier.rs:590-614,
  1369-1403): uniform subdivision of a curve whose second derivative is bounded by M keeps
  every chord within M h^2 / 8 of the curve (h = 1/n), so the fewest uniform segments that
  meet `tolerance` is n* = sqrt(M / (8 tolerance)).  lyon takes ceil(n*) while n* <= 24 and
  the integer part of n* above 24 (its count is a non-integer float turned into u32).  M
  comes from finite differences of the curve evaluated by de Casteljau in f64 (exact for
  these polynomial degrees), never from lyon's coefficients.
* Round joins and caps (stroke.rs:1968-2021, 2459-2524, 2612-2653, 2749-2753): the arc
  vertices lie on the circle of radius width / 2 around the join (cap) point, equally
  spaced; a chord of angle theta has sagitta r (1 - cos(theta / 2)), so the largest angle
  within `tolerance` is theta* = 2 acos(1 - tolerance / r); lyon uses 2^k chords with k the
  nearest integer to log2(ceil(angle / theta*)).  Checked on the oracle's stroke output.
* Sharp turns (stroke.rs:3178-3212): when the oracle splits a quadratic, the curve's
  derivative along the long axis (the control vector, or the baseline when the control
  point projects outside it) vanishes at the split t; whenever the control vector is
  longer than sqrt(30) baselines the curve is split, and never when it is shorter.
* Arcs (arc.rs:339-388): every chord of a flattened circular arc has sagitta at most
  `tolerance` and the count is ceil(|sweep| / theta*) (theta* as above).

Inputs within a relative 1e-4 of a rounding boundary of these formulas are skipped (the
f32 evaluation may legitimately fall on either side there); every test asserts that most
of its inputs were checked.
"""
import math

import numpy as np
import pytest

from lyon_amd.path import PathBuilder
from oracle import pyoracle as O

import ctypes as C


def _lib():
    L = O.lib()
    L.orc_find_sharp_turn.argtypes = [C.c_float] * 6 + [C.POINTER(C.c_float)]
    L.orc_find_sharp_turn.restype = C.c_int
    return L


def _bezier(ctrl, t):
    """de Casteljau in f64; ctrl (k, 2)."""
    p = np.asarray(ctrl, np.float64).copy()
    while len(p) > 1:
        p = (1 - t) * p[:-1] + t * p[1:]
    return p[0]


def _max_second_derivative(ctrl):
    if len(ctrl) == 3:
        # B(0) - 2 B(1/2) + B(1) = B'' / 4 for a quadratic
        d = _bezier(ctrl, 0.0) - 2 * _bezier(ctrl, 0.5) + _bezier(ctrl, 1.0)
        return float(np.hypot(*(4 * d)))
    # cubic: B'' is linear in t, so its norm peaks at an end; the one-sided 4-point
    # stencil with h = 1/3 is exact for cubics
    b = [_bezier(ctrl, k / 3) for k in range(4)]
    d0 = 9 * (2 * b[0] - 5 * b[1] + 4 * b[2] - b[3])
    d1 = 9 * (2 * b[3] - 5 * b[2] + 4 * b[1] - b[0])
    return float(max(np.hypot(*d0), np.hypot(*d1)))


def _expected_count(m, tol):
    n_star = math.sqrt(m / (8.0 * tol))
    near = lambda x: abs(x - round(x)) <= 1e-4 * max(1.0, abs(x))  # noqa: E731
    if near(n_star) or abs(n_star - 24.0) <= 1e-3:
        return None
    if n_star <= 24.0:
        return max(1, math.ceil(n_star))
    return int(math.floor(n_star))


def test_flattened_counts_match_uniform_subdivision_bound():
    rng = np.random.default_rng(21)
    n = 4000
    kinds = rng.integers(1, 3, n).astype(np.int32)
    scale = 10.0 ** rng.uniform(-1, 3.5, n)
    ctrl = (rng.random((n, 4, 2)) * scale[:, None, None]).astype(np.float32)
    checked = regimes_small = regimes_large = 0
    for tol in (0.01, 0.05, 0.1, 0.5):
        _, off = O.flatten_segments(kinds, ctrl, tol)
        counts = np.diff(off)
        for s in range(n):
            k = 3 if kinds[s] == 1 else 4
            m = _max_second_derivative(ctrl[s][:k])
            exp = _expected_count(m, tol)
            if exp is None:
                continue
            assert counts[s] == exp, (s, tol, int(counts[s]), exp, m)
            checked += 1
            regimes_small += exp <= 24
            regimes_large += exp > 24
    assert checked > 0.99 * 4 * n
    assert regimes_small > 1000 and regimes_large > 1000


def _outer_arc(vertices, center, radius):
    d = np.hypot(vertices[:, 0] - center[0], vertices[:, 1] - center[1])
    on = vertices[np.abs(d - radius) <= 1e-4 * radius + 1e-4]
    ang = np.sort(np.arctan2(on[:, 1] - center[1], on[:, 0] - center[0]))
    return ang


# euclid's fast atan2 (stroke.rs takes the arc's end angles from it) is within 2.1e-4 rad
# of atan2: the arc's own vertices are evenly spaced between its approximate end angles,
# and only the steps to the exact side vertices absorb the two angle errors.
FAST_ATAN2_ERR = 2.1e-4


def _assert_even_steps(steps, expected, what):
    assert np.all(np.abs(steps - expected) <= 2 * FAST_ATAN2_ERR), what
    if len(steps) > 2:
        inner = steps[1:-1]
        assert inner.max() - inner.min() <= 1e-5, what


def _expected_depth(angle, radius, tol):
    theta = 2.0 * math.acos(1.0 - min(tol, radius) / radius)
    m = angle / theta
    if abs(m - round(m)) <= 1e-4 * max(1.0, m):
        return None
    lg = math.log2(math.ceil(m))
    if abs(lg - math.floor(lg) - 0.5) <= 1e-3:
        return None
    return int(round(lg))


def test_round_join_subdivisions_follow_the_sagitta_bound():
    rng = np.random.default_rng(3)
    checked = 0
    depths = set()
    for _ in range(300):
        alpha = rng.uniform(0.05, 2.6) * rng.choice([-1, 1])
        width = float(rng.uniform(0.5, 60.0))
        tol = float(rng.choice([0.01, 0.05, 0.1, 0.25]))
        L = 100.0
        pb = PathBuilder()
        pb.begin((0.0, 0.0))
        pb.line_to((L, 0.0))
        pb.line_to((L + L * math.cos(alpha), L * math.sin(alpha)))
        pb.end(False)
        st, v, _ = O.stroke(pb.build(), width, 4.0, tol, 0, 0, 2)  # Round join, Butt caps
        assert st == 0
        v = np.asarray(v, np.float64).reshape(-1, 2)
        f32 = np.float32
        ang = _outer_arc(v, (L, 0.0), float(f32(width) * f32(0.5)))
        k = _expected_depth(abs(alpha), width / 2, tol)
        if k is None:
            continue
        assert len(ang) == 2 ** k + 1, (alpha, width, tol, len(ang), k)
        _assert_even_steps(np.diff(ang), abs(alpha) / 2 ** k, (alpha, width, tol))
        checked += 1
        depths.add(k)
    assert checked > 280 and len(depths) >= 4


def test_round_cap_subdivisions_follow_the_sagitta_bound():
    rng = np.random.default_rng(4)
    checked = 0
    depths = set()
    for _ in range(200):
        width = float(rng.uniform(0.5, 80.0))
        tol = float(rng.choice([0.01, 0.05, 0.1, 0.25]))
        if width / 2 < tol:
            continue
        pb = PathBuilder()
        pb.begin((0.0, 0.0))
        pb.line_to((100.0, 0.0))
        pb.end(False)
        st, v, _ = O.stroke(pb.build(), width, 4.0, tol, 2, 2, 0)  # Round caps
        assert st == 0
        v = np.asarray(v, np.float64).reshape(-1, 2)
        k = _expected_depth(math.pi / 2, width / 2, tol)
        if k is None:
            continue
        for c in ((0.0, 0.0), (100.0, 0.0)):
            ang = _outer_arc(v, c, float(np.float32(width) * np.float32(0.5)))
            # the half circle: two side vertices, the middle one and 2^k - 1 per quarter
            assert len(ang) == 2 ** (k + 1) + 1, (width, tol, len(ang), k)
            a = np.unwrap(np.sort(np.mod(ang - (0.0 if c[0] else math.pi) + math.pi / 2,
                                         2 * math.pi)))
            _assert_even_steps(np.diff(a), math.pi / 2 ** (k + 1), (width, tol))
        checked += 1
        depths.add(k)
    assert checked > 150 and len(depths) >= 4


def test_sharp_turn_split_is_the_extremum_along_the_long_axis():
    L = _lib()
    rng = np.random.default_rng(7)
    split = kept = 0
    for _ in range(20000):
        f = rng.uniform(-100, 100, 2)
        b = rng.normal(size=2) * 10 ** rng.uniform(-1, 2)
        v = rng.normal(size=2) * 10 ** rng.uniform(-1, 3)
        pts = np.float32([f, f + v, f + b])
        P = pts.astype(np.float64)
        t = C.c_float(0)
        r = L.orc_find_sharp_turn(*[float(x) for x in pts.reshape(-1)], C.byref(t))
        bb = P[2] - P[0]
        vv = P[1] - P[0]
        vb = vv @ bb
        n = np.array([-bb[1], bb[0]])
        inside = (0 <= vb <= bb @ bb) or abs(vv @ n) * 2 >= abs(vb)
        ratio = (vv @ vv) / max(bb @ bb, 1e-300)
        if inside and abs(ratio / 30 - 1) < 1e-3:
            continue  # the f32 comparison may fall either way
        if inside and ratio < 30:
            assert r == 0, (pts, ratio)
            kept += 1
            continue
        assert r == 1, (pts, ratio, inside)
        axis = vv if inside else bb
        axis = axis / np.linalg.norm(axis)
        tt = float(t.value)
        assert 0.0 < tt < 1.0
        d = 2 * (1 - tt) * (P[1] - P[0]) + 2 * tt * (P[2] - P[1])  # B'(t)
        dmax = max(np.linalg.norm(P[1] - P[0]), np.linalg.norm(P[2] - P[1])) * 2
        # euclid's fast atan2 rotates the axis by up to FAST_ATAN2_ERR
        assert abs(d @ axis) <= 1.5 * FAST_ATAN2_ERR * dmax, (pts, tt, d @ axis, dmax)
        split += 1
    assert split > 5000 and kept > 5000


def test_arc_flattening_sagitta_and_count():
    rng = np.random.default_rng(9)
    n = 3000
    r = 10 ** rng.uniform(-0.5, 3, n)
    tol = rng.choice([0.01, 0.05, 0.1, 0.3], n)
    sweep = rng.uniform(-2 * math.pi, 2 * math.pi, n)
    arcs = np.stack([rng.uniform(-500, 500, n), rng.uniform(-500, 500, n), r, r,
                     rng.uniform(-math.pi, math.pi, n), sweep, rng.uniform(-1, 1, n)], 1)
    arcs = arcs.astype(np.float32)
    keep = arcs[:, 2] > 2.5 * tol
    arcs, tol = arcs[keep], tol[keep]
    checked = 0
    for tv in np.unique(tol):
        sel = arcs[tol == tv]
        pts, off = O.flatten_arcs(sel, float(tv))
        for i, a in enumerate(sel):
            A = a.astype(np.float64)
            cx, cy, rad, s0, sw, xr = A[0], A[1], A[2], A[4], A[5], A[6]
            theta = 2 * math.acos(1 - float(np.float32(tv)) / rad)
            m = abs(sw) / theta
            seg = pts[off[i]:off[i + 1]].astype(np.float64)
            start = np.array([cx + rad * math.cos(s0 + xr), cy + rad * math.sin(s0 + xr)])
            poly = np.vstack([start, seg])
            d = np.hypot(poly[:, 0] - cx, poly[:, 1] - cy)
            assert np.all(np.abs(d - rad) <= 1e-5 * max(rad, 600)), (i, a)
            # sagitta of every chord: r - distance from the centre to the chord's midpoint
            mid = (poly[:-1] + poly[1:]) / 2
            sag = rad - np.hypot(mid[:, 0] - cx, mid[:, 1] - cy)
            # lyon evaluates acos((r - tol) / r) in f32: the quotient's rounding (half an
            # ulp of 1) moves theta* by a relative 3e-8 r / tol, so counts within that of
            # an integer are not decided by the geometry
            cond = 3e-8 * rad / tv + 1e-6
            assert sag.max() <= tv * (1 + 1e-3 + 2 * cond) + 2e-5 * max(rad, 600), \
                (i, a, sag.max())
            if m > 1e4 or abs(m - round(m)) <= max(1e-4, 4 * cond) * max(1.0, m):
                continue
            assert off[i + 1] - off[i] == math.ceil(m), (i, a, off[i + 1] - off[i], m)
            checked += 1
    assert checked > 0.85 * len(arcs)  # the rest are ill-conditioned in f32 (large r / tol)
