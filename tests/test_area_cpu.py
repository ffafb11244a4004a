"""Pins the oracle's exact fill area (oracle/refarea.cpp: the region's boundary pieces
integrated with the divergence theorem) against closed forms and against the triangle
areas of the lyon fill restatement (itself pinned by fill_tests.rs's known answers).
CPU only."""
import math

import numpy as np
import pytest

from helpers import fill_corpus
from lyon_amd.path import PathBuilder
from oracle import pyoracle as O


def _poly(pts, closed=True):
    b = PathBuilder()
    b.add_polygon([tuple(map(float, p)) for p in pts], closed)
    return b.build()


def _cross(a, b):
    return a[..., 0] * b[..., 1] - a[..., 1] * b[..., 0]


def _tri_area(v, i):
    if len(i) == 0:
        return 0.0
    v = np.asarray(v, np.float64).reshape(-1, 2)
    t = np.asarray(i, np.int64).reshape(-1, 3)
    a, b, c = v[t[:, 0]], v[t[:, 1]], v[t[:, 2]]
    return float(np.abs(_cross(b - a, c - a)).sum() / 2)


def test_squares():
    sq = _poly([(0, 0), (10, 0), (10, 10), (0, 10)])
    for rule in (0, 1):
        a, k = O.fill_area(sq, 0.1, rule)
        assert a == pytest.approx(100.0, rel=1e-12) and k == 0
    # two squares overlapping in a 5 x 5 corner, both counter-clockwise
    b = PathBuilder()
    b.add_polygon([(0.0, 0.0), (10.0, 0.0), (10.0, 10.0), (0.0, 10.0)], True)
    b.add_polygon([(5.0, 5.0), (15.0, 5.0), (15.0, 15.0), (5.0, 15.0)], True)
    p = b.build()
    assert O.fill_area(p, 0.1, 0)[0] == pytest.approx(200 - 2 * 25, rel=1e-12)
    assert O.fill_area(p, 0.1, 1)[0] == pytest.approx(200 - 25, rel=1e-12)
    # the second one reversed: a hole under NonZero too
    b = PathBuilder()
    b.add_polygon([(0.0, 0.0), (10.0, 0.0), (10.0, 10.0), (0.0, 10.0)], True)
    b.add_polygon([(5.0, 15.0), (15.0, 15.0), (15.0, 5.0), (5.0, 5.0)], True)
    p = b.build()
    assert O.fill_area(p, 0.1, 1)[0] == pytest.approx(200 - 2 * 25, rel=1e-12)


def test_pentagram():
    """{5/2} star: NonZero covers the star's outline (a decagon through the five tips and
    the five chord crossings), EvenOdd the outline minus the inner pentagon; both from the
    crossing points of the chords, computed here."""
    R = 100.0
    tips = [(R * math.cos(math.pi / 2 + 2 * math.pi * k / 5),
             R * math.sin(math.pi / 2 + 2 * math.pi * k / 5)) for k in range(5)]
    star = [tips[(2 * k) % 5] for k in range(5)]
    p = _poly(np.float32(star))
    st = np.float64(np.float32(star))

    def meet(a, b, c, d):
        r, s = b - a, d - c
        t = _cross(c - a, s) / _cross(r, s)
        return a + t * r

    chords = [(st[k], st[(k + 1) % 5]) for k in range(5)]
    inner = []
    for k in range(5):
        # chord k meets chord k+2 inside
        a, b = chords[k]
        c, d = chords[(k + 2) % 5]
        inner.append(meet(a, b, c, d))
    inner = np.array(inner)
    cen = inner.mean(0)
    inner = inner[np.argsort(np.arctan2(inner[:, 1] - cen[1], inner[:, 0] - cen[0]))]
    outer = np.vstack([st, inner])
    outer = outer[np.argsort(np.arctan2(outer[:, 1] - cen[1], outer[:, 0] - cen[0]))]

    def shoelace(q):
        return abs(np.sum(q[:, 0] * np.roll(q[:, 1], -1) - np.roll(q[:, 0], -1) * q[:, 1])) / 2

    a_nz, k = O.fill_area(p, 0.1, 1)
    a_eo, _ = O.fill_area(p, 0.1, 0)
    assert k == 5
    assert a_nz == pytest.approx(shoelace(outer), rel=1e-9)
    assert a_eo == pytest.approx(shoelace(outer) - shoelace(inner), rel=1e-9)


def test_matches_lyon_triangle_areas():
    """The fill restatement's triangles cover the same area on the corpus inputs in general
    position (the KAT polygons, the plain workload path, curves with holes, self-intersecting
    stars, cubic loops of the fill workload) under both rules: to 1e-9 without crossings;
    with crossings lyon rounds each crossing vertex to f32 and snaps vertices lying within
    tolerance / 2 of an edge onto it (fill.rs:1095, 1360-1368), each moving the area by at
    most about tolerance^2 / 2, so the areas agree within 1e-9 + k * tolerance^2 / 2 for k
    crossings (the exact area is the oracle's).  The fuzz and earcut fixtures (spikes,
    collinear overlaps) are left to the coverage checks."""
    from lyon_amd import workloads
    paths = [(n, p) for n, p in fill_corpus()
             if not n.startswith(("fuzz", "earcut"))]
    fb = workloads.fill_batch(24, seed=3)
    paths += [(f"fill_workload_{k}", fb.path(k)) for k in range(24)]
    checked = crossed = 0
    for name, path in paths:
        for rule in (0, 1):
            st, v, i = O.fill(path, 0.05, rule)
            if st != 0:
                continue
            ref = _tri_area(v, i)
            a, k = O.fill_area(path, 0.05, rule)
            bound = 1e-9 * abs(ref) + 1e-9 + k * 0.05 ** 2 / 2
            assert abs(a - ref) <= bound, (name, rule, a, ref, k)
            checked += 1
            crossed += k > 0
    assert checked >= 60 and crossed >= 20
