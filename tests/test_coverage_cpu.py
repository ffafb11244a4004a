"""Pins the fill checker and the fill oracle on the CPU.

* The coverage checker (oracle/coverage.cpp) must report zero difference between two
  triangulations of the same region, and a difference for a planted mistake (a dropped
  or moved triangle).
* The fill oracle's triangles must cover exactly the samples that the textbook winding
  number of the path's flattened edges puts inside (even-odd or non-zero), for every
  corpus path.  The tolerance is scaled to the path (1e-5 of its bbox diagonal) so lyon's
  vertex snapping (within tolerance / 2) stays inside the excluded band around edges
  (1e-4 of the diagonal).
"""
import numpy as np
import pytest

from helpers import bbox_diag, fill_corpus
from lyon_amd.path import PathBuilder, polygon
from oracle import pyoracle as O

SQUARE = np.array([[0, 0], [10, 0], [10, 10], [0, 10]], np.float32)


def test_checker_same_region_two_triangulations():
    a = np.array([0, 1, 2, 0, 2, 3], np.uint32)
    b = np.array([0, 1, 3, 1, 2, 3], np.uint32)
    r = O.coverage_compare(SQUARE, a, SQUARE, b, box=(-1, -1, 11, 11), n=512)
    assert r["mismatch"] == 0
    assert r["covered_a"] == r["covered_b"] > 0
    assert abs(r["area_a"] - 100.0) < 1e-12 and abs(r["area_b"] - 100.0) < 1e-12
    assert r["overlap_a"] == 0 and r["overlap_b"] == 0


def test_checker_planted_mistakes():
    a = np.array([0, 1, 2, 0, 2, 3], np.uint32)
    dropped = np.array([0, 1, 2], np.uint32)
    r = O.coverage_compare(SQUARE, a, SQUARE, dropped, box=(-1, -1, 11, 11), n=512)
    assert r["mismatch"] > 0.4 * r["covered_a"]
    assert abs(r["area_b"] - 50.0) < 1e-12
    moved = SQUARE.copy()
    moved[2] += 0.05
    r = O.coverage_compare(SQUARE, a, moved, a, box=(-1, -1, 11, 11), n=1024)
    assert r["mismatch"] > 0
    overlap = np.array([0, 1, 2, 0, 2, 3, 0, 1, 2], np.uint32)
    r = O.coverage_compare(SQUARE, a, SQUARE, overlap, box=(-1, -1, 11, 11), n=512)
    assert r["mismatch"] == 0 and r["overlap_b"] > 0 and r["area_b"] > r["area_a"]


def test_winding_checker_rules():
    """Overlapping squares: even-odd leaves the overlap empty, non-zero fills it."""
    b = PathBuilder()
    b.begin((0, 0)); b.line_to((10, 0)); b.line_to((10, 10)); b.line_to((0, 10)); b.end(True)
    b.begin((5, 5)); b.line_to((15, 5)); b.line_to((15, 15)); b.line_to((5, 15)); b.end(True)
    p = b.build()
    for rule, area in ((0, 150.0), (1, 175.0)):
        st, v, i = O.fill(p, 0.01, rule)
        assert st == 0
        r = O.winding_compare(p, 0.01, rule, v, i, n=1024)
        assert r["mismatch"] == 0
        assert abs(r["area_a"] - area) < 1e-9
        # the other rule's triangles disagree with this rule's winding
        st2, v2, i2 = O.fill(p, 0.01, 1 - rule)
        r2 = O.winding_compare(p, 0.01, rule, v2, i2, n=1024)
        assert r2["mismatch"] > 0


@pytest.mark.parametrize("rule", [0, 1])
def test_oracle_matches_winding_number(rule):
    worst = 0
    for name, p in fill_corpus():
        d = bbox_diag(p.points)
        tol = 1e-5 * d
        st, v, i = O.fill(p, tol, rule)
        assert st == 0, f"{name}: status {st}"
        r = O.winding_compare(p, tol, rule, v, i, n=1024, eps=1e-4 * d)
        assert r["mismatch"] == 0, f"{name}: {r['mismatch']} samples differ from the winding rule"
        worst = max(worst, r["overlap_a"])
    assert worst == 0


def test_oracle_planted_winding_mistake():
    """A plausible oracle mistake (a dropped triangle) is caught by the winding check."""
    p = polygon([(0, 0), (4, 0), (5, 3), (2, 6), (-1, 3)], True)
    st, v, i = O.fill(p, 0.01, 0)
    r = O.winding_compare(p, 0.01, 0, v, i[3:], n=512)
    assert r["mismatch"] > 0
