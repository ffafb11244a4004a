"""Shared comparison helpers for the parity tests."""
import numpy as np

from lyon_amd.path import PathBuilder


def bbox_diag(points):
    p = np.asarray(points, np.float64).reshape(-1, 2)
    p = p[np.isfinite(p).all(axis=1)]
    if len(p) == 0:
        return 1.0
    d = p.max(axis=0) - p.min(axis=0)
    return float(max(np.hypot(d[0], d[1]), 1e-30))


def assert_same_geometry(gpu, ref, diag, rel_tol=1e-5, what=""):
    """Index topology bit-exact; positions within rel_tol * diag (f32)."""
    gv, gi = gpu
    rv, ri = ref
    assert len(gv) == len(rv), f"{what}: vertex count {len(gv)} != {len(rv)}"
    assert len(gi) == len(ri), f"{what}: index count {len(gi)} != {len(ri)}"
    if len(gi):
        bad = np.nonzero(np.asarray(gi) != np.asarray(ri))[0]
        assert len(bad) == 0, f"{what}: {len(bad)} indices differ, first at {bad[0]}: " \
                              f"{gi[bad[0]]} != {ri[bad[0]]}"
    if len(gv):
        g = np.asarray(gv, np.float64)
        r = np.asarray(rv, np.float64)
        both_nan = np.isnan(g) & np.isnan(r)
        err = np.where(both_nan, 0.0, np.abs(g - r))
        m = float(np.nanmax(err)) if err.size else 0.0
        assert m <= rel_tol * diag, f"{what}: max position error {m} > {rel_tol} * {diag}"
    return exact_fraction(gv, rv)


def exact_fraction(gv, rv):
    if len(gv) == 0:
        return 1.0
    a = np.asarray(gv, np.float32).view(np.uint32)
    b = np.asarray(rv, np.float32).view(np.uint32)
    return float(np.mean(np.all(a == b, axis=1)))


def square_paths():
    """stroke.rs:2852-2910 test_square: two open sub-paths of 4 points."""
    b = PathBuilder()
    b.begin((-1, 1)); b.line_to((1, 1)); b.line_to((1, -1)); b.line_to((-1, -1)); b.end(False)
    b.begin((-1, -1)); b.line_to((1, -1)); b.line_to((1, 1)); b.line_to((-1, 1)); b.end(False)
    return b.build()


def empty_caps_path():
    """stroke.rs:2921-2974 test_empty_caps."""
    b = PathBuilder()
    b.begin((1, 0)); b.end(True)
    b.begin((2, 0)); b.line_to((2, 0)); b.end(False)
    b.begin((3, 0)); b.line_to((3, 0)); b.end(True)
    b.begin((3, 0)); b.line_to((3, 0)); b.line_to((3, 0)); b.line_to((3, 0)); b.end(True)
    b.begin((4, 0)); b.end(False)
    return b.build()
