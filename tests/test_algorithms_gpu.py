"""lyon_algorithms' validators on the GPU (hit_test.rs, area.rs) against the oracle:
winding numbers bit-exact (integers), areas bit-exact (lyon's serial f32 sum)."""
import numpy as np
import pytest

from helpers import star_paths
from lyon_amd import FillRule, PathBatch, PathBuilder, TessellationError
from lyon_amd import algorithms as A
from lyon_amd.batch import DeviceBatch, DeviceBuffer
from lyon_amd.workloads import fill_batch, stroke_batch
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu


def _kat_path():
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
    """hit_test.rs:135-197."""
    p = _kat_path()
    for q in [(-1.0, 0.5), (2.0, 0.5), (2.0, 0.0), (0.5, -1.0), (0.5, 2.0), (0.5, 0.5)]:
        assert not A.hit_test_path(q, p, FillRule.EvenOdd, 0.1), q
    assert A.hit_test_path((0.5, 0.5), p, FillRule.NonZero, 0.1)
    assert A.hit_test_path((0.2, 0.5), p, FillRule.EvenOdd, 0.1)
    assert A.hit_test_path((0.8, 0.5), p, FillRule.EvenOdd, 0.1)


def test_signed_area_reference_kats():
    """area.rs:80-130."""
    b = PathBuilder()
    b.add_polygon([(0.0, 0.0), (1.0, 0.0), (1.0, 1.0), (0.0, 1.0)], True)
    b.add_polygon([(0.0, 0.0), (0.0, 1.0), (1.0, 1.0), (1.0, 0.0)], True)
    two = b.build()
    assert A.approximate_signed_area(0.01, two) == 0.0  # +1 and -1
    b = PathBuilder()
    b.add_polygon([(0.0, 0.0), (1.0, 0.0), (1.0, 1.0), (0.0, 1.0)], True)
    assert A.approximate_signed_area(0.01, b.build()) == 1.0
    cross = [(0.0, 1.0), (1.0, 1.0), (1.0, 0.0), (2.0, 0.0), (2.0, 1.0), (3.0, 1.0), (3.0, 2.0),
             (2.0, 2.0), (2.0, 3.0), (1.0, 3.0), (1.0, 2.0), (0.0, 2.0)]
    b = PathBuilder()
    b.add_polygon(cross, True)
    assert A.approximate_signed_area(0.01, b.build()) == 5.0


def _paths():
    out = [p for _, p in star_paths(5, seed=8, max_points=120)]
    sb = stroke_batch(8, 4, 64, 1024.0, seed=2)
    out += [sb.path(i) for i in range(len(sb))]
    fb = fill_batch(6, seed=1)
    out += [fb.path(i) for i in range(len(fb))]
    b = PathBuilder()
    b.add_circle((100.0, 50.0), 40.0, True)
    b.add_ellipse((300.0, 60.0), (90.0, 30.0), 0.3, False)
    b.add_rectangle((10.0, 10.0, 80.0, 30.0), True)
    out.append(b.build())
    return out


@pytest.mark.parametrize("tol", [0.05, 0.5])
def test_winding_grid_matches_oracle(tol):
    rng = np.random.default_rng(11)
    for k, p in enumerate(_paths()):
        pts = np.asarray(p.points, np.float32)
        lo, hi = pts.min(axis=0) - 3, pts.max(axis=0) + 3
        gx, gy = np.meshgrid(np.linspace(lo[0], hi[0], 48), np.linspace(lo[1], hi[1], 48))
        xy = np.concatenate([np.stack([gx.ravel(), gy.ravel()], 1).astype(np.float32),
                             pts[rng.integers(0, len(pts), 64)],  # on vertices exactly
                             (pts[:-1][:64] + pts[1:][:64]) * np.float32(0.5)])
        for rule in (FillRule.EvenOdd, FillRule.NonZero):
            w, ins = A.path_winding_numbers(xy, p, tol, rule)
            ref = O.winding_numbers(p, xy, tol)
            bad = np.nonzero(w != ref)[0]
            assert len(bad) == 0, f"path {k}: {len(bad)} winding numbers differ, first " \
                                  f"{xy[bad[0]]}: {w[bad[0]]} != {ref[bad[0]]}"
            want = (ref % 2 != 0) if rule == FillRule.EvenOdd else (ref != 0)
            assert np.array_equal(ins, want)


def test_area_matches_oracle_bit_exact():
    for tol in (0.01, 0.05, 1.0):
        for k, p in enumerate(_paths()):
            a = A.approximate_signed_area(tol, p)
            r = O.signed_area(p, tol)
            assert np.float32(a).view(np.uint32) == np.float32(r).view(np.uint32), (k, tol, a, r)


def test_batched_forms_match_oracle():
    """Batched winding numbers and areas over device-resident paths, with ragged inputs:
    empty paths, a path of one Begin/End, and a query naming a path past the batch."""
    paths = _paths()
    b0 = PathBuilder()
    b0.begin((5.0, 5.0))
    b0.end(False)
    lone = b0.build()
    empty = PathBuilder().build()
    paths = [empty] + paths[:10] + [lone] + paths[10:] + [empty]
    batch = PathBatch.from_paths(paths)
    db = DeviceBatch(batch)
    rng = np.random.default_rng(3)
    n = 6000
    qp = rng.integers(0, len(paths) + 1, n).astype(np.uint32)  # len(paths) is past the end
    xy = np.zeros((n, 2), np.float32)
    for i in range(n):
        if qp[i] < len(paths) and len(paths[qp[i]].points):
            pts = np.asarray(paths[qp[i]].points)
            lo, hi = pts.min(axis=0), pts.max(axis=0)
            xy[i] = lo + rng.random(2) * (hi - lo)
    dq, dp = DeviceBuffer.from_numpy(xy.reshape(-1)), DeviceBuffer.from_numpy(qp)
    dw, di = DeviceBuffer(4 * n), DeviceBuffer(n)
    A.winding_numbers_device(db, dq, dp, n, 0.05, FillRule.NonZero, dw, di)
    w = dw.download(np.int32, n)
    ins = di.download(np.uint8, n)
    ref = np.zeros(n, np.int32)
    for j, p in enumerate(paths):
        sel = np.nonzero(qp == j)[0]
        if len(sel):
            ref[sel] = O.winding_numbers(p, xy[sel], 0.05)
    assert np.array_equal(w, ref)
    assert np.array_equal(ins.astype(bool), ref != 0)
    assert np.any(ref != 0)
    da = DeviceBuffer(4 * len(paths))
    A.signed_areas_device(db, 0.05, da)
    a = da.download(np.float32, len(paths))
    r = np.array([O.signed_area(p, 0.05) for p in paths], np.float32)
    assert np.array_equal(a.view(np.uint32), r.view(np.uint32))
    assert a[0] == 0.0 and a[11] == 0.0


def test_headline_batch_areas():
    """Areas of the 10,000-path stroke workload (open paths, implicitly closed) against the
    oracle, bit-exact."""
    batch = stroke_batch(10000, 8, 256, 2048.0, seed=0)
    db = DeviceBatch(batch)
    da = DeviceBuffer(4 * len(batch))
    A.signed_areas_device(db, 0.05, da)
    a = da.download(np.float32, len(batch))
    idx = np.arange(0, len(batch), 7)
    r = np.array([O.signed_area(batch.path(int(i)), 0.05) for i in idx], np.float32)
    assert np.array_equal(a[idx].view(np.uint32), r.view(np.uint32))


def test_tolerance_precondition():
    p = _kat_path()
    with pytest.raises(TessellationError) as e:
        A.approximate_signed_area(0.0, p)
    assert e.value.code == 4
    with pytest.raises(TessellationError):
        A.hit_test_path((0.5, 0.5), p, FillRule.EvenOdd, float("nan"))
