"""GPU parity of lyon_geom flattening against the oracle: point counts and positions
bit-exact for lines/quadratics/cubics; arcs within 1e-5 of the bbox diagonal."""
import numpy as np
import pytest

from lyon_amd import arc_to_quadratics, flatten_arcs, flatten_segments
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu


def random_segments(n, seed, scale):
    rng = np.random.default_rng(seed)
    kinds = rng.integers(0, 3, n).astype(np.int32)
    ctrl = (rng.random((n, 4, 2)) * scale).astype(np.float32)
    return kinds, ctrl


@pytest.mark.parametrize("tol", [0.05, 0.1, 0.001, 1.0])
def test_segments_bit_exact(tol):
    kinds, ctrl = random_segments(3000, 1, 2048.0)
    gp, go = flatten_segments(kinds, ctrl, tol)
    rp, ro = O.flatten_segments(kinds, ctrl, tol)
    assert np.array_equal(go, ro), "point counts differ"
    assert np.array_equal(gp.view(np.uint32), rp.view(np.uint32)), "positions differ"


def test_segments_many_per_wave():
    """Runs of lines (one point each) put up to 1,024 segments under one wave's points,
    past the wave's LDS window of segment offsets; curves between the runs."""
    kinds, ctrl = random_segments(6000, 7, 2048.0)
    kinds[:2500] = 0
    kinds[3000:5500] = 0
    kinds[2700] = 2
    for tol in (0.05, 1.0):
        gp, go = flatten_segments(kinds, ctrl, tol)
        rp, ro = O.flatten_segments(kinds, ctrl, tol)
        assert np.array_equal(go, ro), "point counts differ"
        assert np.array_equal(gp.view(np.uint32), rp.view(np.uint32)), "positions differ"


def test_segments_degenerate():
    ctrl = np.zeros((4, 4, 2), np.float32)
    ctrl[1] = [[1, 1], [1, 1], [1, 1], [1, 1]]
    ctrl[2] = [[0, 0], [5, 0], [10, 0], [10, 0]]  # straight quadratic
    ctrl[3] = [[0, 0], [1e6, 1e6], [-1e6, 1e6], [3, 0]]  # huge
    kinds = np.array([1, 2, 1, 2], np.int32)
    gp, go = flatten_segments(kinds, ctrl, 0.05)
    rp, ro = O.flatten_segments(kinds, ctrl, 0.05)
    assert np.array_equal(go, ro)
    assert np.array_equal(gp.view(np.uint32), rp.view(np.uint32))
    # quadratic_bezier.rs:1381-1419: empty and straight-line quadratics flatten to 1 segment
    assert go[1] - go[0] == 1


def test_arcs():
    rng = np.random.default_rng(3)
    arcs = np.zeros((200, 7), np.float32)
    arcs[:, 0:2] = rng.random((200, 2)) * 1000
    arcs[:, 2:4] = rng.random((200, 2)) * 300 + 1
    arcs[:, 4] = rng.random(200) * 6.28
    arcs[:, 5] = (rng.random(200) - 0.5) * 12.0
    arcs[:, 6] = rng.random(200) * 3.0
    gp, go = flatten_arcs(arcs, 0.05)
    rp, ro = O.flatten_arcs(arcs, 0.05)
    assert np.array_equal(go, ro)
    diag = 2000.0
    assert np.max(np.abs(gp.astype(np.float64) - rp)) <= 1e-5 * diag
    gq, gqo = arc_to_quadratics(arcs)
    rq, rqo = O.arc_to_quads(arcs)
    assert np.array_equal(gqo, rqo)
    assert np.max(np.abs(gq.astype(np.float64) - rq)) <= 1e-5 * diag


def test_device_segments_match_host_and_oracle():
    """The device-memory flattening entry point gives the host entry point's bytes and the
    oracle's on the segments of a stroke workload batch."""
    from lyon_amd import workloads
    from lyon_amd.batch import DeviceBuffer, flatten_segments_device
    b = workloads.stroke_batch(600, 8, 256, 2048.0, seed=4)  # > 32768 segments: tiled scan
    kinds, ctrl = workloads.segments_of(b)
    n = len(kinds)
    dk = DeviceBuffer.from_numpy(kinds)
    dc = DeviceBuffer.from_numpy(ctrl.reshape(-1))
    doff = DeviceBuffer(4 * (n + 1))
    try:
        flatten_segments_device(dk, dc, n, 0.05, None, doff)
        raise AssertionError("expected a capacity error")
    except Exception as e:  # the count comes back with the capacity status
        total = e.needed
    out = DeviceBuffer(8 * total)
    assert flatten_segments_device(dk, dc, n, 0.05, out, doff) == total
    dp = out.download(np.float32, 2 * total).reshape(-1, 2)
    do = doff.download(np.uint32, n + 1)
    hp, ho = flatten_segments(kinds, ctrl.reshape(-1, 4, 2), 0.05)
    assert np.array_equal(do, ho)
    assert np.array_equal(dp.view(np.uint32), hp.view(np.uint32))
    sel = np.arange(0, n, 7)
    rp, ro = O.flatten_segments(kinds[sel], ctrl.reshape(-1, 4, 2)[sel], 0.05)
    for j, i in enumerate(sel):
        a = dp[do[i]:do[i + 1]]
        r = rp[ro[j]:ro[j + 1]]
        assert np.array_equal(a.view(np.uint32), r.view(np.uint32)), f"segment {i}"


def test_issue_652_degenerate_cubic_terminates():
    """cubic_bezier.rs:1648-1666: a cubic whose control points lie on one vertical line
    and return to its start flattens (and stops) at tolerances 1, 0.1 and 0.01; its last
    point is `to` exactly."""
    ctrl = np.array([[[-1061.0, -3327.0], [-1061.0, -3177.0], [-1061.0, -3477.0],
                      [-1061.0, -3327.0]]], np.float32)
    for tol in (1.0, 0.1, 0.01):
        gp, go = flatten_segments(np.array([2], np.int32), ctrl, tol)
        rp, ro = O.flatten_segments(np.array([2], np.int32), ctrl, tol)
        assert np.array_equal(go, ro) and 1 <= go[1] < 10000
        assert np.array_equal(gp.view(np.uint32), rp.view(np.uint32))
        assert gp[-1].tolist() == [-1061.0, -3327.0]


def test_arc_negative_flattening_step():
    """arc.rs:1207-1236: an arc of sweep -898 rad flattens to a finite list, and a
    negative sweep of -2 rad flattens to more than one segment."""
    arcs = np.array([[-100.0, -150.0, 50.0, 50.0, 0.982944787, -898.0, 0.0],
                     [0.0, 0.0, 100.0, 10.0, 0.2, -2.0, 0.0]], np.float32)
    tols = (0.100000001, 0.1)
    for k in range(2):
        gp, go = flatten_arcs(arcs[k:k + 1], tols[k])
        rp, ro = O.flatten_arcs(arcs[k:k + 1], tols[k])
        assert np.array_equal(go, ro)
        assert np.max(np.abs(gp.astype(np.float64) - rp)) <= 1e-5 * 300
    assert go[1] - go[0] > 1
