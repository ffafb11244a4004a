"""Shared comparison helpers and input corpora for the parity tests."""
import json
import os

import numpy as np

from lyon_amd.path import PathBuilder, polygon

HERE = os.path.dirname(os.path.abspath(__file__))


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


def fixture_path(subs, earcut=False):
    """A path from a fuzz_tests.rs / earcut_tests.rs fixture (tests/golden/*.json)."""
    b = PathBuilder()
    for s in subs:
        pts = s if earcut else s["points"]
        if not pts:
            continue
        if earcut:
            b.begin((pts[0][0], pts[0][0]))  # earcut_tests.rs:20192 begins at (x0, x0)
            for p in pts:
                b.line_to(p)
            b.end(True)
        else:
            b.begin(pts[0])
            for p in pts[1:]:
                b.line_to(p)
            b.end(s["closed"])
    return b.build()


def fuzz_fixtures():
    cases = json.load(open(os.path.join(HERE, "golden", "fuzz_polygons.json")))
    return [(name, fixture_path(subs)) for name, subs in cases.items()]


def earcut_fixtures():
    cases = json.load(open(os.path.join(HERE, "golden", "earcut_polygons.json")))
    return [("earcut_" + name, fixture_path(rings, earcut=True)) for name, rings in cases.items()]


# fill_tests.rs KAT polygons: (name, points, rotation step, EvenOdd triangle count)
FILL_KATS = [
    ("simple_triangle", [(0, 0), (1, 1), (0, 1)], 0.01, 1),
    ("simple_split", [(0, 0), (2, 1), (2, 3), (1, 2), (0, 3)], 0.01, 3),
    ("simple_merge_split", [(0, 0), (1, 1), (2, 0), (2, 3), (1, 2), (0, 3)], 0.01, 4),
    ("simple_aligned", [(0, 0), (1, 0), (2, 0), (2, 1), (2, 2), (1, 2), (0, 2), (0, 1)], 0.01, 6),
    ("simple_1", [(0, 0), (1, 1), (2, 0), (1, 3), (0.5, 4), (0, 3)], 0.01, 4),
    ("simple_2", [(0, 0), (1, 0), (2, 0), (3, 0), (3, 1), (3, 2), (3, 3), (2, 3), (1, 3), (0, 3),
                  (0, 2), (0, 1)], 0.01, 10),
]


def hole_path():
    """fill_tests.rs test_hole_1."""
    b = PathBuilder()
    b.begin((-11, 5)); b.line_to((0, -5)); b.line_to((10, 5)); b.end(True)
    b.begin((-5, 2)); b.line_to((0, -2)); b.line_to((4, 2)); b.end(True)
    return b.build()


def overlapping_squares_path():
    """fill_tests.rs test_simple_double_merge-like: two overlapping squares."""
    b = PathBuilder()
    b.begin((20, 20)); b.line_to((60, 20)); b.line_to((60, 60)); b.line_to((20, 60)); b.end(True)
    b.begin((40, 10)); b.line_to((70, 40)); b.line_to((40, 70)); b.line_to((10, 40)); b.end(True)
    return b.build()


def star_paths(n_paths, seed=5, max_points=80):
    """Self-intersecting random stars: points at unsorted random angles."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n_paths):
        n = int(rng.integers(5, max_points))
        ang = rng.uniform(0, 2 * np.pi, n)
        r = rng.uniform(0.2, 1.0, n) * 100
        pts = np.stack([200 + r * np.cos(ang), 200 + r * np.sin(ang)], 1).astype(np.float32)
        out.append((f"star{k}", polygon([tuple(p) for p in pts], True)))
    return out


def fill_corpus():
    """Every fill input of the parity tests: KAT polygons, lyon's fuzz and earcut fixtures,
    the plain workload path and random self-intersecting stars."""
    from lyon_amd import workloads
    out = [(name, polygon(pts, True)) for name, pts, _, _ in FILL_KATS]
    out.append(("hole", hole_path()))
    out.append(("overlapping_squares", overlapping_squares_path()))
    out.append(("plain", workloads.plain_path(0)))
    out += fuzz_fixtures()
    out += earcut_fixtures()
    out += star_paths(12)
    return out


def flattened_subpaths(path, tolerance):
    """The flattened sub-paths of a path as float32 polylines [first, ..., last], each
    implicitly closed, from the oracle's segment flattening (one call for all curves)."""
    from oracle import pyoracle as O
    verbs = np.asarray(path.verbs)
    pts = np.asarray(path.points, np.float32).reshape(-1, 2)
    npv = np.array([1, 2, 3, 1, 1, 0], np.int64)[verbs]
    first = np.cumsum(npv) - npv
    curves = [i for i, v in enumerate(verbs) if v in (1, 2)]
    flat = {}
    if curves:
        kinds = np.array([verbs[i] for i in curves], np.int32)
        ctrl = np.zeros((len(curves), 4, 2), np.float32)
        for k, i in enumerate(curves):
            ctrl[k, 0] = pts[first[i] - 1]
            for j in range(verbs[i] + 1):
                ctrl[k, j + 1] = pts[first[i] + j]
        fp, fo = O.flatten_segments(kinds, ctrl, tolerance)
        for k, i in enumerate(curves):
            flat[i] = fp[fo[k]:fo[k + 1]]
    subs, cur = [], None
    for i, v in enumerate(verbs):
        if v == 3:
            cur = [pts[first[i]]]
        elif v == 0:
            cur.append(pts[first[i]])
        elif v in (1, 2):
            cur.extend(flat[i])
        else:
            subs.append(np.asarray(cur, np.float32).reshape(-1, 2))
            cur = None
    return subs


def area64(subs):
    """Shoelace area of implicitly closed polylines, in float64."""
    a = 0.0
    for s in subs:
        d = s.astype(np.float64) - s[0].astype(np.float64)
        a += 0.5 * float(np.sum(d[:-1, 0] * d[1:, 1] - d[:-1, 1] * d[1:, 0]))
    return a


def winding64(subs, xy):
    """Textbook winding numbers of points xy (k, 2) against the closed polylines: edges
    whose y range [min, max) holds the point's y and that cross its horizontal line at or
    left of it count +1 going down (+y), -1 going up; float64.  Also returns each point's
    distance to the nearest edge."""
    q = np.asarray(xy, np.float64).reshape(-1, 2)
    w = np.zeros(len(q), np.int64)
    dmin = np.full(len(q), np.inf)
    for s in subs:
        a = s.astype(np.float64)
        b = np.roll(a, -1, axis=0)
        for (x0, y0), (x1, y1) in zip(a, b):
            lo, hi = min(y0, y1), max(y0, y1)
            dy = y1 - y0
            ex, ey = x1 - x0, y1 - y0
            L2 = ex * ex + ey * ey
            if L2 > 0:
                t = np.clip(((q[:, 0] - x0) * ex + (q[:, 1] - y0) * ey) / L2, 0, 1)
                dmin = np.minimum(dmin, np.hypot(q[:, 0] - (x0 + t * ex), q[:, 1] - (y0 + t * ey)))
            if dy == 0:
                continue
            inr = (q[:, 1] >= lo) & (q[:, 1] < hi)
            xc = x0 + (q[:, 1] - y0) * (x1 - x0) / dy
            w += np.where(inr & (xc <= q[:, 0]), 1 if dy > 0 else -1, 0)
    return w, dmin


def f32_exact(text):
    """Correctly rounded (half to even) float32 of a decimal string, by exact rational
    arithmetic: an independent restatement of what Rust's str::parse::<f32> returns."""
    from fractions import Fraction
    neg = text.startswith("-")
    x = Fraction(text.lstrip("-") if not text.lstrip("-").startswith(".") else "0" + text.lstrip("-"))
    if x == 0:
        return np.float32(-0.0 if neg else 0.0)
    k = x.numerator.bit_length() - x.denominator.bit_length()
    if Fraction(2) ** k > x:
        k -= 1
    if k > 127:
        bits = 0x7f800000
    else:
        s = 23 - k if k >= -126 else 149
        y = x * Fraction(2) ** s
        q, r = divmod(y.numerator, y.denominator)
        c = 2 * r - y.denominator
        if c > 0 or (c == 0 and q & 1):
            q += 1
        if k >= -126:
            if q == 1 << 24:
                q, k = 1 << 23, k + 1
            bits = 0x7f800000 if k > 127 else ((k + 127) << 23) | (q - (1 << 23))
        else:
            bits = q
    if neg:
        bits |= 0x80000000
    return np.array([bits], np.uint32).view(np.float32)[0]


def number_corpus(n=3000, seed=0):
    """Decimal strings in the parser's number syntax: random digit counts (1-25), points,
    exponents from -60 to 45, and the hard cases (ties, FLT_MAX, the subnormal edge)."""
    rng = np.random.default_rng(seed)
    out = ["0", "-0", "0.1", "1", "16777216", "16777217", "16777218", "16777219", "33554435",
           "3.4028235e38", "3.4028236e38", "3.40282357e38", "3.4028234e38", "1e39", "1e-45",
           "7e-46", "7.1e-46", "7.006e-46", "1.401298464324817e-45", "1.1754943e-38",
           "1.17549435e-38", "1.1754942e-38", "0.000000000000000000000000000000000000000000001",
           "123456789012345678901234567890", "1.5", "2.5", "0.3", "1e10", "1.e5", ".5", "-.25",
           "1e-5", "9999999999999999999", "99999999999999999999", "4.9e-324", "1e-400",
           "1e400", "0e999", "8388608.5", "8388609.5", "0.1e1", "100000000000000000000e-20"]
    for _ in range(n):
        nd = int(rng.integers(1, 26))
        digits = "".join(str(int(d)) for d in rng.integers(0, 10, nd))
        p = int(rng.integers(0, nd + 1))
        s = digits[:p] + ("." + digits[p:] if p < nd else "")
        if s.startswith("."):
            s = "0" + s if rng.random() < 0.5 else s
        if rng.random() < 0.6:
            s += ("e" if rng.random() < 0.5 else "E") + ("-" if rng.random() < 0.5 else "") + \
                 str(int(rng.integers(0, 46)))
        if rng.random() < 0.3:
            s = "-" + s
        out.append(s)
    return out


def to_svg(path):
    """A path as absolute SVG path data; '%.9g' gives every f32 back exactly under a
    correctly rounded parse."""
    verbs = np.asarray(path.verbs)
    pts = np.asarray(path.points, np.float32).reshape(-1, 2)
    f = lambda p: "%.9g %.9g" % (float(p[0]), float(p[1]))
    out, pi = [], 0
    for v in verbs:
        if v == 3:
            out.append("M " + f(pts[pi]))
            pi += 1
        elif v == 0:
            out.append("L " + f(pts[pi]))
            pi += 1
        elif v == 1:
            out.append("Q " + f(pts[pi]) + " " + f(pts[pi + 1]))
            pi += 2
        elif v == 2:
            out.append("C " + f(pts[pi]) + " " + f(pts[pi + 1]) + " " + f(pts[pi + 2]))
            pi += 3
        elif v == 4:
            out.append("Z")
            pi += 1
    return " ".join(out)


def random_svg(rng, n_cmds=12, arcs=True, corrupt=0.0):
    """Random SVG path data over every command (absolute and relative, implicit repeats,
    mixed separators, exponents, signs glued to the previous number), optionally with
    stray characters."""
    def num():
        r = rng.random()
        if r < 0.3:
            s = str(int(rng.integers(-500, 500)))
        elif r < 0.7:
            s = "%.*f" % (int(rng.integers(0, 5)), rng.uniform(-500, 500))
        elif r < 0.85:
            s = ("%.3e" % rng.uniform(-500, 500)).replace("e+", "e")
        else:
            s = ("-" if rng.random() < 0.5 else "") + "." + str(int(rng.integers(0, 1000)))
        return s

    def sep(nxt):
        if nxt.startswith("-") and rng.random() < 0.5:
            return ""
        return [" ", ",", ", ", "\n", "\t", "  "][int(rng.integers(0, 6))]

    def args(n, flags=()):
        out = ""
        for k in range(n):
            tok = str(int(rng.integers(0, 2))) if k in flags else num()
            out += sep(tok) + tok
        return out

    spec = {"M": (2, ()), "L": (2, ()), "H": (1, ()), "V": (1, ()), "Q": (4, ()), "T": (2, ()),
            "C": (6, ()), "S": (4, ()), "A": (7, (3, 4)), "Z": (0, ())}
    letters = [c for c in spec if arcs or c != "A"]
    s = "M" + args(2)
    for _ in range(n_cmds):
        c = letters[int(rng.integers(0, len(letters)))]
        if rng.random() < 0.5:
            c = c.lower()
        n, fl = spec[c.upper()]
        s += sep("x") + c + args(n, fl)
        if n and rng.random() < 0.3:  # implicit repetition
            s += args(n, fl)
        if c in "Zz" and rng.random() < 0.8:
            s += " M" + args(2)
    if corrupt and rng.random() < corrupt:
        k = int(rng.integers(0, len(s)))
        s = s[:k] + ["x", "*", "+", "#", "e"][int(rng.integers(0, 5))] + s[k:]
    return s
