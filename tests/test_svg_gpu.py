"""SVG path data ingest on the GPU (lyon_extra's PathParser, parser.rs) against the oracle:
verbs, errors, numbers and points equal (arcs with correctly rounded sin / cos / tan on both
sides)."""
import numpy as np
import pytest

from helpers import number_corpus, random_svg, to_svg
from lyon_amd import svg as S
from lyon_amd.path import PathBatch
from lyon_amd.workloads import stroke_batch
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu


def _err(e):
    return [0, 0, 0, 0] if e is None else [e.kind, e.line, e.column, ord(e.char) if e.char else 0]


def test_reference_kats():
    """parser.rs:498-700."""
    p, e = S.parse_path("M 0 0 L 1 0 L 1 1 L 0 1 Z")
    assert e is None and p.verbs.tolist() == [3, 0, 0, 0, 4]
    assert S.parse_path("M 0 0 1.0 L 1 0 2.0 L 1 1 3.0 L 0 1 4.0 Z", S.ParserOptions(1))[1] is None
    p, e = S.parse_path("0 0 0 1 1 1.0 2 2 2.0 3 3 3", S.ParserOptions(1))
    assert e is None and p.verbs.tolist() == [3, 0, 0, 0, 5]
    e = S.parse_path("x 0 0 0", S.ParserOptions(1))[1]
    assert (e.kind, e.line, e.column, e.char) == (S.ERR_COMMAND, 0, 0, "x")
    e = S.parse_path("\n M 0 \n0 1 x 1 1 1", S.ParserOptions(1))[1]
    assert (e.kind, e.line, e.column, e.char) == (S.ERR_COMMAND, 2, 4, "x")
    p, e = S.parse_path("M 0.6.5")
    assert e is None and p.points.tolist() == [[np.float32(0.6), 0.5]]
    assert S.parse_path("M 1e-2 -1E3")[1] is None
    for bad in ["M 0 --1", "M 0 1ee2", "M 0 1e--1", "M 0 *2", "M 0 e", "M 0 1e", "M 0 +1"]:
        assert S.parse_path(bad)[1].kind == S.ERR_NUMBER, bad
    for s in ["M 0 0 | xxxxxx", "M 0 0| xxxxxx", "| xxxxxx", "    | xxxxxx"]:
        assert S.parse_path(s, S.ParserOptions(stop_at="|"))[1] is None, s
    p1, e1 = S.parse_path("M 0 0 Z L 1 1 2 2 L 3 3 Z M 4 4", S.ParserOptions(stop_at="|"))
    p2, e2 = S.parse_path("M 0 0 Z", S.ParserOptions(stop_at="|"))
    assert e1.kind == S.ERR_MISSING_MOVE_TO and e2 is None
    assert p1.verbs.tolist() == p2.verbs.tolist() and p1.points.tolist() == p2.points.tolist()
    for s in ["", " "]:
        p, e = S.parse_path(s)
        assert e is None and len(p.verbs) == 0


def test_numbers_bit_exact():
    corpus = number_corpus(4000, seed=5)
    batch, errs = S.parse_paths([f"M {s} 0" for s in corpus])
    assert all(e is None for e in errs)
    xs = batch.points[batch.point_offsets[:-1], 0]
    want = np.array([O.parse_f32(s) for s in corpus], np.float32)
    bad = np.nonzero(xs.view(np.uint32) != want.view(np.uint32))[0]
    assert len(bad) == 0, [(corpus[i], xs[i], want[i]) for i in bad[:5]]


def _exact_f32(fr):
    """Round a positive Fraction to f32 bits, ties to even (the rounding of str::parse)."""
    from fractions import Fraction
    if fr == 0:
        return 0
    lo, hi = 0, 0x7f800000
    while hi - lo > 1:  # largest bits whose value <= fr
        mid = (lo + hi) // 2
        if Fraction(float(np.uint32(mid).view(np.float32))) <= fr:
            lo = mid
        else:
            hi = mid
    if lo == 0x7f7fffff:
        up = Fraction(2) ** 128  # the next step past the largest float rounds to infinity
    else:
        up = Fraction(float(np.uint32(lo + 1).view(np.float32)))
    m = (Fraction(float(np.uint32(lo).view(np.float32))) + up) / 2
    if fr < m or (fr == m and lo % 2 == 0):
        return lo
    return lo + 1


def _decimal(fr, digits_after):
    """Plain decimal text of a Fraction whose expansion ends within digits_after places."""
    ip = fr.numerator // fr.denominator
    rest = fr - ip
    out = []
    for _ in range(digits_after):
        rest *= 10
        d = rest.numerator // rest.denominator
        out.append(str(d))
        rest -= d
        if rest == 0:
            break
    assert rest == 0
    return f"{ip}." + "".join(out)


def test_numbers_past_19_digits_on_a_midpoint():
    """Numbers of 25 to 160 significant digits that are an f32 midpoint exactly, or lie one
    unit of their 60th (or last + 20th) digit above or below it: rounded like str::parse
    (correctly, ties to even), checked against exact rational rounding and the oracle's
    strtof; values from subnormals to the largest float."""
    from fractions import Fraction
    rng = np.random.default_rng(21)
    bits = np.concatenate([rng.integers(1, 0x7f7fffff, 60, dtype=np.int64),
                           [1, 2, 0x7fffff, 0x800000, 0x3f800000, 0x7f7ffffe, 0x7f7fffff]])
    strings, want = [], []
    for b in bits.tolist():
        lo = Fraction(float(np.uint32(b).view(np.float32)))
        up = Fraction(2) ** 128 if b == 0x7f7fffff else \
            Fraction(float(np.uint32(b + 1).view(np.float32)))
        m = (lo + up) / 2
        text = _decimal(m, 200)
        sig = len(text.replace(".", "").lstrip("0"))
        if sig <= 19:
            continue
        eps = Fraction(1, 10 ** (len(text.split(".")[1]) + 20))
        for v, t in ((m, text), (m + eps, _decimal(m + eps, 260)), (m - eps, _decimal(m - eps, 260))):
            strings.append(t)
            want.append(_exact_f32(v))
    assert len(strings) > 150
    batch, errs = S.parse_paths([f"M {t} 0" for t in strings])
    assert all(e is None for e in errs), [(t, e) for t, e in zip(strings, errs) if e][:3]
    got = batch.points[batch.point_offsets[:-1], 0].view(np.uint32).astype(np.int64)
    want = np.array(want, np.int64)
    bad = np.nonzero(got != want)[0]
    assert len(bad) == 0, [(strings[i][:40], hex(got[i]), hex(want[i])) for i in bad[:5]]
    ref = np.array([O.parse_f32(t) for t in strings], np.float32).view(np.uint32)
    assert np.array_equal(ref.astype(np.int64), want)


@pytest.mark.parametrize("arcs", [False, True])
def test_random_svg_matches_oracle(arcs):
    rng = np.random.default_rng(17 if arcs else 4)
    strings = [random_svg(rng, int(rng.integers(1, 40)), arcs=arcs, corrupt=0.25)
               for _ in range(1500)]
    strings += ["", "   ", "M 1 2", "Z", "z 1 1", "M 0 0 Z Z", "m 1 1 2 2 3 3 z 4 4"]
    if arcs:  # the densest output per byte (7 quadratics per 12 characters)
        strings.append("M0 0A1 1 0 111 0" + " 1 1 0 110 0 1 1 0 111 0" * 50)
    batch, errs = S.parse_paths(strings)
    n_err = 0
    for i, s in enumerate(strings):
        rv, rp, re = O.parse_svg(s)
        p = batch.path(i)
        assert _err(errs[i]) == re.tolist(), (s, errs[i], re)
        n_err += re[0] != 0
        assert p.verbs.tolist() == rv.tolist(), s
        assert p.points.shape == rp.shape, s
        if not arcs:
            assert np.array_equal(p.points.view(np.uint32), rp.view(np.uint32)), s
        else:
            # arcs: sin / cos / tan rounded correctly to f32 on both sides (NaN where an
            # overflowing number made one)
            assert np.array_equal(p.points, rp, equal_nan=True), s
    assert 0 < n_err < len(strings)


def test_round_trip_of_the_stroke_workload():
    """Full size: the 10,000-path stroke workload (and a closed variant) written as SVG
    path data with 9 significant digits parses back to the same verbs and bit-identical
    points."""
    for closed in (False, True):
        src = stroke_batch(10000 if not closed else 2000, 8, 256, 2048.0, seed=0, closed=closed)
        strings = [to_svg(src.path(i)) for i in range(len(src))]
        batch, errs = S.parse_paths(strings)
        assert all(e is None for e in errs)
        assert np.array_equal(batch.verbs, src.verbs)
        assert np.array_equal(batch.verb_offsets, src.verb_offsets)
        assert np.array_equal(batch.point_offsets, src.point_offsets)
        assert np.array_equal(batch.points.view(np.uint32), src.points.view(np.uint32))


def test_non_ascii_and_options():
    p, e = S.parse_path("M 0 0 L 1 1 L 2 é")
    assert e.kind == S.ERR_NON_ASCII
    assert O.parse_svg("M 0 0 L 1 1 L 2 é")[2][0] == S.ERR_NON_ASCII
    # text after the stop character is never read
    p, e = S.parse_path("M 0 0 L 1 1 | é", S.ParserOptions(stop_at="|"))
    assert e is None and p.verbs.tolist() == [3, 0, 5]
    with pytest.raises(ValueError):
        S.parse_path("M 0 0", S.ParserOptions(stop_at="é"))


def test_attributes_syntax_matches_oracle():
    """The extended syntax (ParserOptions::num_attributes numbers after every endpoint) on
    random strings: verbs, points and errors as the oracle parses them."""
    rng = np.random.default_rng(23)
    for nattr in (1, 3):
        strings = []
        for _ in range(300):
            s = random_svg(rng, int(rng.integers(1, 20)), arcs=True, corrupt=0.1)
            strings.append(s)
        batch, errs = S.parse_paths(strings, S.ParserOptions(num_attributes=nattr))
        for i, s in enumerate(strings):
            rv, rp, re = O.parse_svg(s, num_attributes=nattr)
            p = batch.path(i)
            assert _err(errs[i]) == re.tolist(), s
            assert p.verbs.tolist() == rv.tolist(), s
            assert np.array_equal(p.points, rp, equal_nan=True), s


def test_text_to_triangles_on_the_device():
    """SVG text parsed on the device feeds the batched stroke and fill directly; the
    triangles are byte-identical to those of the original paths."""
    from lyon_amd.batch import DeviceBatch, DeviceBuffer, FillBatcher, StrokeBatcher
    from lyon_amd.workloads import fill_batch, fill_options, stroke_options
    cases = [(stroke_batch(800, 8, 128, 2048.0, seed=4), stroke_options(800, tolerance=0.05),
              StrokeBatcher),
             (fill_batch(300, seed=2), fill_options(300, tolerance=0.05), FillBatcher)]
    for src, opts, Batcher in cases:
        strings = [to_svg(src.path(i)) for i in range(len(src))]
        text, offs = S.pack_strings(strings)
        parsed = S.parse_paths_device(DeviceBuffer.from_numpy(text), DeviceBuffer.from_numpy(offs),
                                      len(strings)).with_options(opts)
        assert parsed.n_errors == 0
        a = Batcher().run(parsed)
        b = Batcher().run(DeviceBatch(src, opts))
        assert a.n_vertices == b.n_vertices and a.n_indices == b.n_indices
        assert np.array_equal(a.all_vertices().view(np.uint32), b.all_vertices().view(np.uint32))
        assert np.array_equal(a.all_indices(), b.all_indices())
