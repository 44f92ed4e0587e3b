"""The oracle's floatexp series mode (an extension of the reference: the series
test and skip of mandelbrot.cpp:138-142, 160-178 with each real held as a double
mantissa and an int exponent), pinned on the CPU:
* where the reference's doubles are normal numbers it gives the reference's
  results bit for bit (stop index, skip, counts and smoothing), so on those
  views the reference itself pins it;
* its coefficients equal the reference's series recurrence (mandelbrot.cpp:112-131,
  with the new-index quirk) evaluated in mpmath, also where the doubles overflow;
* its first unstable index equals the first index at which the stability test
  holds in exact arithmetic on the same operands;
* with its skip, the reference's HP continuation gives the counts of direct
  iteration from z = c on deep views where the double test degenerates."""
import json
import os

import mpmath as mp
import numpy as np
import pytest

import oracle_lib as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _views():
    return json.load(open(os.path.join(GOLD, "views.json")))


def _view(name, nr, nc, width=None, N=None):
    v = _views()[name]
    o = O.OView(nr, nc)
    o.set_view(int(N or v["N"]), v["re"], v["im"], width or v["width"])
    return o, v


@pytest.mark.parametrize("name,nr,nc,width", [("t11", 24, 32, None), ("t13", 24, 32, None),
                                              ("c2_seed1", 18, 32, None), ("t13", 12, 16, "3e-23")])
def test_floatexp_equals_the_reference_where_its_doubles_are_normal(name, nr, nc, width):
    """Both methods (the reference's HP continuation and the FP64 perturbation),
    the stop index, the skip point and the skip term are identical in the two
    series modes on views whose series terms stay in the double range."""
    o, v = _view(name, nr, nc, width)
    o.precompute(4)
    rc = np.array([(r, c) for r in range(nr) for c in range(nc)], dtype=np.int32)
    ref = o.render_samples(rc, 4)
    pert = o.render_samples(rc, 4, pert=True)
    starts = [o.sa_start(int(r), int(c)) for r, c in rc[::7]]
    o.set_series_mode(1)
    assert np.array_equal(o.render_samples(rc, 4), ref)
    assert np.array_equal(o.render_samples(rc, 4, pert=True), pert)
    for (r, c), (ua, found, d) in zip(rc[::7], starts):
        ua2, found2, d2 = o.sa_start(int(r), int(c))
        assert (ua2, found2) == (ua, found) and np.array_equal(d2, d)
    assert len({s[0] for s in starts}) > 1  # the stop index varies over the view


def test_floatexp_coefficients_are_the_truncated_hp_values():
    """The mantissa is the HP value truncated to 53 bits and scaled into [1/2, 1):
    times 2^exponent it is the reference's descend()ed double wherever that is a
    normal double."""
    o, v = _view("t13", 8, 8)
    o.precompute(2)
    _, _, A, B, Cc = o.series()
    for (m, e), d in zip(o.series_fx(), (A, B, Cc)):
        nz = m != 0
        assert np.all((np.abs(m[nz]) >= 0.5) & (np.abs(m[nz]) < 1.0))
        normal = nz & (np.abs(d) >= 2.0 ** -1022) & np.isfinite(d)
        assert normal.sum() > 0.9 * len(d)
        assert np.array_equal(np.ldexp(m[normal], e[normal]), d[normal])
        assert np.all(d[~nz] == 0)


def _deep_view_name():
    views = _views()
    for name in ("c3_detail1", "c5_detail1"):
        if name in views:
            return name
    pytest.skip("no detail-walk view in views.json")


def test_floatexp_coefficients_against_mpmath_beyond_the_double_range():
    """At width 1e-200 (the detail walk's 2560x1440 view, N 50000) the series
    coefficients leave the double range along the orbit (the reference's
    descend()ed doubles become inf); the floatexp values equal the recurrence
    evaluated from an mpmath orbit of the same reference point to 2^-50 relative,
    at every 97th index and at every index where the doubles overflow."""
    if "c3_detail1" not in _views():
        pytest.skip("c3_detail1 is not in views.json")
    o, v = _view("c3_detail1", 8, 8)
    re, im = v["re"], v["im"]
    o.precompute_at(re, im)
    L = o.ref_len()
    _, _, A, B, Cc = o.series()
    fx = o.series_fx()
    bad = ~np.isfinite(Cc[0::2]) | ~np.isfinite(Cc[1::2]) | ~np.isfinite(B[0::2]) | ~np.isfinite(B[1::2])
    check = set(range(1, L, 97)) | set(np.flatnonzero(bad)[::50].tolist())
    mp.mp.prec = int(o.bits()) + 64
    c = mp.mpc(re, im)
    worst = 0.0
    # the recurrence of mandelbrot.cpp:112-131 (new-index A, B on the right)
    x, a, b, cc = c, mp.mpc(1), mp.mpc(0), mp.mpc(0)
    for n in range(1, L):
        a = 2 * x * a + 1
        b = 2 * x * b + a * a
        cc = 2 * x * cc + 2 * a * b
        x = x * x + c
        if n in check:
            for (m, e), want in zip(fx, (a, b, cc)):
                got = mp.mpc(mp.ldexp(mp.mpf(m[2 * n]), int(e[2 * n])), mp.ldexp(mp.mpf(m[2 * n + 1]), int(e[2 * n + 1])))
                if want != 0:
                    worst = max(worst, float(abs(got - want) / abs(want)))
    assert worst < 2.0 ** -50, worst
    assert bad.sum() > 0  # the doubles left the range where the floatexp values did not


def test_floatexp_stop_index_is_the_exact_test():
    """On the deep detail-walk view: the first unstable index of the floatexp scan
    is the first i at which |B_i e^2|^2 tol < |C_i e^3|^2 holds in exact
    arithmetic on the same operands (the floatexp coefficients and e, e^2, e^3
    formed as the reference forms them), unless the exact test lies within 1e-12
    relative of its threshold there."""
    name = _deep_view_name()
    o, v = _view(name, 6, 8)
    o.precompute_at(v["re"], v["im"])
    L = o.ref_len()
    (_, _), (Bmn, Bex), (Cmn, Cex) = o.series_fx()
    mp.mp.prec = 200
    Bv = [mp.mpc(mp.ldexp(mp.mpf(Bmn[2 * i]), int(Bex[2 * i])), mp.ldexp(mp.mpf(Bmn[2 * i + 1]), int(Bex[2 * i + 1])))
          for i in range(L)]
    Cv = [mp.mpc(mp.ldexp(mp.mpf(Cmn[2 * i]), int(Cex[2 * i])), mp.ldexp(mp.mpf(Cmn[2 * i + 1]), int(Cex[2 * i + 1])))
          for i in range(L)]
    tol = mp.mpf(1e-10)
    fx_found = 0
    for r, c in [(0, 0), (2, 5), (5, 7), (3, 1)]:
        e = o.pixel_eps(r, c)
        ec = mp.mpc(float(e[0]), float(e[1]))
        e2 = ec * ec  # exact: the floatexp e^2, e^3 differ from these by rounding only
        e3 = ec * e2
        first = 0
        for i in range(1, L):
            if abs(Bv[i] * e2) ** 2 * tol < abs(Cv[i] * e3) ** 2:
                first = i
                break
        o.set_series_mode(1)
        ua = o.sa_start(r, c)[0]
        o.set_series_mode(0)
        if first != ua:
            i = min(first or L, ua or L)
            ratio = abs(Bv[i] * e2) ** 2 * tol / abs(Cv[i] * e3) ** 2
            assert abs(ratio - 1) < 1e-12, (r, c, ua, first)
        fx_found += int(ua != 0)
    assert fx_found > 0


def test_floatexp_counts_equal_direct_iteration_where_the_double_test_degenerates():
    """Width 1e-200 (the detail walk's view, N 50000), 6 x 8 samples: with the
    double series every sample skips to the orbit's end (e^2 underflows, no index
    tests unstable), so every count is N; with the floatexp series the reference
    method (skip + HP continuation) and the GPU method (skip + FP64 perturbation)
    give, on the first 12 samples, the counts of direct iteration from z = c at
    750 bits."""
    if "c3_detail1" not in _views():
        pytest.skip("c3_detail1 is not in views.json")
    nr, nc = 6, 8
    o, v = _view("c3_detail1", nr, nc)
    o.precompute(4)
    rc = np.array([(r, c) for r in range(nr) for c in range(nc)], dtype=np.int32)
    dbl = o.render_samples(rc, 4, pert=True)
    assert (dbl["iterations"] == v["N"]).all() and o.sa_start(1, 1)[0] == 0
    o.set_series_mode(1)
    ref = o.render_samples(rc, 4)
    pert = o.render_samples(rc, 4, pert=True)
    assert len(np.unique(ref["iterations"])) > 10
    N = int(v["N"])
    with mp.workprec(750):
        sz = mp.mpf(v["width"]) / nc
        for k, (r, c) in enumerate(rc[:12]):
            cr = mp.mpf(v["re"]) + (int(c) - nc // 2) * sz
            ci = mp.mpf(v["im"]) + (nr // 2 - int(r) - 1) * sz
            zr, zi, want = cr, ci, N
            for i in range(N):
                a, b = float(zr), float(zi)
                if a * a + b * b > 1024.0 ** 2:
                    want = i
                    break
                zr, zi = zr * zr - zi * zi + cr, 2 * zr * zi + ci
            assert ref["iterations"][k] == want and pert["iterations"][k] == want, (r, c, want)
