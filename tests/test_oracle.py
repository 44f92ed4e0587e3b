"""Pins the CPU oracle (test infrastructure) against known answers that do not come
from the oracle itself: mpmath arithmetic, closed forms, brute force on tiny
inputs, and invariants.  The reference has no tests, fixtures or golden vectors
(it ships none), so these are the oracle's anchors.
"""
import json
import math
import os

import mpmath as mp
import numpy as np
import pytest

import oracle_lib as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")


# --------------------------------------------------------------- HP arithmetic
def test_hp_ops_match_mpmath_round_down():
    """ohp restates GMP mpf's truncating arithmetic; mpmath with rounding 'd' at
    the same precision is an independent computation of the same results."""
    recs = json.load(open(os.path.join(GOLD, "hp_ops.json")))["records"]
    assert len(recs) >= 500
    for r in recs:
        got = O.hp_op(r["op"], r["a"], r["b"], r["iv"], r["dv"], r["bits"])
        assert got == (r["sign"], r["exp"], r["words"]), r


def test_hp_truncation_direction():
    """1/3 at 64-bit request: truncation gives a value below 1/3 that is within one
    unit of the last place; negating the input mirrors it (toward zero)."""
    s, e, w = O.hp_op(4, "1", None, 3, 0.0, 64)
    P = 32 * len(w)
    M = 0
    for x in w:
        M = (M << 32) | x
    with mp.workprec(1000):
        val = mp.mpf(M) * mp.mpf(2) ** (e - P)
        assert val < mp.mpf(1) / 3
        assert (mp.mpf(1) / 3 - val) < mp.mpf(2) ** (e - P)
    s2, e2, w2 = O.hp_op(4, "-1", None, 3, 0.0, 64)
    assert (s2, e2, w2) == (-1, e, w)


# ------------------------------------------------------ plain-double path KATs
def _hw_brute(c, N=1024, bailout=1024.0):
    """getIterationsHW semantics in Python floats (no FMA, same op order)."""
    z0 = complex(c)
    zr, zi = z0.real, z0.imag
    cr, ci = zr, zi
    for it in range(N):
        nr = (zr * zr - zi * zi) + cr
        ni = (2.0 * zr) * zi + ci
        zr, zi = nr, ni
        r2 = zr * zr + zi * zi
        if r2 > bailout * bailout:
            return it, np.float32(1.0 - math.log2(0.5 * math.log(r2) / math.log(bailout)))
    return N, np.float32(0.0)


@pytest.mark.parametrize("re,im", [("1", "0"), ("2", "0"), ("-2.5", "0"), ("0.5", "0.5"), ("0.3", "0"),
                                   ("-0.1", "0.9"), ("0.37", "-0.2")])
def test_plain_path_kats(re, im):
    it, sm = O.point_hw(re, im, 1024, 64)
    want_it, want_sm = _hw_brute(complex(float(re), float(im)))
    assert it == want_it
    assert np.float32(sm) == want_sm


def test_plain_path_closed_forms():
    # c = 2: 2 -> 6 -> 38 -> 1446 escapes on the third update (count 2)
    it, sm = O.point_hw("2", "0", 1024, 64)
    assert it == 2
    assert abs(sm - (1.0 - math.log2(math.log(1446.0) / math.log(1024.0)))) < 1e-6
    # c = 1: 1, 2, 5, 26, 677, 458330: count 4
    assert O.point_hw("1", "0", 1024, 64)[0] == 4
    # bounded orbits: -2 (fixed at 2), i (period 2), 0.25 (parabolic) never escape
    for re, im in [("-2", "0"), ("0", "1"), ("0.25", "0")]:
        assert O.point_hw(re, im, 1024, 64) == (1024, 0.0)


def test_cardioid_closed_forms_real_axis():
    """On the real axis the main cardioid is (-0.75, 0.25) and the period-2 disk
    is (-1.25, -0.75); the boundary points themselves are not inside."""
    inside = ["-1.2499999", "-1", "-0.7500001", "-0.7499999", "0", "0.2499999"]
    outside = ["-1.2500001", "-1.25", "-0.75", "0.25", "0.2500001", "-2"]
    for x in inside:
        assert O.point_in_cardioid(x, "0", 64), x
    for x in outside:
        assert not O.point_in_cardioid(x, "0", 64), x


def test_cardioid_parametric_boundary():
    """mu -> mu/2 - mu^2/4 maps the unit disk onto the main cardioid: |mu| = 1 -+ 1e-6
    lands just inside / just outside (away from the period-2 disk)."""
    for t in np.linspace(0.3, 2.6, 9):
        for rad, want in [(1 - 1e-6, True), (1 + 1e-6, False)]:
            mu = rad * np.exp(1j * t)
            z = complex(mu / 2 - mu * mu / 4)
            assert O.point_in_cardioid(repr(float(z.real)), repr(float(z.imag)), 64) == want, (t, rad)


# ---------------------------------------------------------------- precision
@pytest.mark.parametrize("width,cols,bits", [("3.0", 640, 64), ("1e-30", 7680, 109), ("1e-200", 2560, 672),
                                             ("1e-100", 3840, 341), ("1e-40", 1280, 140)])
def test_set_precision_table(width, cols, bits):
    """bits = max(64, (int)(64 - e + log2(1e-20))) with sz.re = m 2^e, m in [1/2, 1)."""
    o = O.OView(4, cols)
    o.set_view(100, "-0.5", "0", width)
    sz = mp.mpf(width) / cols
    e = int(mp.floor(mp.log(sz, 2))) + 1
    assert bits == max(64, int(64 - e + math.log2(1e-20)))
    assert o.bits() == bits


# ------------------------------------------------------------ series (A, B, C)
def test_series_at_zero_reference():
    """Reference point 0: X = 0 forever, so A = 1, B = A^2 = 1, C = 2 A B = 2."""
    o = O.OView(8, 8)
    o.set_view(50, "0", "0", "1e-20")
    n = o.precompute_at("0", "0")
    assert n == 50
    Xt, Xn, A, B, C = o.series()
    assert np.all(Xn[:100] == 0)
    assert np.all(A[2::2] == 1) and np.all(A[3::2] == 0)
    assert np.all(B[2::2] == 1) and np.all(C[2::2] == 2)


def test_series_A_is_the_derivative_and_B_has_the_quirk():
    """A_i = dX_i/dc exactly (Taylor); B_1 = A_1^2 = (2c+1)^2 by the reference's
    new-index recurrence (the Taylor coefficient would be A_0^2 = 1)."""
    c = mp.mpc("0.1234", "-0.3")
    o = O.OView(8, 8)
    o.set_view(12, "0.1234", "-0.3", "1e-20")
    o.precompute_at("0.1234", "-0.3")
    Xt, Xn, A, B, C = o.series()
    mp.mp.prec = 300
    x, dx = c, mp.mpc(1)
    for i in range(1, 10):
        dx = 2 * x * dx + 1
        x = x * x + c
        assert abs(complex(A[2 * i], A[2 * i + 1]) - complex(dx)) <= 1e-15 * abs(complex(dx))
    b1 = (2 * c + 1) ** 2
    assert abs(complex(B[2], B[3]) - complex(b1)) <= 1e-15 * abs(complex(b1))


# -------------------------------------------------------------- invariants
def test_plain_path_row_mirror_symmetry():
    """centre.im = 0: pixel im = (rows/2 - r - 1) sz, so rows r and rows-2-r are
    complex conjugates; truncation is odd and sq(conj z) = conj(sq z) bitwise."""
    o = O.OView(40, 60)
    o.set_N(500)
    g = o.render_rows(np.arange(40))
    for r in range(39):
        assert np.array_equal(g[r], g[38 - r])


def test_scale_up_then_down_round_trip():
    rng = np.random.default_rng(3)
    g = np.zeros((12, 16), dtype=O.ESCAPE_DTYPE)
    g["iterations"] = rng.integers(1, 2000, size=(12, 16))
    g["smoothing"] = rng.random((12, 16)).astype(np.float32)
    for sc in (2, 3, 4):
        back = O.downsample(O.upsample(g, sc), sc)
        tot = g["iterations"] + g["smoothing"].astype(np.float64)
        got = back["iterations"] + back["smoothing"].astype(np.float64)
        assert np.allclose(got, tot, rtol=0, atol=2e-3 * sc * sc)


def test_downsample_float_order():
    """at(r,c,sc) sums iterations + smoothing in float, r1 outer, c1 inner."""
    g = np.zeros((2, 2), dtype=O.ESCAPE_DTYPE)
    g["iterations"] = [[16777216, 1], [1, 1]]
    g["smoothing"] = [[0.5, 0.5], [0.5, 0.5]]
    out = O.downsample(g, 2)
    s = np.float32(0)
    for v in (16777216.5, 1.5, 1.5, 1.5):
        s = np.float32(s + np.float32(v))
    s = np.float32(s / np.float32(4))
    assert out["iterations"][0, 0] == int(s)
    assert out["smoothing"][0, 0] == np.float32(s - np.float32(int(s)))


# ----------------------------------------------- HP continuation, brute force
def test_hp_continuation_vs_mpmath_brute_force():
    """With a tolerance that stops the series at i = 1 (found = 0), the reference
    method iterates the whole orbit in HP from Y_0 = X_0 + descend(Y0 - X0); mpmath
    iterating z <- z^2 + c from the same start gives the same escape index."""
    v = json.load(open(os.path.join(GOLD, "views.json")))["t11"]
    nr, nc = 6, 8
    o = O.OView(nr, nc)
    o.set_view(v["N"], v["re"], v["im"], "1e-15")
    o.set_tol(1e-300)
    o.precompute()
    rc = np.array([(r, c) for r in range(nr) for c in range(nc)], dtype=np.int32)
    got = o.render_samples(rc)
    mp.mp.prec = 400
    sz = mp.mpf("1e-15") / nc
    for k, (r, c) in enumerate(rc):
        ua, found, d = o.sa_start(int(r), int(c))
        eps = o.pixel_eps(int(r), int(c))
        # the reference point itself has eps = 0: the test never fires and the
        # skip runs to the end of the (exact) reference orbit
        assert found == 0 or not eps.any()
        y0 = mp.mpc(mp.mpf(v["re"]) + (c - nc // 2) * sz, mp.mpf(v["im"]) + (nr // 2 - r - 1) * sz)
        z = y0
        n = v["N"]
        for i in range(1, v["N"]):
            z = z * z + y0
            zz = complex(z)
            if zz.real * zz.real + zz.imag * zz.imag > 1024.0 ** 2:
                n = i
                break
        assert got["iterations"][k] == n, (r, c)


def test_method_restatement_matches_reference_method():
    """pert64 (the GPU method) against the reference method on a deep view."""
    v = json.load(open(os.path.join(GOLD, "views.json")))["t13"]
    nr, nc = 24, 32
    o = O.OView(nr, nc)
    o.set_view(v["N"], v["re"], v["im"], "1e-24")
    o.precompute()
    rc = np.array([(r, c) for r in range(nr) for c in range(nc)], dtype=np.int32)
    a = o.render_samples(rc)
    b = o.render_samples(rc, pert=True)
    assert np.array_equal(a["iterations"], b["iterations"])


# ---- the learned skip predictor's oracle (oracle/mlp.c) and the tile probes ----

def _mlp_offsets():
    IN, H = 132, 64
    oW1 = 0
    oB1 = oW1 + H * IN
    oW2 = oB1 + H
    oB2 = oW2 + H * H
    oW3 = oB2 + H
    oB3 = oW3 + H
    return IN, H, oW1, oB1, oW2, oB2, oW3, oB3


def test_mlp_oracle_forward_closed_form():
    """W1 routes input j to unit j, W2 = I, W3 = 1, b3 = 0.5: y = 0.5 + sum_j relu(x_j)."""
    IN, H, oW1, oB1, oW2, oB2, oW3, oB3 = _mlp_offsets()
    p = np.zeros(O.lib().orc_mlp_nparams(), dtype=np.float32)
    for j in range(H):
        p[oW1 + j * IN + j] = 1.0
        p[oW2 + j * H + j] = 1.0
        p[oW3 + j] = 1.0
    p[oB3] = 0.5
    rng = np.random.default_rng(3)
    x = rng.normal(size=(7, IN)).astype(np.float32)
    y = O.mlp_forward(p, x)
    want = 0.5 + np.maximum(x[:, :H].astype(np.float64), 0.0).sum(axis=1)
    assert np.allclose(y, want, rtol=0, atol=1e-12)


def test_mlp_oracle_gradient_matches_finite_differences():
    """The oracle's reverse-mode gradient against central differences of its own
    forward pass (double), on every layer's weights and biases."""
    IN, H, oW1, oB1, oW2, oB2, oW3, oB3 = _mlp_offsets()
    rng = np.random.default_rng(7)
    p = (rng.normal(size=O.lib().orc_mlp_nparams()) * 0.2).astype(np.float32)
    x = rng.normal(size=(9, IN)).astype(np.float32)
    t = rng.normal(size=9).astype(np.float32)
    loss, g = O.mlp_grad(p, x, t)

    def L(q):
        y = O.mlp_forward(q, x)
        return np.mean((y - t.astype(np.float64)) ** 2)

    assert abs(loss - L(p)) < 1e-12 * max(1.0, loss)
    idx = list(rng.choice(oB1, 12, replace=False)) + [oB1 + 3, oW2 + 5 * H + 9, oB2 + 7, oW3 + 11, oB3]
    for i in idx:
        h = np.float32(1e-3)
        qp = p.copy()
        qm = p.copy()
        qp[i] += h
        qm[i] -= h
        fd = (L(qp) - L(qm)) / (np.float64(qp[i]) - np.float64(qm[i]))
        assert abs(fd - g[i]) <= 2e-4 * max(1.0, abs(g[i])), (i, fd, g[i])


def test_tile_probe_is_the_worst_sample_of_its_tile():
    """The probe (the corner farthest from the reference point) stops the
    reference's scan first: every sample of the tile has a first unstable index
    >= the probe's (0 = never), by brute force over the tile."""
    v = json.load(open(os.path.join(GOLD, "views.json")))["t13"]
    o = O.OView(40, 56)
    o.set_view(v["N"], v["re"], v["im"], "3e-24")
    o.precompute()
    ua, rc = o.tile_probe_ua(16)
    for tr in range(ua.shape[0]):
        for tc in range(ua.shape[1]):
            pr, pc = rc[tr, tc]
            u_probe = o.sa_start(int(pr), int(pc))[0]
            assert u_probe == ua[tr, tc]
            for r in range(tr * 16, min(40, tr * 16 + 16), 5):
                for c in range(tc * 16, min(56, tc * 16 + 16), 5):
                    u = o.sa_start(r, c)[0]
                    assert u == 0 or (ua[tr, tc] != 0 and u >= ua[tr, tc])


# ------------------------------------- the series skip, pinned independently
def _orbit_mp(c, n):
    x = [c]
    for _ in range(n):
        x.append(x[-1] * x[-1] + c)
    return x


def test_series_B_C_as_sums_over_the_orbit_at_a_nonzero_reference():
    """computeSeries (mandelbrot.cpp:112-131) written as the solution of its linear
    recurrences: with G(k, n) = prod_{j=k}^{n-1} 2 X_j,
      A_n = sum_{k=0}^{n} G(k, n),  B_n = sum_{k=1}^{n} G(k, n) A_k^2,
      C_n = sum_{k=1}^{n} G(k, n) 2 A_k B_k,
    evaluated from an mpmath orbit at 300 bits; A_n also equals the Taylor
    coefficient of the orbit in c, taken from power-series iteration (truncated
    at degree 3), and B_n, C_n differ from the Taylor b_n, c_n (the reference's
    new-index quirk)."""
    mp.mp.prec = 300
    c = mp.mpc("-0.7391", "0.1837")
    n_max = 6
    o = O.OView(8, 8)
    o.set_view(40, "-0.7391", "0.1837", "1e-20")
    o.precompute_at("-0.7391", "0.1837")
    Xt, Xn, A, B, Cc = o.series()
    X = _orbit_mp(c, n_max)

    def G(k, n):
        p = mp.mpc(1)
        for j in range(k, n):
            p *= 2 * X[j]
        return p

    # Taylor coefficients of Y_n(e) = orbit of c + e: Y_0 = c + e, Y_{n+1} = Y_n^2 + c + e
    poly = [c, mp.mpc(1), mp.mpc(0), mp.mpc(0)]
    taylor = [poly]
    for _ in range(n_max):
        p = taylor[-1]
        sq = [sum(p[i] * p[k - i] for i in range(k + 1)) for k in range(4)]
        taylor.append([sq[0] + c, sq[1] + 1, sq[2], sq[3]])
    Am = [sum(G(k, n) for k in range(n + 1)) for n in range(n_max + 1)]
    Bm = [sum((G(k, n) * Am[k] ** 2 for k in range(1, n + 1)), mp.mpc(0)) for n in range(n_max + 1)]
    Cm = [sum((G(k, n) * 2 * Am[k] * Bm[k] for k in range(1, n + 1)), mp.mpc(0)) for n in range(n_max + 1)]

    def close(arr, n, want, rel=1e-14):
        got = complex(arr[2 * n], arr[2 * n + 1])
        return abs(got - complex(want)) <= rel * max(abs(complex(want)), 1e-300)

    for n in range(1, n_max + 1):
        assert abs(Am[n] - taylor[n][1]) <= mp.mpf(10) ** -60 * abs(Am[n])  # A is the derivative
        assert close(A, n, Am[n]) and close(B, n, Bm[n]) and close(Cc, n, Cm[n]), n
        assert not close(B, n, taylor[n][2], 1e-3) and not close(Cc, n, taylor[n][3], 1e-3), n


def _direct_count(cr, ci, N, bailout=1024.0):
    """z_0 = c, z_{k+1} = z_k^2 + c at the current mpmath precision; the first k
    whose descend()ed |z_k|^2 exceeds bailout^2 (the series path's count)."""
    z = mp.mpc(cr, ci)
    c = z
    for k in range(N):
        zz = complex(z)
        if zz.real * zz.real + zz.imag * zz.imag > bailout * bailout:
            return k
        z = z * z + c
    return N


# c* on the real axis where the orbit's element X[99] (X[0] = c) equals the
# bailout 1024: c slightly below escapes at 100, slightly above at 99 (found by
# bisection of the escape index at 400 bits)
_C_STAR = "0.25100340768234008000104192638925597691522136096723"


def test_escape_inside_the_skip_equals_direct_iteration():
    """A reference just below c* (its orbit keeps X[99] a hair under the bailout
    and escapes at 100): samples on its right cross the bailout at X[99] + d[99]
    with the series still stable at the default tolerance, so the reference's
    binary search (mandelbrot.cpp:186-206) decides their count inside the skip;
    samples on its left continue in HP and escape at 100.  Every count equals
    direct iteration from z = c at 400 bits."""
    mp.mp.prec = 400
    ref = mp.nstr(mp.mpf(_C_STAR) - mp.mpf("1e-24"), 60)
    nr, nc, N, width = 6, 8, 300, "1e-15"
    o = O.OView(nr, nc)
    o.set_view(N, ref, "0", width)
    assert not o.use_hardware()
    assert o.precompute_at(ref, "0") == 100
    rc = np.array([(r, c) for r in range(nr) for c in range(nc)], dtype=np.int32)
    got = o.render_samples(rc)
    sz = mp.mpf(width) / nc
    inside = 0
    for k, (r, c) in enumerate(rc):
        ua, found, d = o.sa_start(int(r), int(c))
        assert ua == 0 and found == 99  # stable over the whole reference orbit
        cr = mp.mpf(ref) + (int(c) - nc // 2) * sz
        ci = (nr // 2 - int(r) - 1) * sz
        assert got["iterations"][k] == _direct_count(cr, ci, N), (r, c)
        inside += int(got["iterations"][k] == 99)
    assert 0 < inside < len(rc)  # both branches ran


class _Trunc:
    """mpf arithmetic as the reference uses it (GMP mpf: every result truncated
    toward zero at the working precision), from mpmath's low-level layer with
    rounding mode 'd' at P = 64 ((bits + 127) // 64 + 1) bits (the precision
    tests/golden/make_hp_goldens.py pins against both HP implementations)."""

    def __init__(self, bits):
        from mpmath import libmp
        self.L = libmp
        self.P = 64 * ((bits + 127) // 64 + 1)

    def s(self, x):
        return self.L.from_str(x, self.P, "d")

    def d(self, x):
        return self.L.from_float(float(x))

    def add(self, a, b):
        return self.L.mpf_add(a, b, self.P, "d")

    def sub(self, a, b):
        return self.L.mpf_sub(a, b, self.P, "d")

    def mul(self, a, b):
        return self.L.mpf_mul(a, b, self.P, "d")

    def mul_i(self, a, k):
        return self.L.mpf_mul(a, self.L.from_int(k), self.P, "d")

    def div_i(self, a, k):
        return self.L.mpf_div(a, self.L.from_int(k), self.P, "d")

    def get_d(self, a):  # mpf_get_d: truncation toward zero
        return self.L.to_float(self.L.mpf_pos(a, 53, "d"))


def test_series_stop_index_and_continuation_pinned_independently():
    """The default tolerance on a view whose samples continue past the skip (t11
    at width 8e-16, N 3000, reference point at the view's centre):
    * the first unstable index equals the first i at which isUnstable
      (mandelbrot.cpp:138-142) holds when evaluated in exact arithmetic on the
      same doubles (the series terms descend()ed, eps and its powers as the
      reference forms them), unless the exact test lies within 1e-12 relative of
      its threshold, and found = max(i - 3, 1) - 1;
    * the count equals the reference's HP continuation (mandelbrot.cpp:209-224)
      redone in mpmath with truncation at the same precision: the pixel
      coordinate, the reference orbit and the continuation from
      Y = X[found] + d[found], each operation truncated as mpf truncates.
    Counts are not compared with iteration from z = c: the reference's series
    (its new-index B, C recurrence) moves the skip point by ~1e-10 relative, and
    chaotic tails turn that into different counts for some samples (printed);
    good = max(i - 3, 1) is the reference's convention, which only the reference
    itself could pin."""
    v = json.load(open(os.path.join(GOLD, "views.json")))["t11"]
    nr, nc, width, N = 4, 6, "8e-16", v["N"]
    o = O.OView(nr, nc)
    o.set_view(N, v["re"], v["im"], width)
    assert not o.use_hardware()  # the series path (sample spacing below 1.5e-16)
    o.precompute_at(v["re"], v["im"])
    L = o.ref_len()
    Xt, Xn, A, B, Cc = o.series()
    rc = np.array([(r, c) for r in range(nr) for c in range(nc)], dtype=np.int32)
    got = o.render_samples(rc)
    T = _Trunc(o.bits())
    cre, cim = T.s(v["re"]), T.s(v["im"])
    sz = T.div_i(T.s(width), nc)
    # the reference orbit as computeOrbit forms it (mandelbrot.cpp:97-110)
    X = [(cre, cim)]
    for i in range(1, L + 1):
        pr, pi = X[-1]
        xr = T.add(T.sub(T.mul(pr, pr), T.mul(pi, pi)), cre)
        xi = T.add(T.mul_i(T.mul(pr, pi), 2), cim)
        X.append((xr, xi))
    assert all(T.get_d(X[i][0]) == Xt[2 * i] for i in range(0, L, 97))
    tol = mp.mpf(1e-10)
    mp.mp.prec = 300
    differ_from_z0 = checked = 0
    for k, (r, c) in enumerate(rc):
        ua, found, d = o.sa_start(int(r), int(c))
        e = o.pixel_eps(int(r), int(c))
        e0, e1 = float(e[0]), float(e[1])
        # eps2 = sq(eps), eps3 = eps * eps2 as the reference rounds them (complex.h)
        e2 = (e0 * e0 - e1 * e1, (2.0 * e0) * e1)
        e3 = (e0 * e2[0] - e1 * e2[1], e0 * e2[1] + e1 * e2[0])
        first = 0
        for i in range(1, L):
            b = mp.mpc(B[2 * i], B[2 * i + 1]) * mp.mpc(*e2)
            cc = mp.mpc(Cc[2 * i], Cc[2 * i + 1]) * mp.mpc(*e3)
            if abs(b) ** 2 * tol < abs(cc) ** 2:
                first = i
                break
        if first != ua:
            i = min(first or L, ua or L)
            b = mp.mpc(B[2 * i], B[2 * i + 1]) * mp.mpc(*e2)
            cc = mp.mpc(Cc[2 * i], Cc[2 * i + 1]) * mp.mpc(*e3)
            assert abs(abs(b) ** 2 * tol / abs(cc) ** 2 - 1) < 1e-12, (r, c, ua, first)
        assert found == (max(ua - 3, 1) if ua else L) - 1
        # the pixel (mandelbrot.cpp:269-283) and the HP continuation, truncated
        y0r = T.add(cre, T.mul_i(sz, int(c) - nc // 2))
        y0i = T.add(cim, T.mul_i(sz, nr // 2 - int(r) - 1))
        yr, yi = T.add(X[found][0], T.d(d[0])), T.add(X[found][1], T.d(d[1]))
        dr, di = T.get_d(yr), T.get_d(yi)
        if dr * dr + di * di > 1024.0 ** 2:
            continue  # escaped inside the skip: the binary search (pinned above)
        want = N
        for i in range(found + 1, N):
            nr_ = T.add(T.sub(T.mul(yr, yr), T.mul(yi, yi)), y0r)
            ni_ = T.add(T.mul_i(T.mul(yr, yi), 2), y0i)
            yr, yi = nr_, ni_
            dr, di = T.get_d(yr), T.get_d(yi)
            if dr * dr + di * di > 1024.0 ** 2:
                want = i
                break
        assert got["iterations"][k] == want, (r, c, found)
        checked += 1
        cr = mp.mpf(v["re"]) + (int(c) - nc // 2) * (mp.mpf(width) / nc)
        ci = mp.mpf(v["im"]) + (nr // 2 - int(r) - 1) * (mp.mpf(width) / nc)
        with mp.workprec(400):
            differ_from_z0 += int(got["iterations"][k] != _direct_count(cr, ci, N))
    assert checked == len(rc)
    print("t11 at 8e-16: %d of %d counts differ from 400-bit iteration from z = c" % (differ_from_z0, len(rc)))
