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
