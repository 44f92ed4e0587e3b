"""The learned series skip on the GPU: network kernels against the CPU oracle,
per-tile skips against the oracle, and what the trained network gains.

A tile skip starts every sample of a tile at the series skip of the tile's probe
(its sample farthest from the reference point).  The probe's first unstable index
comes from the reference's scan, the threshold tables or the network; a network
prediction is used only when the reference's stability test confirms it exactly,
so all three give the same tiles and the same grids.
"""
import json
import os

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu

_VIEWS = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "views.json")))


def _view(nm, width, nr=96, nc=128, name="t13"):
    v = _VIEWS[name]
    m = nm.Mandelbrot(nr, nc)
    m.set_view(v["N"], v["re"], v["im"], width)
    m.precompute()
    o = O.OView(nr, nc)
    o.set_view(v["N"], v["re"], v["im"], width)
    o.precompute()
    return m, o


def test_skipnet_forward_and_gradient_match_cpu_oracle(gpu):
    """GPU forward outputs and batch gradients (FP32) against the oracle's double
    reverse-mode differentiation: relative 1e-4 on the loss and the outputs, and
    1e-3 of the largest gradient entry on every parameter."""
    nm = gpu
    rng = np.random.default_rng(11)
    net = nm.SkipNet()
    p = (rng.normal(size=nm.SKIPNET_PARAMS) * 0.15).astype(np.float32)
    net.set(p, np.zeros(nm.SKIPNET_INPUTS, np.float32), np.ones(nm.SKIPNET_INPUTS, np.float32))
    for n in (1, 37, 256):
        x = rng.normal(size=(n, nm.SKIPNET_INPUTS)).astype(np.float32)
        t = rng.normal(size=n).astype(np.float32)
        y = net.predict(x)
        y_ref = O.mlp_forward(p, x)
        assert np.allclose(y, y_ref, rtol=1e-4, atol=1e-4 * np.abs(y_ref).max())
        loss, g = net.gradient(x, t)
        loss_ref, g_ref = O.mlp_grad(p, x, t)
        assert abs(loss - loss_ref) <= 1e-4 * loss_ref
        assert np.abs(g - g_ref).max() <= 1e-3 * np.abs(g_ref).max(), np.abs(g - g_ref).max()


def test_skipnet_training_fits_a_smooth_target(gpu):
    """Adam on the GPU drives the loss of a smooth target far below its variance."""
    nm = gpu
    rng = np.random.default_rng(5)
    x = rng.normal(size=(4096, nm.SKIPNET_INPUTS)).astype(np.float32)
    t = (0.5 + 0.3 * np.tanh(x[:, 0] - 0.5 * x[:, 1]) + 0.1 * x[:, 2] ** 2).astype(np.float32)
    net = nm.SkipNet()
    curve = net.train(x, t, epochs=60, batch=256, lr=3e-3, seed=2)
    y = net.predict(x)
    mse = float(np.mean((y - t) ** 2))
    assert curve[-1] < 0.1 * curve[0]
    assert mse < 0.05 * float(np.var(t)), (mse, np.var(t))


@pytest.mark.parametrize("width,T", [("3e-24", 16), ("1e-22", 32)])
def test_tile_skips_scan_and_tables_equal_the_oracle(gpu, width, T):
    nm = gpu
    m, o = _view(nm, width)
    want, _ = o.tile_probe_ua(T)
    scan, how0, st0 = m.tile_skips(T, mode=0)
    table, how1, st1 = m.tile_skips(T, mode=1)
    assert np.array_equal(scan, want.ravel())
    assert np.array_equal(table, want.ravel())
    assert st0["ntiles"] == want.size and st0["evals"] >= st1["evals"]


@pytest.mark.parametrize("mode", [1, 2])
def test_tile_skip_render_bit_exact_vs_oracle(gpu, mode):
    """Every sample starts at its tile's skip: identical to the oracle's method
    restatement with that skip, sample for sample."""
    nm = gpu
    nr, nc, T = 96, 128, 16
    m, o = _view(nm, "3e-24", nr, nc)
    ua, _ = o.tile_probe_ua(T)
    m.set_skip_mode(mode, T)
    m.render()
    g = m.grid()
    st = m.tile_stats()
    assert st["ntiles"] == ua.size
    rc = np.array([(r, c) for r in range(nr) for c in range(nc)], dtype=np.int32)
    per = np.array([ua[r // T, c // T] for r, c in rc], dtype=np.int32)
    ref = o.pert_samples_ua(rc, per).reshape(nr, nc)
    assert np.array_equal(g["iterations"], ref["iterations"])
    assert np.array_equal(g["smoothing"].view(np.uint32), ref["smoothing"].view(np.uint32))


def _brute_count(v, width, nr, nc, r, c):
    """Escape count of sample (r, c) by plain HP iteration from z = c (no series),
    the reference's count convention (mandelbrot.cpp:208-228): mpmath, 300 bits."""
    import mpmath
    with mpmath.workprec(300):
        sz = mpmath.mpf(width) / nc
        cr = mpmath.mpf(v["re"]) + (c - nc // 2) * sz
        ci = mpmath.mpf(v["im"]) + (nr // 2 - r - 1) * sz
        zr, zi = cr, ci
        for k in range(v["N"]):
            zr, zi = zr * zr - zi * zi + cr, 2 * zr * zi + ci
            if zr * zr + zi * zi > 1024 ** 2:
                return k + 1
    return v["N"]


def test_tile_skip_render_vs_reference_hp(gpu):
    """A tile starts at its probe's skip, never later than a sample's own, so its
    series error is smaller than the reference's per-sample skip.  Counts agree
    with the reference method (SA scan + HP continuation) on >= 99.9% of samples,
    within 2 elsewhere, and where they differ plain HP iteration from z = c (no
    series at all) sides with the tile skip on >= 80% of them: the differences are
    the reference's own series error."""
    nm = gpu
    nr, nc, width = 96, 128, "1e-22"
    m, o = _view(nm, width)
    m.set_skip_mode(nm.SKIP_TILE_TABLE, 32)
    m.render()
    g = m.grid()
    ref = o.render_rows(np.arange(nr))
    same = g["iterations"] == ref["iterations"]
    assert same.mean() >= 0.999
    assert np.abs(g["iterations"].astype(np.int64) - ref["iterations"]).max() <= 2
    diff = np.argwhere(~same)
    if len(diff):
        v = _VIEWS["t13"]
        tile_right = sum(_brute_count(v, width, nr, nc, int(r), int(c)) == g["iterations"][r, c] for r, c in diff)
        assert tile_right >= 0.8 * len(diff), (tile_right, len(diff))


def _training_set(nm, widths, T, nr=192, nc=256):
    xs, ts, views = [], [], []
    for w in widths:
        m, _ = _view(nm, w, nr, nc)
        ua, _, _ = m.tile_skips(T, mode=0)
        xs.append(m.tile_features(T))
        ts.append(m.tile_targets(ua))
        views.append((m, ua))
    return np.concatenate(xs), np.concatenate(ts), views


def test_learned_skip_gain_over_the_plain_probe_search(gpu):
    """Trained on the probe searches of 16 views of one zoom (widths 1e-25 ..
    3e-22), the network predicts the tiles of 4 held-out widths.  A prediction
    that the check accepts (every index below it certainly stable) starts the scan
    there, one that it rejects falls back to the scan from 1, so the tiles equal
    the scan's; the gain is the exact stability tests the probes no longer
    evaluate."""
    nm = gpu
    T = 16
    train_w = ["%.3e" % (1e-25 * 10 ** (k * 3.5 / 15)) for k in range(16)]
    test_w = ["2e-25", "7e-25", "4e-24", "9e-23"]
    x, t, _ = _training_set(nm, train_w, T)
    net = nm.SkipNet()
    curve = net.train(x, t, epochs=150, batch=128, lr=2e-3, seed=1)
    assert curve[-1] < curve[0]
    scan_evals = learned_evals = accepted = total = 0
    for w in test_w:
        m, _ = _view(nm, w, 192, 256)
        scan, _, s0 = m.tile_skips(T, mode=0)
        learned, how, s2 = m.tile_skips(T, mode=2, net=net)
        assert np.array_equal(learned, scan)  # validation makes the learned path exact
        assert set(np.unique(how)) <= {1, 2, 3, 4}
        assert s2["accepted"] + s2["rejected"] + s2["bounded"] + int((how == 1).sum()) == s2["ntiles"]
        scan_evals += s0["evals"]
        learned_evals += s2["evals"]
        accepted += s2["accepted"] + s2["bounded"]
        total += s2["ntiles"]
    print("learned skip: %d/%d predictions accepted; exact stability tests %d (scan) -> %d (learned)"
          % (accepted, total, scan_evals, learned_evals))
    assert accepted > 0
    assert learned_evals < scan_evals
