"""Parity of the HIP kernels (through the C ABI) with the CPU oracle.

Bars: integer escape counts bit-exact where the method is the same arithmetic
(plain-double path; series start + FP64 perturbation vs the oracle's restatement
of the same method); >= 99.99% identical counts against the reference method
with HP continuation; smoothing within 1e-6 relative (absolute floor 1e-6 for
values near 0), stored as float32 like grid.h.
"""
import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu

SMOOTH_RTOL = 1e-6


def _smooth_ok(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return np.abs(a - b) <= SMOOTH_RTOL * np.maximum(np.abs(b), 1.0)


def _pair(nr, nc, N, cre, cim, width, bailout=None):
    import newman_amd as nm
    m = nm.Mandelbrot(nr, nc)
    m.set_view(N, cre, cim, width)
    o = O.OView(nr, nc)
    o.set_view(N, cre, cim, width)
    if bailout is not None:
        m.set_bailout(bailout)
        o.set_bailout(bailout)
    return m, o


def test_plain_path_config1_full_frame(gpu):
    """The plain-path view: 640x480, centre -0.75+0i, width 3.0, N=1024, bailout 1024: the
    reference's plain-double path (mandelbrot.cpp:231-254), every sample."""
    m, o = _pair(480, 640, 1024, "-0.75", "0", "3.0")
    assert m.useHardware() and o.use_hardware()
    st = m.render()
    g = m.grid()
    ref = o.render_rows(np.arange(480))
    assert np.array_equal(g["iterations"], ref["iterations"])
    ok = _smooth_ok(g["smoothing"], ref["smoothing"])
    assert ok.all(), np.argwhere(~ok)[:5]
    # cardioid/bulb samples counted apart from credited work
    assert st["n_cardioid"] > 0.15 * 480 * 640
    esc = ref["iterations"].astype(np.int64)
    assert st["credited"] == esc.sum() - st["n_cardioid"] * 1024


# (radius, escape test once per group of four updates): the frame's |c| reaches
# |-2.25 + 1.125i| = 2.52, so the grouped test needs 0.999 (R^2 - R) >= 2.52 and R <= 2^30
@pytest.mark.parametrize("bailout,grouped", [(2.0, False), (1.5, False), (3.0, True), (1e6, True),
                                             (1e9, True), (2.0 ** 31, False)])
def test_plain_path_bailout_2(gpu, bailout, grouped):
    """Other bailout radii (the configuration names radius 2): the smoothing's
    series logarithms for radii >= 2, the library's below 2; the escape test per
    update and per group of four updates."""
    m, o = _pair(120, 160, 1024, "-0.75", "0", "3.0", bailout=bailout)
    st = m.render()
    assert st["grouped"] == int(grouped)
    g = m.grid()
    ref = o.render_rows(np.arange(120))
    assert np.array_equal(g["iterations"], ref["iterations"])
    assert _smooth_ok(g["smoothing"], ref["smoothing"]).all()


@pytest.mark.parametrize("bailout,grouped", [(1024.0, True), (2.0, False)])
def test_plain_path_unfolded_update(gpu, bailout, grouped):
    """A frame whose row coordinates include 1e-70 (nonzero and below 2^-200)
    runs the three-operation update y' = (2x) y + c (the fold into one fma is
    exact only for coordinates 0 or >= 2^-200), with and without the grouped test."""
    m, o = _pair(61, 80, 700, "-0.75", "1e-70", "3.0", bailout=bailout)
    assert m.useHardware()
    st = m.render()
    assert st["fma2"] == 0 and st["grouped"] == int(grouped)
    g = m.grid()
    ref = o.render_rows(np.arange(61))
    assert np.array_equal(g["iterations"], ref["iterations"])
    assert _smooth_ok(g["smoothing"], ref["smoothing"]).all()


def test_plain_path_symmetry(gpu):
    """centre.im = 0: rows r and nr-2-r are bitwise mirror images."""
    import newman_amd as nm
    m = nm.Mandelbrot(200, 300)
    m.set_view(700, "-0.6", "0", "2.8")
    m.render()
    g = m.grid()
    for r in range(0, 199):
        assert np.array_equal(g[r], g[198 - r])


# points on the cardioid boundary c(t) = e^{it}/2 - e^{2it}/4 and the period-2
# disk |c + 1| = 1/4, at widths where many samples sit within the double filter's
# error bound of the boundary (the device then decides in double-double)
_BOUNDARY = [("0.36552730612364814", "0.29715374784582538", "1e-3"),  # t = 1.2
             ("-0.044662513057668218", "0.64384933723982285", "1e-9"),  # t = 2.0
             ("-1.0", "0.25", "1e-13"), ("-1.1767766952966369", "0.1767766952966369", "1e-11"),
             # the cardioid's rightmost (t = pi/3) and topmost (t = 2 pi / 3) points,
             # nearest the device's bounding box of the two components
             ("0.375", "0.21650635094610965", "1e-6"), ("-0.125", "0.649519052838329", "1e-6"),
             # wide frames across the edges of that box (x = 0.42, |y| = 0.7, x = -1.3)
             ("0.40", "0.2", "0.1"), ("-0.125", "0.68", "0.1"), ("-1.27", "0.0", "0.1")]


@pytest.mark.parametrize("cre,cim,width", _BOUNDARY)
def test_plain_path_cardioid_boundary_frames(gpu, cre, cim, width):
    """Frames straddling the cardioid / period-2 disk boundary: the device's
    double filter plus double-double test equals the HP test (mandelbrot.cpp:63-71)
    on every sample."""
    m, o = _pair(64, 96, 400, cre, cim, width)
    assert m.useHardware()
    st = m.render()
    g = m.grid()
    ref = o.render_rows(np.arange(64))
    card = np.array([[o.in_cardioid(r, c) for c in range(96)] for r in range(64)])
    assert 0 < card.sum() < card.size  # the boundary crosses the frame
    assert st["n_cardioid"] == card.sum()
    assert np.array_equal(g["iterations"], ref["iterations"])
    assert _smooth_ok(g["smoothing"], ref["smoothing"]).all()


@pytest.mark.parametrize("pre,K", [(1, 16), (16, 16), (200, 16), (1, 8), (1, 64), (16, 64)])
def test_plain_path_pre_iteration_depth(gpu, pre, K):
    """The start kernel's pre-iteration depth and the iterate kernel's block size
    only move work between kernels and blocks."""
    m, o = _pair(120, 160, 700, "-0.75", "0.05", "2.9")
    m.set_launch(pre_iter=pre, block_k=K)
    m.render()
    ref = o.render_rows(np.arange(120))
    g = m.grid()
    assert np.array_equal(g["iterations"], ref["iterations"])
    assert _smooth_ok(g["smoothing"], ref["smoothing"]).all()


def test_plain_path_compute_row_dropin(gpu):
    """computeRow(r) for ascending rows (viewer.cpp:166-171) then at(r, c)."""
    import newman_amd as nm
    m = nm.Mandelbrot(60, 80)
    m.N = 300
    for r in range(60):
        m.computeRow(r)
    o = O.OView(60, 80)
    o.set_N(300)
    ref = o.render_rows(np.arange(60))
    for r in (0, 17, 59):
        for c in (0, 40, 79):
            it, sm = m.at(r, c)
            assert it == ref["iterations"][r, c]


import json
import os

_VIEWS = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "views.json")))
# (view, rows, cols, width): series path (sample spacing < 1.5e-16)
DEEP_CASES = [("t11", 96, 128, "1e-16"), ("t11", 96, 128, "1e-14"), ("t13", 96, 128, "1e-25"),
              ("t13", 96, 128, "1e-22"), ("t13", 45, 61, "3e-24"),
              # series terms under/overflow in the exact test: the deep start of the stop search
              ("c5_seed1", 36, 48, "1e-100"), ("c5_seed1", 36, 48, "3e-97"), ("c4_seed1", 36, 48, "1e-40")]


def _deep(name, nr, nc, width):
    v = _VIEWS[name]
    return _pair(nr, nc, v["N"], v["re"], v["im"], width)


@pytest.mark.parametrize("name,nr,nc,width", DEEP_CASES)
def test_series_path_bit_exact_vs_method_oracle(gpu, name, nr, nc, width):
    """Series start + FP64 perturbation: identical to the oracle's restatement of
    the same arithmetic on every sample (counts and smoothing bits)."""
    m, o = _deep(name, nr, nc, width)
    assert not m.useHardware()
    m.precompute()
    o.precompute()
    assert m.reference_length() == o.ref_len() and m.probe_winner() == o.winner()
    st = m.render()
    g = m.grid()
    rc = np.array([(r, c) for r in range(nr) for c in range(nc)], dtype=np.int32)
    ref = o.render_samples(rc, pert=True).reshape(nr, nc)
    assert np.array_equal(g["iterations"], ref["iterations"]), np.argwhere(g["iterations"] != ref["iterations"])[:5]
    assert _smooth_ok(g["smoothing"], ref["smoothing"]).all()
    assert st["n_iterated"] + st["n_sa_resolved"] + st["n_cardioid"] == nr * nc


@pytest.mark.parametrize("K", [12, 16, 32, 64])
@pytest.mark.parametrize("name,nr,nc,width", [DEEP_CASES[4], DEEP_CASES[7]])
def test_series_block_size_keeps_results(gpu, name, nr, nc, width, K):
    """A block only groups K steps of the same per-step arithmetic (escape record,
    rebase, forced rebase at the orbit's end), so every block size gives the
    oracle's grid bit for bit."""
    m, o = _deep(name, nr, nc, width)
    m.precompute()
    o.precompute()
    st = m.render(block_k=K)
    assert st["block_k"] == K
    g = m.grid()
    rc = np.array([(r, c) for r in range(nr) for c in range(nc)], dtype=np.int32)
    ref = o.render_samples(rc, pert=True).reshape(nr, nc)
    assert np.array_equal(g["iterations"], ref["iterations"])
    assert _smooth_ok(g["smoothing"], ref["smoothing"]).all()


@pytest.mark.parametrize("run", [0, 1, 4, 16, 64])
def test_series_chunk_interleaving_keeps_results(gpu, run):
    """The series kernel's chunk interleaving only changes which wave takes a
    sample, so every run length gives the oracle's grid bit for bit; the render
    reports the order it used (0: the default, runs of 4 or 64 samples)."""
    m, o = _deep(*DEEP_CASES[4])
    m.precompute()
    o.precompute()
    m.set_launch(chunk_run=run)
    st = m.render()
    assert m.get_launch()["chunk_run"] == run
    assert st["chunk_run"] == run if run else st["chunk_run"] in (4, 64)
    nr, nc = DEEP_CASES[4][1], DEEP_CASES[4][2]
    g = m.grid()
    rc = np.array([(r, c) for r in range(nr) for c in range(nc)], dtype=np.int32)
    ref = o.render_samples(rc, pert=True).reshape(nr, nc)
    assert np.array_equal(g["iterations"], ref["iterations"])
    assert _smooth_ok(g["smoothing"], ref["smoothing"]).all()


@pytest.mark.parametrize("name,nr,nc,width", DEEP_CASES[:4])
def test_series_path_vs_reference_hp(gpu, name, nr, nc, width):
    """Against the reference method (SA scan + HP continuation, mandelbrot.cpp:144-229):
    counts identical on >= 99.99% of samples (the rest within one iteration);
    smoothing within 1e-6 relative on >= 99.9% of equal-count samples and within
    1e-4 on all (FP64 deltas drift from a 192-bit continuation only in long
    chaotic tails after a sample leaves the reference)."""
    m, o = _deep(name, nr, nc, width)
    m.precompute()
    o.precompute()
    m.render()
    g = m.grid()
    ref = o.render_rows(np.arange(nr))
    same = g["iterations"] == ref["iterations"]
    assert same.mean() >= 0.9999 or (~same).sum() <= 1, (same.mean(), np.argwhere(~same)[:5])
    assert np.abs(g["iterations"].astype(np.int64) - ref["iterations"]).max() <= 1
    ok6 = _smooth_ok(g["smoothing"], ref["smoothing"])[same]
    assert ok6.mean() >= 0.999
    d = np.abs(g["smoothing"].astype(np.float64) - ref["smoothing"])[same]
    assert d.max() <= 1e-4


def test_downsample_matches_at_avg(gpu):
    """scaleDown on the device == at(r, c, sc) of mandelbrot.cpp:320-332."""
    import newman_amd as nm
    m = nm.Mandelbrot(120, 160)
    m.N = 400
    m.render()
    fine = m.grid().copy()
    want = O.downsample(fine, 4)
    for (r, c) in [(0, 0), (5, 7), (29, 39)]:
        it, sm = m.at(r, c, 4)
        assert (it, np.float32(sm)) == (want["iterations"][r, c], want["smoothing"][r, c])
    m.scaleDown(4)
    got = m.grid()
    assert np.array_equal(got["iterations"], want["iterations"])
    assert np.array_equal(got["smoothing"], want["smoothing"])


@pytest.mark.parametrize("N,width", [(12, "1e-24"), (40, "1e-24"), (300, "3e-23")])
def test_series_path_short_orbits(gpu, N, width):
    """Reference orbits shorter than a block (len(X) <= N < K): every block is an
    end-zone block and rebased lanes can reach the orbit's end again."""
    v = _VIEWS["t13"]
    m, o = _pair(40, 56, N, v["re"], v["im"], width)
    m.precompute()
    o.precompute()
    m.render()
    rc = np.array([(r, c) for r in range(40) for c in range(56)], dtype=np.int32)
    ref = o.render_samples(rc, pert=True).reshape(40, 56)
    g = m.grid()
    assert np.array_equal(g["iterations"], ref["iterations"])
    assert _smooth_ok(g["smoothing"], ref["smoothing"]).all()


@pytest.mark.parametrize("rre,rim", [("0.5", "0.5"), ("-1.8", "0.02"), ("-0.8", "0.2")])
def test_series_path_reference_escapes_early(gpu, rre, rim):
    """An explicit reference point whose orbit escapes after a few steps: samples
    outlive it and rebase at its end (the forced rebase of the iteration)."""
    v = _VIEWS["t13"]
    m, o = _pair(48, 64, 3000, v["re"], v["im"], "1e-23")
    m.precompute_at(rre, rim)
    o.precompute_at(rre, rim)
    assert m.reference_length() == o.ref_len() < 50
    st = m.render()
    rc = np.array([(r, c) for r in range(48) for c in range(64)], dtype=np.int32)
    ref = o.render_samples(rc, pert=True).reshape(48, 64)
    g = m.grid()
    assert np.array_equal(g["iterations"], ref["iterations"])
    assert _smooth_ok(g["smoothing"], ref["smoothing"]).all()
    assert st["n_rebased"] > 0


def test_frame_batch_equals_single_renders(gpu):
    """A zoom-video frame set (shared reference, widths 1.1x apart) in one launch
    equals the frames rendered one by one."""
    import newman_amd as nm
    v = _VIEWS["t13"]
    widths = ["%.6e" % (1e-24 * 1.1 ** k) for k in range(6)]
    m = nm.Mandelbrot(36, 48)
    m.set_view(v["N"], v["re"], v["im"], widths[0])
    m.precompute()
    batch, st = m.render_frames(widths)
    for k, w in enumerate(widths):
        s = nm.Mandelbrot(36, 48)
        s.set_view(v["N"], v["re"], v["im"], widths[0])
        s.precompute()
        one, _ = s.render_frames([w])
        assert np.array_equal(batch[k]["iterations"], one[0]["iterations"])
        assert np.array_equal(batch[k]["smoothing"], one[0]["smoothing"])


def test_plain_batch_over_two_tile_scan_groups(gpu):
    """300 frames of 240 x 240 = 17,280,000 samples = 16,875 start tiles: the tile
    scan splits its counts into two workgroups' ranges (16,384 and 491 counts),
    the second starting at the first's total; every frame equals its single-frame
    render (one range)."""
    import newman_amd as nm
    m = nm.Mandelbrot(240, 240)
    ws = ["3.0", "0.75", "1.5"]
    m.set_view(300, "-0.75", "0.1", "3.0")
    batch, st = m.render_frames([ws[k % 3] for k in range(300)])
    one = []
    for w in ws:
        m.set_view(300, "-0.75", "0.1", w)
        m.render()
        one.append(m.grid().copy())
    for k in range(300):
        assert np.array_equal(batch[k], one[k % 3]), k


# 48 x 64 = 3 start tiles of 1024 samples per frame; 37 x 29 = 1073 samples, so
# start tiles straddle frames (the per-sample frame lookup) and the widths differ
@pytest.mark.parametrize("nr,nc", [(48, 64), (37, 29)])
def test_plain_batch_equals_single_frame(gpu, nr, nc):
    import newman_amd as nm
    m = nm.Mandelbrot(nr, nc)
    widths = ["3.0", "2.5", "3.0", "1.25", "3.0"]
    m.set_view(500, "-0.75", "0", "3.0")
    batch, _ = m.render_frames(widths)
    for k, w in enumerate(widths):
        m.set_view(500, "-0.75", "0", w)
        m.render()
        assert np.array_equal(batch[k], m.grid()), (k, w)
