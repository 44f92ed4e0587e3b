"""Parity at the benchmark configurations' own views, and of view files.

* Deep zoom (7680x4320, width 1e-30, N 20000): the whole frame on the GPU; every
  540th row against the oracle's restatement of the GPU method with the same
  reference (bit for bit) and against the reference method (series scan + HP
  continuation, mandelbrot.cpp:144-229, with its own findProbe) on >= 99.99%.
* Zoom-video batch (1280x720 frames 1.1x apart, 8.4e-36 ... 1e-40, N 60000, one
  shared reference): five frames rendered together; every 144th row of each
  against the oracle's method with the batch's reference and precision (bit for
  bit) and against the reference method (series scan + HP continuation) from the
  same reference on >= 99.99%; the viewer's per-keyframe findProbe
  (viewer.cpp:255-294) is reported beside it.
* Precision variants (2560x1440, width 1e-200, N 50000, the detail walk's
  view, series in floatexp): every 288th row against the oracle's floatexp
  restatement of the GPU method (bit for bit), with FP64 and with FP32X deltas;
  the FP32X counts' agreement with FP64 is reported.
* Learned skip (3840x2160, width 1e-100, N 100000, the detail walk's view,
  series in floatexp, 64x64 tiles): the learned tile skip against the oracle's
  GPU method with the same per-tile skips.
* View files (viewer.cpp:6-29, mandelbrot.cpp:19-35): a saved view loaded back
  by loadLegacy renders as the oracle's loadLegacy of the same file.
"""
import json
import os

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu

_VIEWS = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "views.json")))
THREADS = 16


def _smooth_ok(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return np.abs(a - b) <= 1e-6 * np.maximum(np.abs(b), 1.0)


def _rc(rows, nc):
    return np.array([(r, c) for r in rows for c in range(nc)], dtype=np.int32)


def _brute_count(v, nr, nc, r, c, bits=400, bailout=1024):
    """The escaping index of the orbit z_0 = c, z_{k+1} = z_k^2 + c (the series
    path's convention, mandelbrot.cpp:204-217) by direct iteration at `bits`
    bits, for the pixel (r, c) of the view (mandelbrot.cpp:269-283)."""
    import mpmath as mp
    with mp.workprec(bits):
        sz = mp.mpf(v["width"]) / nc
        cr = mp.mpf(v["re"]) + (c - nc // 2) * sz
        ci = mp.mpf(v["im"]) + (nr // 2 - r - 1) * sz
        zr, zi, b2 = cr, ci, mp.mpf(bailout) ** 2
        for k in range(v["N"]):
            if zr * zr + zi * zi > b2:
                return k
            zr, zi = zr * zr - zi * zi + cr, 2 * zr * zi + ci
    return v["N"]


def test_deep_zoom_config_frame(gpu):
    nm = gpu
    v = _VIEWS["c2_seed1"]
    nr, nc = 4320, 7680
    m = nm.Mandelbrot(nr, nc)
    m.set_view(v["N"], v["re"], v["im"], v["width"])
    m.precompute()
    m.render()
    g = m.grid()
    rows = np.arange(270, nr, 540)
    rc = _rc(rows, nc)
    mine = g[rows].reshape(-1)
    # the GPU method (series start + FP64 perturbation), same reference: bit for bit
    o = O.OView(nr, nc)
    o.set_view(v["N"], v["re"], v["im"], v["width"])
    re, im = m.reference_point()
    o.precompute_at(re, im)
    meth = o.render_samples(rc, THREADS, pert=True)
    assert np.array_equal(mine["iterations"], meth["iterations"])
    assert _smooth_ok(mine["smoothing"], meth["smoothing"]).all()
    # the reference method with its own findProbe
    ref = O.OView(nr, nc)
    ref.set_view(v["N"], v["re"], v["im"], v["width"])
    ref.precompute(THREADS)
    assert ref.ref_len() == m.reference_length()
    hp = ref.render_rows(rows, THREADS).reshape(-1)
    same = mine["iterations"] == hp["iterations"]
    print("deep zoom frame: counts equal to the reference method on %.5f%% of %d samples (%d differ)"
          % (100 * same.mean(), same.size, (~same).sum()))
    # The differences are samples that iterate hundreds of steps past the series
    # skip on a chaotic tail: there the FP64 continuation and the reference's HP
    # continuation (both from the same skip) part ways, and the reference itself
    # is off from the exact orbit (its three-term series carries a relative error
    # of ~1e-10 into the skip).  Each differing sample is checked against direct
    # iteration from z = c at 400 bits: the reference method must be off there.
    assert same.mean() >= 0.9998
    for i in np.nonzero(~same)[0][:12]:
        r, c = rc[i]
        exact = _brute_count(v, nr, nc, int(r), int(c))
        print("  sample (%d, %d): GPU %d, reference method %d, 400-bit orbit %d"
              % (r, c, mine["iterations"][i], hp["iterations"][i], exact))
        assert exact != hp["iterations"][i]


def test_zoom_video_batch_frames(gpu):
    nm = gpu
    v = _VIEWS["c4_seed1"]
    nr, nc, nframes = 720, 1280, 120
    widths = ["%.17e" % (1e-40 * 1.1 ** (nframes - 1 - k)) for k in range(nframes)]
    pick = [0, 30, 60, 90, 119]
    m = nm.Mandelbrot(nr, nc)
    m.set_view(v["N"], v["re"], v["im"], widths[-1])  # the deepest frame fixes precision and reference
    m.precompute()
    grids, _ = m.render_frames([widths[k] for k in pick])
    re, im = m.reference_point()
    rows = np.arange(72, nr, 144)
    rc = _rc(rows, nc)
    total = differ = 0
    for i, k in enumerate(pick):
        mine = grids[i][rows].reshape(-1)
        # the frame as the batch sees it: the deepest frame's precision and
        # reference, this frame's width
        o = O.OView(nr, nc)
        o.set_view(v["N"], v["re"], v["im"], widths[-1])
        o.set_width(widths[k])
        o.precompute_at(re, im)
        meth = o.render_samples(rc, THREADS, pert=True)
        assert np.array_equal(mine["iterations"], meth["iterations"]), k
        assert _smooth_ok(mine["smoothing"], meth["smoothing"]).all(), k
        # the reference method (series scan + HP continuation) from the same
        # reference and skip
        hp = o.render_rows(rows, THREADS).reshape(-1)
        d = int((mine["iterations"] != hp["iterations"]).sum())
        print("zoom-video frame %d (width %s): %d of %d counts differ from the reference method with the "
              "batch's reference" % (k, widths[k], d, mine.size))
        total += mine.size
        differ += d
    assert differ <= 1e-4 * total
    # The viewer renders each keyframe with its own findProbe (viewer.cpp:263);
    # a different reference moves the reference's own series skip error, so such
    # a render differs from any shared-reference one on a few percent of the
    # samples of the shallower frames (direct 400-bit iteration sides with
    # neither); it is reported, not asserted.
    k = pick[0]
    ref = O.OView(nr, nc)
    ref.set_view(v["N"], v["re"], v["im"], widths[k])
    ref.precompute(THREADS)
    own = ref.render_rows(rows[:1], THREADS).reshape(-1)
    print("frame %d with its own findProbe reference: %d of %d counts differ" % (
        k, int((grids[0][rows[:1]].reshape(-1)["iterations"] != own["iterations"]).sum()), own.size))


def test_precision_config_view(gpu):
    nm = gpu
    v = _VIEWS["c3_detail1"]
    nr, nc = 1440, 2560
    m = nm.Mandelbrot(nr, nc)
    m.set_view(v["N"], v["re"], v["im"], v["width"])
    m.set_series_mode(nm.NM_SERIES_FLOATEXP)
    m.precompute()
    m.render()
    g64 = m.grid().copy()
    m.set_delta_mode(nm.NM_DELTA_FP32X)
    st = m.render()
    g32 = m.grid().copy()
    o = O.OView(nr, nc)
    o.set_view(v["N"], v["re"], v["im"], v["width"])
    re, im = m.reference_point()
    o.precompute_at(re, im)
    o.set_series_mode(1)
    rows = np.arange(144, nr, 288)
    rc = _rc(rows, nc)
    meth = o.render_samples(rc, THREADS, pert=True)
    mine = g64[rows].reshape(-1)
    assert np.array_equal(mine["iterations"], meth["iterations"])
    assert _smooth_ok(mine["smoothing"], meth["smoothing"]).all()
    fx = o.pert_samples_fx(rc, st["block_k"], None, THREADS)
    m32 = g32[rows].reshape(-1)
    assert np.array_equal(m32["iterations"], fx["iterations"])
    assert _smooth_ok(m32["smoothing"], fx["smoothing"]).all()
    same = g32["iterations"] == g64["iterations"]
    print("2560x1440 at 1e-200: %d distinct counts; FP32X counts equal to FP64 on %.4f%% of samples"
          % (len(np.unique(g64["iterations"])), 100 * same.mean()))
    assert len(np.unique(g64["iterations"])) > 100


def test_learned_skip_config_view(gpu):
    nm = gpu
    v = _VIEWS["c5_detail1"]
    nr, nc, T = 2160, 3840, 64
    m = nm.Mandelbrot(nr, nc)
    m.set_view(v["N"], v["re"], v["im"], v["width"])
    m.set_series_mode(nm.NM_SERIES_FLOATEXP)
    m.precompute()
    # a network trained on this view's own probe searches (every prediction is
    # validated, so the grid does not depend on how well it fits)
    ua, _, _ = m.tile_skips(T, mode=0)
    net = nm.SkipNet()
    net.train(m.tile_features(T), m.tile_targets(ua), epochs=40, batch=256, lr=2e-3, seed=3)
    m.set_skip_mode(nm.SKIP_TILE_LEARNED, T, net)
    m.render()
    g = m.grid()
    st = m.tile_stats()
    assert st["ntiles"] == ua.size and st["accepted"] + st["rejected"] + st["bounded"] == st["ntiles"]
    o = O.OView(nr, nc)
    o.set_view(v["N"], v["re"], v["im"], v["width"])
    re, im = m.reference_point()
    o.precompute_at(re, im)
    o.set_series_mode(1)
    tua, _ = o.tile_probe_ua(T)
    assert np.array_equal(tua.ravel(), ua)
    rows = np.arange(135, nr, 270)
    rc = _rc(rows, nc)
    sample_ua = tua[rc[:, 0] // T, rc[:, 1] // T]
    meth = o.pert_samples_ua(rc, sample_ua, THREADS)
    mine = g[rows].reshape(-1)
    assert np.array_equal(mine["iterations"], meth["iterations"])
    assert _smooth_ok(mine["smoothing"], meth["smoothing"]).all()


@pytest.mark.parametrize("deep", [False, True])
def test_view_file_renders_like_the_oracle(gpu, tmp_path, deep):
    nm = gpu
    nr, nc = 45, 60
    m = nm.Mandelbrot(nr, nc)
    if deep:
        v = _VIEWS["t13"]
        m.set_view(v["N"], v["re"], v["im"], "1e-24")
    else:
        m.set_view(700, "-0.7453", "0.1127", "0.01")
    p = str(tmp_path / "view.dat")
    m.save_view(p)
    a = nm.Mandelbrot(nr, nc)
    o = O.OView(nr, nc)
    for _ in range(2):  # raise the precision first: loadLegacy parses at the current one
        a.zoom(1e9)
        o.L.orc_zoom(o.v, 1e9)
    a.loadLegacy(p)
    assert o.L.orc_load_legacy(o.v, p.encode()) == 0
    assert a.center == o.center()
    assert a.useHardware() == o.use_hardware() == (not deep)
    if deep:
        a.precompute()
        o.precompute()
        assert a.reference_length() == o.ref_len()
    a.computeAll()
    g = a.grid()
    rc = _rc(range(nr), nc)
    ref = o.render_samples(rc, THREADS, pert=deep).reshape(nr, nc)
    assert np.array_equal(g["iterations"], ref["iterations"])
    assert _smooth_ok(g["smoothing"], ref["smoothing"]).all()


@pytest.mark.parametrize("name,nr,nc,row,bar", [("c3_detail1", 1440, 2560, 719, 0.999),
                                                ("c5_detail1", 2160, 3840, 1079, 0.98)])
def test_floatexp_config_views_against_the_reference_method(gpu, name, nr, nc, row, bar):
    """The 2560x1440 view at 1e-200 (N 50000) and the 3840x2160 view at 1e-100
    (N 100000), series in floatexp, per-sample FP64 path: one middle row against
    the reference method (the same series skip, then the HP continuation,
    mandelbrot.cpp:209-224) with the same reference point.  At 1e-200 (a few
    thousand steps past the skip): >= 99.9% equal.  At 1e-100 samples run up to
    13,000 steps past the skip
    through chaotic tails, where the series' 1e-10 error at the skip point grows
    until a FP64 and an HP continuation of it part ways (98.5% equal on this
    row); there the reference method is itself off from the exact orbit: the
    first differing samples are checked against direct iteration from z = c at
    700 bits, and the reference method must disagree with it at each."""
    nm = gpu
    v = _VIEWS[name]
    m = nm.Mandelbrot(nr, nc)
    m.set_view(v["N"], v["re"], v["im"], v["width"])
    m.set_series_mode(nm.NM_SERIES_FLOATEXP)
    m.precompute()
    m.render()
    mine = m.grid()[row]
    o = O.OView(nr, nc)
    o.set_view(v["N"], v["re"], v["im"], v["width"])
    re, im = m.reference_point()
    o.precompute_at(re, im)
    o.set_series_mode(1)
    hp = o.render_rows(np.array([row], dtype=np.int32), THREADS).reshape(-1)
    same = mine["iterations"] == hp["iterations"]
    print("%s row %d: counts equal to the reference method on %.4f%% of %d samples (%d distinct counts)"
          % (name, row, 100 * same.mean(), same.size, len(np.unique(hp["iterations"]))))
    for c in np.nonzero(~same)[0][:3]:
        exact = _brute_count(v, nr, nc, row, int(c), bits=700)
        print("  column %d: GPU %d, reference method %d, 700-bit orbit %d" % (
            c, mine["iterations"][c], hp["iterations"][c], exact))
        assert exact != hp["iterations"][c]
    assert same.mean() >= bar
