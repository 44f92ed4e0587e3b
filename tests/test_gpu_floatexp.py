"""The floatexp series mode on the GPU (an extension of the reference: its series
test and skip, mandelbrot.cpp:138-142, 160-178, with each real held as a double
mantissa and an int exponent).

* Where the reference's doubles are normal numbers the grids equal the double
  mode's bit for bit (per-sample skips and the tile searches).
* On the detail-walk views (width 1e-100, N 100000 and width 1e-200, N 50000)
  every sample equals the oracle's floatexp restatement of the GPU method with
  the same reference, and the tile searches equal the oracle's probe scans.
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


@pytest.mark.parametrize("name,nr,nc,width", [("t11", 90, 120, None), ("t13", 120, 160, None),
                                              ("t13", 96, 128, "3e-23"), ("c2_seed1", 216, 384, None)])
def test_floatexp_grid_equals_the_double_grid(gpu, name, nr, nc, width):
    nm = gpu
    v = _VIEWS[name]
    m = nm.Mandelbrot(nr, nc)
    m.set_view(v["N"], v["re"], v["im"], width or v["width"])
    m.precompute()
    m.render()
    g64 = m.grid().copy()
    ua64 = [m.tile_skips(16, mode=k)[0].copy() for k in (0, 1)]
    m.set_series_mode(nm.NM_SERIES_FLOATEXP)
    m.render()
    assert np.array_equal(m.grid(), g64)
    for k in (0, 1):
        assert np.array_equal(m.tile_skips(16, mode=k)[0], ua64[k])
    m.set_series_mode(nm.NM_SERIES_DOUBLE)
    m.render()
    assert np.array_equal(m.grid(), g64)


def _detail(name):
    if name not in _VIEWS:
        pytest.skip("%s is not in views.json" % name)
    return _VIEWS[name]


@pytest.mark.parametrize("name,nr,nc", [("c5_detail1", 108, 192), ("c3_detail1", 72, 128)])
def test_floatexp_deep_view_equals_the_oracle(gpu, name, nr, nc):
    nm = gpu
    v = _detail(name)
    m = nm.Mandelbrot(nr, nc)
    m.set_view(v["N"], v["re"], v["im"], v["width"])
    m.set_series_mode(nm.NM_SERIES_FLOATEXP)
    m.precompute()
    st = m.render()
    g = m.grid().reshape(-1)
    o = O.OView(nr, nc)
    o.set_view(v["N"], v["re"], v["im"], v["width"])
    re, im = m.reference_point()
    o.precompute_at(re, im)
    o.set_series_mode(1)
    rc = np.array([(r, c) for r in range(nr) for c in range(nc)], dtype=np.int32)
    meth = o.render_samples(rc, THREADS, pert=True)
    assert np.array_equal(g["iterations"], meth["iterations"])
    assert _smooth_ok(g["smoothing"], meth["smoothing"]).all()
    tua, _ = o.tile_probe_ua(16)
    for k in (0, 1):
        assert np.array_equal(m.tile_skips(16, mode=k)[0], tua.ravel()), k
    skipped = st["skipped"]
    print("%s %dx%d: %d distinct counts, %.1f%% of the iterations skipped by the series, tile stop indices %d..%d"
          % (name, nr, nc, len(np.unique(g["iterations"])), 100.0 * skipped / max(1, int(g["iterations"].sum())),
             tua.min(), tua.max()))
    assert len(np.unique(g["iterations"])) > 10 and tua.min() > 0
