"""Edge shapes and iteration limits on the GPU against the oracle: single-pixel,
single-row and single-column frames, odd sizes that leave ragged start tiles and
partial chunks, and N from 1 (one update, mandelbrot.cpp:240-249 tests k >= N
before each) up, on the plain path and on the series path (series start and FP64
perturbation, the same reference point as the oracle's)."""
import json
import os

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu

_VIEWS = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "views.json")))
SHAPES = [(1, 1), (1, 7), (7, 1), (3, 5), (17, 33)]


def _smooth_ok(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return np.abs(a - b) <= 1e-6 * np.maximum(np.abs(b), 1.0)


@pytest.mark.parametrize("N", [1, 2, 5, 700])
@pytest.mark.parametrize("nr,nc", SHAPES)
def test_plain_path_edge_shapes(gpu, nr, nc, N):
    nm = gpu
    m = nm.Mandelbrot(nr, nc)
    m.set_view(N, "-0.75", "0.1", "3.0")
    assert m.useHardware()
    m.render()
    g = m.grid()
    o = O.OView(nr, nc)
    o.set_view(N, "-0.75", "0.1", "3.0")
    want = o.render_rows(np.arange(nr, dtype=np.int32), 2)
    assert np.array_equal(g["iterations"], want["iterations"])
    assert _smooth_ok(g["smoothing"], want["smoothing"]).all()


@pytest.mark.parametrize("N", [1, 3, 8000])
@pytest.mark.parametrize("nr,nc", SHAPES)
def test_series_path_edge_shapes(gpu, nr, nc, N):
    nm = gpu
    v = _VIEWS["t13"]
    m = nm.Mandelbrot(nr, nc)
    m.set_view(N, v["re"], v["im"], "1e-24")
    assert not m.useHardware()
    m.precompute()
    m.render()
    g = m.grid().reshape(-1)
    o = O.OView(nr, nc)
    o.set_view(N, v["re"], v["im"], "1e-24")
    re, im = m.reference_point()
    o.precompute_at(re, im)
    rc = np.array([(r, c) for r in range(nr) for c in range(nc)], dtype=np.int32)
    want = o.render_samples(rc, 2, pert=True)
    assert np.array_equal(g["iterations"], want["iterations"])
    assert _smooth_ok(g["smoothing"], want["smoothing"]).all()
