"""nm_compute_all_noalloc (no counterpart in the reference): the plain-path
render with every device buffer in the library's static device scratch array,
checked against nm_compute_all (its allocating counterpart) bit for bit, on two
calls in a row, the second with a smaller frame.  The first test calls the
allocation-free entry point before anything else touches the device, so run
alone it exercises the path's own set-up."""
import numpy as np
import pytest

import newman_amd as nm

pytestmark = pytest.mark.gpu


def _view(nr, nc, N, cre, cim, width):
    m = nm.Mandelbrot(nr, nc)
    m.set_view(N, cre, cim, width)
    assert m.useHardware()
    return m


def _same(a, b):
    assert np.array_equal(a["iterations"], b["iterations"])
    assert np.array_equal(a["smoothing"].view(np.uint32), b["smoothing"].view(np.uint32))


def test_noalloc_equals_compute_all_two_calls():
    big = _view(240, 320, 1024, "-0.75", "0", "3.0")       # the 640x480 view at half size
    small = _view(60, 80, 700, "-0.7453", "0.1127", "0.01")  # seahorse valley, smaller frame
    out_big = big.compute_all_noalloc()       # first device call of the process when run alone
    out_small = small.compute_all_noalloc()   # second call, smaller input
    big.computeAll()
    small.computeAll()
    _same(out_big, big.grid())
    _same(out_small, small.grid())
    # the scratch array keeps the last call's contents: a repeat of the larger
    # frame after the smaller one gives the same grid again
    again = big.compute_all_noalloc(np.zeros((240, 320), dtype=nm.ESCAPE_DTYPE))
    _same(again, big.grid())
    assert (out_big["iterations"] < 1024).any() and (out_big["iterations"] == 1024).any()


def test_noalloc_bailout_and_odd_sizes():
    m = _view(37, 53, 300, "-1.25", "0.05", "0.7")
    m.set_bailout(2.0)
    out = m.compute_all_noalloc()
    m.computeAll()
    _same(out, m.grid())


def test_noalloc_rejects_what_it_cannot_render():
    m = nm.Mandelbrot(1024, 1024)  # 2^20 samples: above the scratch array's limit
    m.N = 64
    with pytest.raises(nm.NewmanError) as e:
        m.compute_all_noalloc()
    assert e.value.code == nm.NM_ERR_RANGE
    d = nm.Mandelbrot(20, 30)
    d.set_view(200, "-0.75", "0.1", "1e-20")  # deep view: needs the device reference orbit
    with pytest.raises(nm.NewmanError) as e:
        d.compute_all_noalloc()
    assert e.value.code == nm.NM_ERR_STATE
