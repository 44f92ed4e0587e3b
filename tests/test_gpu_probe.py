"""findProbe with the probes screened on the GPU.

The device estimates every probe's orbit length by FP64 perturbation against the
frame centre's orbit; only probes within 3 iterations of the best estimate are
iterated in HP.  The winner (first longest, mandelbrot.cpp:73-95), the reference
length and the rendered grid must equal the host search's, which iterates every
probe in HP, and the oracle's restatement of findProbe.
"""
import json
import os
import time

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu

_VIEWS = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "views.json")))
_EXTRA = {
    # a view inside the main cardioid (every probe reaches N: the HP scan stops
    # at its first batch), a seahorse-valley view with lengths spread over
    # hundreds of iterations, and a view outside the set where every probe
    # escapes after one step (every estimate ties)
    "interior": {"N": 3000, "re": "-0.1", "im": "0.1", "width": "1e-20"},
    "valley": {"N": 2000, "re": "-0.7453", "im": "0.1127", "width": "1e-14"},
    "outside": {"N": 200, "re": "2.5", "im": "2.5", "width": "1e-20"},
}
CASES = [
    ("interior", 48, 64, None), ("valley", 61, 83, None), ("outside", 30, 40, None),
    ("t13", 48, 64, "1e-25"), ("t13", 47, 65, "3e-23"), ("t11", 60, 80, "1e-15"),
    ("c2_seed1", 72, 128, "1e-30"), ("c4_seed1", 36, 48, "1e-40"), ("c5_seed1", 24, 32, "1e-100"),
]


def _view(name):
    return _EXTRA[name] if name in _EXTRA else _VIEWS[name]


@pytest.mark.parametrize("name,nr,nc,width", CASES)
def test_gpu_probe_search_equals_host_and_oracle(gpu, name, nr, nc, width):
    nm = gpu
    v = _view(name)
    w = width or v["width"]
    out = {}
    for mode in (nm.PROBE_HOST, nm.PROBE_GPU):
        m = nm.Mandelbrot(nr, nc)
        m.set_view(v["N"], v["re"], v["im"], w)
        m.set_probe_mode(mode)
        t0 = time.perf_counter()
        m.precompute()
        dt = time.perf_counter() - t0
        m.render()
        out[mode] = (m.probe_winner(), m.reference_length(), m.probe_candidates(), dt, m.grid().copy())
    o = O.OView(nr, nc)
    o.set_view(v["N"], v["re"], v["im"], w)
    o.precompute()
    hw, hl, hc, ht, hg = out[nm.PROBE_HOST]
    gw, gl, gc, gt, gg = out[nm.PROBE_GPU]
    print("%s %dx%d: winner %d len %d; HP orbits %d (host) -> %d (GPU screened); precompute %.3f s -> %.3f s"
          % (name, nr, nc, gw, gl, hc, gc, ht, gt))
    assert not m.useHardware()
    assert hw == o.winner()
    assert gw == hw and gl == hl
    assert 1 <= gc <= hc
    assert np.array_equal(gg["iterations"], hg["iterations"])
    assert np.array_equal(gg["smoothing"].view(np.uint32), hg["smoothing"].view(np.uint32))

