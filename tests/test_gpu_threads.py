"""Two threads rendering through different handles at once: every engine call
runs under the C boundary's process-wide lock, so each thread's grids equal its
single-thread render (without the lock the shared engine's tables, results and
grow-only buffers were overwritten across threads)."""
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _views(nm):
    import json
    import os
    v = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "views.json")))["t13"]
    a = nm.Mandelbrot(120, 160)
    a.set_view(1024, "-0.75", "0", "3.0")  # plain path
    b = nm.Mandelbrot(48, 64)
    b.set_view(v["N"], v["re"], v["im"], "1e-24")  # series path
    b.precompute()
    return a, b


def test_two_threads_two_handles(gpu):
    nm = gpu
    a, b = _views(nm)
    a.computeAll()
    ra = a.grid().copy()
    b.computeAll()
    rb = b.grid().copy()
    errors = []

    def loop(m, ref, reps):
        try:
            for _ in range(reps):
                m.render(0, 0, 1, copy_back=True)
                g = m.grid()
                if not (np.array_equal(g["iterations"], ref["iterations"])
                        and np.array_equal(g["smoothing"], ref["smoothing"])):
                    errors.append("grid differs from the single-thread render")
                    return
        except Exception as e:  # noqa: BLE001 - reported below
            errors.append(repr(e))

    ts = [threading.Thread(target=loop, args=(a, ra, 12)), threading.Thread(target=loop, args=(b, rb, 12))]
    for t in ts:
        t.start()
    for t in ts:
        t.join(120)
    assert not any(t.is_alive() for t in ts)
    assert not errors, errors
