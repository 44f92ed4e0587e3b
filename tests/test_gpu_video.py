"""Zoom-video frame sets on the GPU: sample orders, HIP-graph replay, autotuner.

Every launch setting must give the same grids bit for bit: they change only the
order in which samples are handed to the kernels, not the arithmetic.
"""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

_VIEWS = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "views.json")))


def _video(nm, nr=45, nc=70, nframes=9, name="t13", deepest="1e-24"):
    """A zoom sequence ending at `deepest`, each frame 1.1x wider than the next,
    sharing the deepest frame's reference (viewer.cpp:255-294 zooms around one
    centre)."""
    v = _VIEWS[name]
    d = float(deepest)
    widths = ["%.6e" % (d * 1.1 ** (nframes - 1 - k)) for k in range(nframes)]
    m = nm.Mandelbrot(nr, nc)
    m.set_view(v["N"], v["re"], v["im"], deepest)
    m.precompute()
    return m, widths


@pytest.mark.parametrize("tile", [8, 32, 64, 128])
def test_tiled_sample_order_equals_row_major(gpu, tile):
    nm = gpu
    m, widths = _video(nm)
    ref, st0 = m.render_frames(widths)
    m.set_launch(tile=tile)
    got, st = m.render_frames(widths)
    assert st["tile"] == tile
    assert np.array_equal(got["iterations"], ref["iterations"])
    assert np.array_equal(got["smoothing"].view(np.uint32), ref["smoothing"].view(np.uint32))
    assert st["credited"] == st0["credited"] and st["executed"] == st0["executed"]


@pytest.mark.parametrize("tile", [0, 32])
def test_tiled_order_plain_path(gpu, tile):
    nm = gpu
    m = nm.Mandelbrot(50, 67)
    m.set_view(600, "-0.75", "0.1", "2.5")
    ref, _ = m.render_frames(["2.5", "2.0", "1.5"])
    m.set_launch(tile=tile, block_k=16, waves_per_simd=2)
    got, st = m.render_frames(["2.5", "2.0", "1.5"])
    assert np.array_equal(got, ref)


def test_graph_replays_equal_eager(gpu):
    """repeats > 1 with graph=1: pass 1 eager, passes 2.. replayed from one
    captured graph (no memset nodes: counters are cleared by a kernel).  Each
    call's output and counters are those of its last replay; eager work and
    synchronisation run between the calls."""
    nm = gpu
    m, widths = _video(nm)
    eager, st0 = m.render_frames(widths)
    m.set_launch(tile=32, graph=True)
    for reps in (2, 4, 3):
        got, st = m.render_frames(widths, repeats=reps)
        assert st["graph"] == 1
        assert np.array_equal(got["iterations"], eager["iterations"])
        assert np.array_equal(got["smoothing"].view(np.uint32), eager["smoothing"].view(np.uint32))
        assert (st["credited"], st["executed"], st["skipped"], st["n_iterated"]) == \
            (st0["credited"], st0["executed"], st0["skipped"], st0["n_iterated"])
        one, _ = m.render_frames(widths[:2])  # eager work between graph calls
        assert np.array_equal(one, eager[:2])


def test_autotuner_picks_a_measured_setting(gpu):
    nm = gpu
    m, widths = _video(nm, nframes=20)
    ref, _ = m.render_frames(widths)
    best, table = m.autotune_frames(widths)
    # the sweep, then its fastest setting with interleaved chunks (series sets)
    assert len(table) == 4 * 4 * 3 + 1
    assert {r["tile"] for r in table} == {0, 32, 64, 128}
    assert {r["block_k"] for r in table} == {16, 32, 64}
    assert [r["chunk_run"] for r in table] == [0] * 48 + [4]
    fastest = min(table[:48], key=lambda r: r["ms"])
    assert {k: table[48][k] for k in ("block_k", "waves_per_simd", "tile")} == \
        {k: fastest[k] for k in ("block_k", "waves_per_simd", "tile")}
    assert best["ms"] == min(r["ms"] for r in table) and best["ms"] > 0
    assert (m.get_launch()["tile"], m.get_launch()["chunk_run"]) == (best["tile"], best["chunk_run"])
    got, st = m.render_frames(widths)
    assert (st["block_k"], st["waves_per_simd"], st["tile"]) == (best["block_k"], best["waves_per_simd"],
                                                                  best["tile"])
    assert np.array_equal(got, ref)


def test_autotuner_on_a_plain_frame_set(gpu):
    """A frame set on the plain path (shallow widths): the sweep covers 16-, 32-
    and 64-update blocks and the tuned launch renders the same grids."""
    nm = gpu
    m = nm.Mandelbrot(96, 128)
    m.set_view(512, "-0.75", "0", "3.0")
    widths = ["3.0", "2.5", "2.0", "1.5"]
    ref, _ = m.render_frames(widths)
    best, table = m.autotune_frames(widths)
    assert len(table) == 4 * 4 * 3
    assert {r["block_k"] for r in table} == {16, 32, 64}
    assert best["ms"] == min(r["ms"] for r in table) and best["ms"] > 0
    got, _ = m.render_frames(widths)
    assert np.array_equal(got, ref)


def test_device_frame_tables_follow_the_view(gpu):
    """The engine keeps a frame set's tables on the device between renders of the
    same set; any change of view, widths, sample order or object re-uploads."""
    nm = gpu
    m, widths = _video(nm)
    a, _ = m.render_frames(widths)
    b, _ = m.render_frames(widths[:4])
    assert np.array_equal(m.render_frames(widths)[0], a)
    m.set_launch(tile=64)
    assert np.array_equal(m.render_frames(widths)[0], a)
    m.set_launch(tile=0)
    p = nm.Mandelbrot(45, 70)  # another object (its own frame set) in between
    p.set_view(300, "-0.75", "0", "3.0")
    p.render()
    assert np.array_equal(m.render_frames(widths[:4])[0], b)
    m.render()  # the single-view set of the same object
    g1 = m.grid().copy()
    m.N = m.N - 1  # changes the view snapshot
    m.render()
    m.N = m.N + 1
    m.render()
    assert np.array_equal(m.grid(), g1)
