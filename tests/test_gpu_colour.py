"""The viewer's colouring (viewer.cpp:89-122) on the GPU against the oracle's
restatement, byte for byte: palette lookup, the smoothing interpolation, black
at or beyond N, and the sc x sc colour average; from a caller's grid, from the
device copy of a render (no host round trip) and from frames of a batch."""
import json
import os

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")
DEFAULT_PAL = open(os.path.join(GOLD, "default.pal")).read()
_VIEWS = json.load(open(os.path.join(GOLD, "views.json")))


def _grid(rng, nr, nc, N):
    g = np.zeros((nr, nc), dtype=O.ESCAPE_DTYPE)
    its = rng.integers(0, N, size=(nr, nc))
    special = rng.random((nr, nc))
    its[special < 0.05] = 0          # pal[-1] stands in as pal[0]
    its[(special >= 0.05) & (special < 0.1)] = N - 1
    its[(special >= 0.1) & (special < 0.2)] = N      # black
    its[(special >= 0.2) & (special < 0.22)] = N + 7
    g["iterations"] = its
    g["smoothing"] = rng.random((nr, nc)).astype(np.float32)
    g["smoothing"][special > 0.97] = 1.0
    return g


@pytest.mark.parametrize("sc", [1, 2, 3, 4, 5])
@pytest.mark.parametrize("smooth", [False, True])
def test_colour_grid_equals_oracle(gpu, sc, smooth):
    nm = gpu
    rng = np.random.default_rng(10 * sc + smooth)
    N = 500
    pal = nm.Palette(text=DEFAULT_PAL).cache(N)
    g = _grid(rng, 50, 71, N)  # 71 is not a multiple of sc: trailing columns drop, as in colorLine
    img, ms = nm.colour_grid(g, N, pal, smooth, sc)
    want = O.colour(g, N, pal, smooth, sc)
    assert img.shape == (50 // sc, 71 // sc, 3)
    assert np.array_equal(img, want), np.argwhere(img != want)[:5]
    assert ms >= 0.0


def test_colour_of_a_render_from_the_device(gpu):
    """A deep view rendered with and without copying the grid back: nm_colour
    reads the device copy, equal to the oracle's colouring of the host grid; after
    scaleUp(2) and a render, the 2 x 2 average of the fine grid."""
    nm = gpu
    v = _VIEWS["t13"]
    m = nm.Mandelbrot(48, 64)
    m.set_view(v["N"], v["re"], v["im"], "3e-24")
    m.precompute()
    m.render()
    g = m.grid().copy()
    pal = nm.Palette(text=DEFAULT_PAL).cache(v["N"])
    for smooth in (False, True):
        img, _ = m.colour(pal, smooth, 1)
        assert np.array_equal(img, O.colour(g, v["N"], pal, smooth, 1))
    m.render(copy_back=False)  # the host grid is not refreshed; the device holds the result
    img, _ = m.colour(pal, True, 1)
    assert np.array_equal(img, O.colour(g, v["N"], pal, True, 1))
    m.scaleUp(2)
    m.precompute()
    m.render()
    fine = m.grid().copy()
    img, _ = m.colour(pal, True, 2)
    assert img.shape == (48, 64, 3)
    assert np.array_equal(img, O.colour(fine, v["N"], pal, True, 2))
    m.scaleDown(2)  # the view changed: the host grid (the 2 x 2 average) is coloured
    img, _ = m.colour(pal, False, 1)
    assert np.array_equal(img, O.colour(m.grid(), v["N"], pal, False, 1))


def test_colour_of_batch_frames(gpu):
    nm = gpu
    v = _VIEWS["c4_seed1"]
    m = nm.Mandelbrot(36, 48)
    widths = ["1e-40", "3e-40", "1e-39"]
    m.set_view(v["N"], v["re"], v["im"], widths[0])
    m.precompute()
    frames, _ = m.render_frames(widths)
    pal = nm.Palette(text=DEFAULT_PAL).cache(v["N"])
    for k in range(len(widths)):
        img, _ = m.colour(pal, True, 1, frame=k)
        assert np.array_equal(img, O.colour(frames[k], v["N"], pal, True, 1))
    with pytest.raises(nm.NewmanError):
        m.colour(pal, True, 1, frame=len(widths))
    with pytest.raises(nm.NewmanError):
        m.colour(pal[: v["N"] - 1], True, 1)  # the palette must cover N entries


@pytest.mark.parametrize("kr,kc,nr,nc,rate", [(27, 36, 18, 24, 5), (45, 60, 30, 40, 45), (31, 50, 20, 33, 7)])
def test_video_interpolate_equals_oracle(gpu, kr, kc, nr, nc, rate):
    nm = gpu
    rng = np.random.default_rng(kr + rate)
    prev = rng.integers(0, 256, size=(kr, kc, 3), dtype=np.uint8)
    nxt = rng.integers(0, 256, size=(kr, kc, 3), dtype=np.uint8)
    frames, ms = nm.video_interpolate(prev, nxt, nr, nc, rate)
    want = O.video_frames(prev, nxt, nr, nc, rate)
    assert frames.shape == (rate, nr, nc, 3)
    assert np.array_equal(frames, want), np.argwhere(frames != want)[:5]


def test_zoom_video_keyframes(gpu):
    """The viewer's zoom loop from the full view: initAutoZoom, then keyframes 1.5x
    apart; each keyframe is the 2 x 2 colour average of the 3 x 3-sampled grid,
    and the frames between two keyframes are VideoZoom::nextFrame's."""
    nm = gpu
    m = nm.Mandelbrot(12, 16)
    m.N = 256
    m.zoom_video_start(1)
    assert (m.rows(), m.cols()) == (36, 48)
    pal = nm.Palette(text=DEFAULT_PAL).cache(256)
    key1, frames, _ = m.zoom_video_step(pal, True, 9)
    assert frames is None and key1.shape == (18, 24, 3)
    assert np.array_equal(key1, O.colour(m.grid(), 256, pal, True, 2))
    key2, frames, _ = m.zoom_video_step(pal, True, 9, prev_key=key1)
    assert np.array_equal(key2, O.colour(m.grid(), 256, pal, True, 2))
    assert np.array_equal(frames, O.video_frames(key1, key2, 12, 16, 9))


def test_beauty_render(gpu):
    """beautyRender (viewer.cpp:186-253): the view re-rendered on a 3240 x 5760
    grid and coloured 3 x 3-averaged.  On a 3240-row handle the beauty grid is the
    handle's own view, so the image equals the handle's render coloured with
    sc = 3, byte for byte; the handle itself (view and grid) is unchanged."""
    nm = gpu
    v = _VIEWS["t13"]
    m = nm.Mandelbrot(3240, 5760)
    m.set_view(v["N"], v["re"], v["im"], "3e-24")
    m.precompute()
    m.render()
    g = m.grid().copy()
    pal = nm.Palette(text=DEFAULT_PAL).cache(v["N"])
    img = m.beauty_render(pal, True)
    assert img.shape == (1080, 1920, 3)
    assert np.array_equal(img, O.colour(g, v["N"], pal, True, 3))
    assert (m.rows(), m.cols()) == (3240, 5760)
    assert np.array_equal(m.grid(), g)
    small = nm.Mandelbrot(12, 16)  # the full view, sampled 270 times finer
    small.N = 256
    pal256 = nm.Palette(text=DEFAULT_PAL).cache(256)
    img = small.beauty_render(pal256, False)
    assert img.shape == (1080, 1920, 3)
    assert (small.rows(), small.cols()) == (12, 16)
    assert img[540, 960].tolist() != [0, 0, 0] or img[0, 0].tolist() != [0, 0, 0]
    with pytest.raises(nm.NewmanError):
        m.beauty_render(pal[: v["N"] - 1], True)
