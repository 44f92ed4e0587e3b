"""The viewer's palette (multiwave.cpp) and colour helpers on the CPU: the
oracle pinned to closed forms, and the product's palette (host C++) equal to the
oracle's.  hsl2rgb / interp / clip belong to the viewer's image library, which is
not in the reference sources; the closed forms below are the checks that pin the
documented conventions (standard HSL hexcone with hue in degrees, HSL channels
rounded to the nearest byte, interpolation clamped and truncated)."""
import colorsys
import math
import os

import numpy as np
import pytest

import oracle_lib as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")
DEFAULT_PAL = open(os.path.join(GOLD, "default.pal")).read()  # the reference's default.pal


def _nm():
    import newman_amd as nm
    return nm


@pytest.mark.parametrize("h,rgb", [(0, (255, 0, 0)), (60, (255, 255, 0)), (120, (0, 255, 0)), (180, (0, 255, 255)),
                                   (240, (0, 0, 255)), (300, (255, 0, 255)), (360, (255, 0, 0)), (-120, (0, 0, 255))])
def test_hsl_primaries(h, rgb):
    assert tuple(O.hsl2rgb(h, 1.0, 0.5)) == rgb


def test_hsl_greys_and_extremes():
    for h in (0.0, 37.5, 221.226532):
        assert tuple(O.hsl2rgb(h, 0.0, 0.5)) == (128, 128, 128)  # 127.5 + 0.5
        assert tuple(O.hsl2rgb(h, 0.7, 0.0)) == (0, 0, 0)
        assert tuple(O.hsl2rgb(h, 0.7, 1.0)) == (255, 255, 255)


def test_hsl_matches_colorsys():
    """Against Python's colorsys (double precision, hue as a fraction of a turn),
    rounded to the nearest byte: equal except where float32 rounding crosses a
    half (then within 1)."""
    rng = np.random.default_rng(3)
    diffs = []
    for _ in range(2000):
        h, s, l = rng.uniform(0, 360), rng.uniform(0, 1), rng.uniform(0, 1)
        want = np.floor(np.array(colorsys.hls_to_rgb(h / 360.0, l, s)) * 255.0 + 0.5)
        got = O.hsl2rgb(h, s, l).astype(float)
        diffs.append(np.abs(got - want).max())
    diffs = np.array(diffs)
    assert diffs.max() <= 1 and (diffs == 0).mean() > 0.99


def test_interp_endpoints_and_truncation():
    a, b = np.array([10, 200, 255], np.uint8), np.array([250, 0, 3], np.uint8)
    assert tuple(O.interp(a, b, 0.0)) == (10, 200, 255)
    assert tuple(O.interp(a, b, 1.0)) == (250, 0, 3)
    assert tuple(O.interp([0, 0, 0], [255, 255, 255], 0.5)) == (127, 127, 127)  # 127.5 truncates


def test_palette_closed_forms():
    # one hue, one saturation, no luminance waves: every entry is hsl(120, 0.5, 0.5)
    assert (O.pal_cache("1\n1 120 1\n1\n1\n0.5 1\n0\n", 5) == [64, 191, 64]).all()
    # a two-value hue cycle of period 4: red, halfway (truncated), green, halfway
    p = O.pal_cache("1\n2 0 120 4\n1\n1\n1 1\n0\n", 4)
    assert p.tolist() == [[255, 0, 0], [127, 127, 0], [0, 255, 0], [127, 127, 0]]
    # a grey palette with one luminance wave of amplitude ln 3, period 4:
    # logistic(0) = 1/2, logistic(ln 3) = 3/4, logistic(-ln 3) = 1/4; at step 2
    # the argument is the float nearest pi, whose sine is -8.7e-8, so lum is just
    # under 1/2 and 127.5 + 0.5 truncates to 127
    g = O.pal_cache("1\n1 0 1\n1\n1\n0 1\n1\n%.9f 4\n" % math.log(3.0), 4)
    assert math.sin(np.float32(math.pi)) < 0
    assert g[:, 0].tolist() == [128, 191, 127, 64] and (g[:, 0] == g[:, 1]).all()
    # the reference's default palette starts at pure blue (hue 240, sat 1, lum 1/2)
    assert O.pal_cache(DEFAULT_PAL, 1).tolist() == [[0, 0, 255]]


@pytest.mark.parametrize("N", [1, 7, 256, 20000])
def test_product_palette_equals_oracle(N):
    nm = _nm()
    assert np.array_equal(nm.Palette(text=DEFAULT_PAL).cache(N), O.pal_cache(DEFAULT_PAL, N))


@pytest.mark.parametrize("text", ["1\n1 120 1\n1\n1\n0.5 1\n0\n", "1\n2 0 120 4\n1\n1\n1 1\n0\n",
                                  "2\n3 10 200 300 7\n2 90 45 5\n11\n3\n0.2 0.9 0.6 13\n2\n1.5 17\n0.25 5\n"])
def test_product_palette_equals_oracle_small(text):
    nm = _nm()
    assert np.array_equal(nm.Palette(text=text).cache(300), O.pal_cache(text, 300))


def test_palette_file_round_trip(tmp_path):
    """save_filename writes the reference's layout: the default palette saves to
    the same text it was read from, and reads back to the same colours."""
    nm = _nm()
    p = nm.Palette(text=DEFAULT_PAL)
    path = tmp_path / "p.pal"
    p.save(path)
    assert path.read_text() == DEFAULT_PAL
    q = nm.Palette(filename=path)
    assert np.array_equal(q.cache(1000), p.cache(1000))


def test_palette_rejects_bad_input(tmp_path):
    nm = _nm()
    with pytest.raises(nm.NewmanError):
        nm.Palette(text="2\n1 5")
    with pytest.raises(nm.NewmanError):
        nm.Palette(filename=tmp_path / "missing.pal")
    with pytest.raises(nm.NewmanError):
        nm.Palette().cache(10)  # nothing loaded


def test_colour_grid_rejects_negative_counts():
    """getColor reads pal[it] and pal[it - 1] (viewer.cpp:89-95): a caller's grid
    with a negative count is refused on the host, before any device work."""
    nm = _nm()
    g = np.zeros((4, 4), dtype=nm.ESCAPE_DTYPE)
    g["iterations"] = 3
    g["iterations"][2, 1] = -1
    pal = np.zeros((16, 3), dtype=np.uint8)
    with pytest.raises(nm.NewmanError) as e:
        nm.colour_grid(g, 16, pal, True, 1)
    assert e.value.code == nm.NM_ERR_ARG


def _keyframes(rng, kr, kc):
    return (rng.integers(0, 256, size=(kr, kc, 3), dtype=np.uint8),
            rng.integers(0, 256, size=(kr, kc, 3), dtype=np.uint8))


def test_video_first_frame_is_the_centre_crop():
    """Frame 0 of VideoZoom::nextFrame: lg = 1 (the previous keyframe at its own
    size, which bilinear scaling leaves unchanged) blitted centred, and the next
    keyframe blended with t = 0: the centre crop of the previous keyframe."""
    rng = np.random.default_rng(4)
    prev, nxt = _keyframes(rng, 27, 36)
    f = O.video_frames(prev, nxt, 18, 24, 5)
    r0, c0 = (27 - 18) // 2, (36 - 24) // 2
    assert np.array_equal(f[0], prev[r0:r0 + 18, c0:c0 + 24])


def test_video_constant_keyframes():
    """Constant keyframes A and B: outside the scaled next keyframe every frame is
    A; inside it, (1 - t) A + t B truncated, with t = i / rate."""
    prev = np.full((30, 45, 3), 200, np.uint8)
    nxt = np.full((30, 45, 3), 10, np.uint8)
    rate = 6
    f = O.video_frames(prev, nxt, 20, 30, rate)
    v, sm = np.float32(1.5 ** (1.0 / rate)), np.float32(2.0 / 3.0)
    for i in range(rate):
        t = np.float32(i) / np.float32(rate)
        srr, scc = int(np.float32(30) * sm), int(np.float32(45) * sm)
        r0, c0 = (20 - srr) // 2, (30 - scc) // 2
        inner = int((np.float32(1.0) - t) * np.float32(200) + t * np.float32(10))
        assert (f[i][max(r0, 0):r0 + srr, max(c0, 0):c0 + scc] == inner).all()
        mask = np.ones((20, 30), bool)
        mask[max(r0, 0):r0 + srr, max(c0, 0):c0 + scc] = False
        assert (f[i][mask] == 200).all()
        sm = np.float32(sm * v)
