"""Host side of the product against the oracle, and the C ABI surface (CPU only:
no kernel is launched here)."""
import ctypes as C
import json
import os
import random
import re

import numpy as np
import pytest

import oracle_lib as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")


def _nm():
    import newman_amd as nm
    return nm


def test_library_exports_every_declared_symbol():
    """Every function declared in include/newman_amd.h is exported by the library
    and bound by the Python mirror."""
    nm = _nm()
    hdr = open(os.path.join(ROOT, "include", "newman_amd.h")).read()
    declared = set(re.findall(r"^\s*(?:int|void|long|double|const char\*)\s+(nm_\w+)\s*\(", hdr, re.M))
    assert len(declared) >= 40
    L = C.CDLL(nm.LIB_PATH)
    for name in declared:
        assert hasattr(L, name), name
    assert declared == set(nm.EXPORTED)


def test_escape_value_layout():
    nm = _nm()
    assert C.sizeof(nm.EscapeValue) == 8 and nm.ESCAPE_DTYPE.itemsize == 8


def test_hp_product_matches_mpmath_goldens():
    nm = _nm()
    recs = json.load(open(os.path.join(GOLD, "hp_ops.json")))["records"]
    for r in recs:
        got = nm.hp_op(r["op"], r["a"], r["b"], r["iv"], r["dv"], r["bits"])
        assert got == (r["sign"], r["exp"], r["words"]), r


def test_hp_product_matches_oracle_random():
    nm = _nm()
    rng = random.Random(5)
    for _ in range(600):
        bits = rng.choice([64, 109, 200, 341, 672, 1200])
        a = "%s%de%d" % (rng.choice(["", "-"]), rng.randint(1, 10 ** 40), rng.randint(-90, 30))
        b = "%s%de%d" % (rng.choice(["", "-"]), rng.randint(1, 10 ** 40), rng.randint(-90, 30))
        op = rng.randint(0, 6)
        iv = rng.randint(1, 9999) if op == 4 else rng.randint(-9999, 9999)
        dv = rng.uniform(-5, 5)
        assert nm.hp_op(op, a, b, iv, dv, bits) == O.hp_op(op, a, b, iv, dv, bits)


def test_default_view_and_precision():
    """Mandelbrot(nr, nc): N 256, tol 1e-10, centre -0.5, sz 4/nc x 3/nr (mandelbrot.cpp:8-17)."""
    nm = _nm()
    m = nm.Mandelbrot(600, 800)
    assert m.N == 256 and m.error_tolerance == 1e-10
    assert m.sz_d() == (4.0 / 800, 3.0 / 600)
    assert float(m.center[0]) == -0.5 and float(m.center[1]) == 0.0
    assert m.useHardware() and m.precision_bits() == 64


def test_view_ops_match_oracle():
    """zoom, zoomAt, translate and the precision they set, against the oracle."""
    nm = _nm()
    m = nm.Mandelbrot(60, 80)
    o = O.OView(60, 80)
    for k in range(12):
        m.zoomAt(900.0, 20 + k, 33 - k, 1)
        o.L.orc_zoom_at(o.v, 900.0, 20 + k, 33 - k, 1)
        m.translate(3, -5, 1)
        o.L.orc_translate(o.v, 3, -5, 1)
    assert m.precision_bits() == o.bits()
    mc, oc = m.center, o.center()
    assert mc == oc
    re_, im_ = C.c_double(), C.c_double()
    o.L.orc_get_size_d(o.v, C.byref(re_), C.byref(im_))
    assert m.sz_d() == (re_.value, im_.value)
    assert not m.useHardware()


def test_view_file_round_trip(tmp_path):
    """The view file of viewer.cpp:6-29 read back by loadLegacy (mandelbrot.cpp:19-35),
    which rescales sz by 800/cols and 600/rows."""
    nm = _nm()
    m = nm.Mandelbrot(600, 800)
    v = json.load(open(os.path.join(GOLD, "views.json")))["t13"]
    m.set_view(v["N"], v["re"], v["im"], "1e-20")
    p = str(tmp_path / "view.dat")
    m.save_view(p)
    m2 = nm.Mandelbrot(600, 800)
    m2.zoom(1e9)  # raise the precision first: loadLegacy parses at the current one
    m2.zoom(1e9)
    m2.loadLegacy(p)
    assert m2.N == v["N"]
    assert m2.sz_d() == m.sz_d()
    o = O.OView(600, 800)
    o.L.orc_zoom(o.v, 1e9)
    o.L.orc_zoom(o.v, 1e9)
    assert o.L.orc_load_legacy(o.v, p.encode()) == 0
    assert m2.center == o.center()


@pytest.mark.parametrize("name,nr,nc,width", [("t11", 96, 128, "1e-16"), ("t13", 120, 160, "1e-24")])
def test_precompute_matches_oracle(name, nr, nc, width):
    """findProbe (winner, len) and the series as the device receives them are
    identical to the oracle's independent HP computation."""
    nm = _nm()
    v = json.load(open(os.path.join(GOLD, "views.json")))[name]
    m = nm.Mandelbrot(nr, nc)
    m.set_view(v["N"], v["re"], v["im"], width)
    o = O.OView(nr, nc)
    o.set_view(v["N"], v["re"], v["im"], width)
    m.precompute()
    o.precompute()
    assert m.reference_length() == o.ref_len()
    assert m.probe_winner() == o.winner()
    X, A, B, Cc = m.reference()
    Xt, Xn, OA, OB, OC = o.series()
    n = m.reference_length()
    assert np.array_equal(X, Xn[:2 * n])
    assert np.array_equal(A, OA) and np.array_equal(B, OB) and np.array_equal(Cc, OC)


def test_point_at_matches_oracle_pixel_map():
    nm = _nm()
    v = json.load(open(os.path.join(GOLD, "views.json")))["t13"]
    m = nm.Mandelbrot(30, 40)
    m.set_view(v["N"], v["re"], v["im"], "1e-22")
    o = O.OView(30, 40)
    o.set_view(v["N"], v["re"], v["im"], "1e-22")
    for r, c in [(0, 0), (15, 20), (29, 39), (7, 31)]:
        re_, im_ = m.pointAt(r, c)
        z = np.zeros(2)
        o.L.orc_pixel_coord_d(o.v, r, c, z.ctypes.data)
        assert float(re_) == pytest.approx(z[0], rel=1e-15) and float(im_) == pytest.approx(z[1], rel=1e-15)


def test_bad_arguments_return_status():
    nm = _nm()
    L = nm.lib()
    h = C.c_void_p()
    assert L.nm_create(C.byref(h), 0, 5) == nm.NM_ERR_ARG
    assert L.nm_create(C.byref(h), 4, 4) == nm.NM_OK
    assert L.nm_set_N(h, 0) == nm.NM_ERR_ARG
    assert L.nm_set_error_tolerance(h, -1.0) == nm.NM_ERR_ARG
    assert L.nm_set_center(h, b"1.2.3", b"0") == nm.NM_ERR_ARG
    assert L.nm_set_view(h, 10, b"0", b"0", b"-1") == nm.NM_ERR_ARG
    assert L.nm_load_legacy(h, b"/nonexistent/view.dat") == nm.NM_ERR_IO
    assert L.nm_compute_row(h, 99) == nm.NM_ERR_ARG
    out = (C.c_uint8 * 3)()
    assert L.nm_beauty_render(h, None, 256, 1, out, None) == nm.NM_ERR_ARG
    assert L.nm_beauty_render(h, out, 1, 1, None, None) == nm.NM_ERR_ARG
    assert L.nm_set_bailout(h, 1.0) == nm.NM_ERR_ARG  # the bailout must exceed 1
    assert L.nm_set_series_mode(h, 2) == nm.NM_ERR_ARG
    assert L.nm_set_series_mode(h, nm.NM_SERIES_FLOATEXP) == nm.NM_OK
    assert L.nm_set_series_mode(h, nm.NM_SERIES_DOUBLE) == nm.NM_OK
    L.nm_destroy(h)


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a HIP device may be present")
def test_no_device_fails_loudly():
    """No CPU fallback: without a HIP device a render is an error."""
    nm = _nm()
    m = nm.Mandelbrot(8, 8)
    with pytest.raises(nm.NewmanError) as e:
        m.computeAll()
    assert e.value.code == nm.NM_ERR_NODEVICE


def test_deep_render_before_precompute_is_a_state_error():
    nm = _nm()
    v = json.load(open(os.path.join(GOLD, "views.json")))["t13"]
    m = nm.Mandelbrot(8, 8)
    m.set_view(v["N"], v["re"], v["im"], "1e-22")
    L = nm.lib()
    rc = L.nm_compute_all(m._h)
    assert rc in (nm.NM_ERR_STATE, nm.NM_ERR_NODEVICE)


def test_launch_chunk_run_argument_checks():
    """The series kernel's chunk interleaving takes runs of 1, 2, 4, 8, 16, 32 or
    64 samples (0: the default, 64 consecutive samples); anything else is
    refused and leaves the launch settings as they were."""
    nm = _nm()
    m = nm.Mandelbrot(8, 8)
    for run in (0, 1, 2, 4, 8, 16, 32, 64):
        m.set_launch(chunk_run=run)
        assert m.get_launch()["chunk_run"] == run
    for bad in (-1, 3, 12, 128):
        with pytest.raises(nm.NewmanError):
            m.set_launch(chunk_run=bad)
    assert m.get_launch()["chunk_run"] == 64


def test_probe_mode_argument_checks():
    """nm_set_probe_mode takes NM_PROBE_HOST or NM_PROBE_GPU only; without a
    device precompute falls back to the host search (every probe in HP)."""
    nm = _nm()
    m = nm.Mandelbrot(8, 8)
    with pytest.raises(nm.NewmanError):
        m.set_probe_mode(2)
    m.set_probe_mode(nm.PROBE_HOST)
    m.set_probe_mode(nm.PROBE_GPU)
