"""ctypes binding of the CPU oracle (oracle/liboracle.so).  Test infrastructure."""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_LIB = os.path.join(ORACLE_DIR, "liboracle.so")

ESCAPE_DTYPE = np.dtype([("iterations", "<i4"), ("smoothing", "<f4")])


class OEscape(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("smoothing", C.c_float)]


_V = C.c_void_p
_SIGS = [
    ("orc_create", _V, [C.c_int, C.c_int]),
    ("orc_destroy", None, [_V]),
    ("orc_set_N", None, [_V, C.c_int]),
    ("orc_set_tol", None, [_V, C.c_double]),
    ("orc_set_bailout", None, [_V, C.c_double]),
    ("orc_set_series_mode", None, [_V, C.c_int]),
    ("orc_get_series_fx", None, [_V, _V, _V, _V, _V, _V, _V]),
    ("orc_set_center", C.c_int, [_V, C.c_char_p, C.c_char_p]),
    ("orc_set_size", C.c_int, [_V, C.c_char_p, C.c_char_p]),
    ("orc_set_width", C.c_int, [_V, C.c_char_p]),
    ("orc_set_view", C.c_int, [_V, C.c_int, C.c_char_p, C.c_char_p, C.c_char_p]),
    ("orc_set_precision", None, [_V]),
    ("orc_precision_bits", C.c_long, [_V]),
    ("orc_use_hardware", C.c_int, [_V]),
    ("orc_load_legacy", C.c_int, [_V, C.c_char_p]),
    ("orc_zoom", None, [_V, C.c_float]),
    ("orc_zoom_at", None, [_V, C.c_float, C.c_int, C.c_int, C.c_int]),
    ("orc_translate", None, [_V, C.c_int, C.c_int, C.c_int]),
    ("orc_scale_size", None, [_V, C.c_int, C.c_int]),
    ("orc_get_center", None, [_V, C.c_char_p, C.c_char_p, C.c_int]),
    ("orc_get_size_d", None, [_V, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    ("orc_precompute", C.c_int, [_V, C.c_int]),
    ("orc_precompute_at", C.c_int, [_V, C.c_char_p, C.c_char_p]),
    ("orc_probe_winner", C.c_int, [_V]),
    ("orc_ref_len", C.c_int, [_V]),
    ("orc_get_series", None, [_V, _V, _V, _V, _V, _V]),
    ("orc_render_rows", None, [_V, _V, C.c_int, _V, C.c_int]),
    ("orc_render_samples", None, [_V, _V, C.c_int, _V, C.c_int]),
    ("orc_pert_samples", None, [_V, _V, C.c_int, _V, C.c_int]),
    ("orc_pert_samples_ua", None, [_V, _V, _V, C.c_int, _V, C.c_int]),
    ("orc_pert_samples_fx", None, [_V, _V, _V, C.c_int, C.c_int, _V, C.c_int]),
    ("orc_tile_probe_ua", C.c_int, [_V, C.c_int, _V, _V]),
    ("orc_mlp_nparams", C.c_int, []),
    ("orc_pal_parse", C.c_int, [C.c_char_p]),
    ("orc_pal_cache", None, [C.c_int, _V]),
    ("orc_hsl2rgb", None, [C.c_float, C.c_float, C.c_float, _V]),
    ("orc_interp", None, [_V, _V, C.c_float, _V]),
    ("orc_clip", C.c_uint8, [C.c_float]),
    ("orc_colour", None, [_V, _V, C.c_int, C.c_int, C.c_int, _V, C.c_int, C.c_int, _V]),
    ("orc_video_frames", C.c_int, [_V, _V, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _V]),
    ("orc_mlp_forward", None, [_V, _V, C.c_int, _V]),
    ("orc_mlp_grad", C.c_double, [_V, _V, _V, C.c_int, _V]),
    ("orc_sa_start", C.c_int, [_V, C.c_int, C.c_int, C.POINTER(C.c_int), _V, C.POINTER(OEscape)]),
    ("orc_pixel_eps", None, [_V, C.c_int, C.c_int, _V]),
    ("orc_pixel_coord_d", None, [_V, C.c_int, C.c_int, _V]),
    ("orc_pixel_in_cardioid", C.c_int, [_V, C.c_int, C.c_int]),
    ("orc_point_hw", OEscape, [C.c_char_p, C.c_char_p, C.c_int, C.c_long]),
    ("orc_point_in_cardioid", C.c_int, [C.c_char_p, C.c_char_p, C.c_long]),
    ("orc_downsample", None, [_V, C.c_int, C.c_int, C.c_int, _V]),
    ("orc_upsample", None, [_V, C.c_int, C.c_int, C.c_int, _V]),
    ("orc_hp_op", C.c_int, [C.c_int, C.c_char_p, C.c_char_p, C.c_long, C.c_double, C.c_long,
                            C.POINTER(C.c_int), C.POINTER(C.c_long), _V]),
    ("orc_zoom_walk", C.c_int, [C.c_uint64, C.c_int, C.c_char_p, C.c_int, C.c_int, C.c_char_p, C.c_char_p, C.c_int,
                                C.POINTER(C.c_int)]),
    ("orc_center_render", None, [C.c_char_p, C.c_char_p, C.c_char_p, C.c_int, C.c_int, C.c_int, _V]),
]

_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", ORACLE_DIR])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_LIB):
            build()
        L = C.CDLL(ORACLE_LIB)
        for name, res, args in _SIGS:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _b(s):
    return s.encode() if isinstance(s, str) else s


def hp_op(op, a, b=None, iv=0, dv=0.0, bits=64):
    words = (C.c_uint32 * 256)()
    sign = C.c_int()
    exp = C.c_long()
    n = lib().orc_hp_op(op, _b(a), _b(b) if b is not None else None, iv, dv, bits, C.byref(sign), C.byref(exp),
                        words)
    assert n > 0
    return sign.value, exp.value, list(words[:n])


class OView:
    """The oracle's restatement of class Mandelbrot (mandelbrot.cpp)."""

    def __init__(self, nr, nc):
        self.L = lib()
        self.v = self.L.orc_create(nr, nc)
        self.nr, self.nc = nr, nc

    def __del__(self):
        if getattr(self, "v", None):
            self.L.orc_destroy(self.v)
            self.v = None

    def set_view(self, N, cre, cim, width):
        assert self.L.orc_set_view(self.v, N, _b(cre), _b(cim), _b(width)) == 0

    def set_N(self, N):
        self.L.orc_set_N(self.v, N)

    def set_tol(self, t):
        self.L.orc_set_tol(self.v, t)

    def set_bailout(self, b):
        self.L.orc_set_bailout(self.v, b)

    def set_series_mode(self, fxs):
        """1: the series test and skip in floatexp (an extension of the reference)"""
        self.L.orc_set_series_mode(self.v, int(fxs))

    def set_center(self, re, im):
        assert self.L.orc_set_center(self.v, _b(re), _b(im)) == 0

    def set_width(self, width):
        """sz = width / cols at the current precision (a batch frame's size)"""
        assert self.L.orc_set_width(self.v, _b(width)) == 0

    def set_size(self, re, im):
        assert self.L.orc_set_size(self.v, _b(re), _b(im)) == 0

    def set_precision(self):
        self.L.orc_set_precision(self.v)

    def bits(self):
        return self.L.orc_precision_bits(self.v)

    def use_hardware(self):
        return bool(self.L.orc_use_hardware(self.v))

    def precompute(self, nthreads=0):
        return self.L.orc_precompute(self.v, nthreads)

    def precompute_at(self, re, im):
        return self.L.orc_precompute_at(self.v, _b(re), _b(im))

    def winner(self):
        return self.L.orc_probe_winner(self.v)

    def ref_len(self):
        return self.L.orc_ref_len(self.v)

    def series(self):
        n = self.ref_len()
        Xt = np.zeros(2 * (n + 1))
        Xn = np.zeros(2 * (n + 1))
        A = np.zeros(2 * n)
        B = np.zeros(2 * n)
        Cc = np.zeros(2 * n)
        self.L.orc_get_series(self.v, Xt.ctypes.data, Xn.ctypes.data, A.ctypes.data, B.ctypes.data, Cc.ctypes.data)
        return Xt, Xn, A, B, Cc

    def series_fx(self):
        """The floatexp series: (mantissas, exponents) of A, B and C, each 2 per index."""
        n = self.ref_len()
        out = []
        for _ in range(3):
            out += [np.zeros(2 * n), np.zeros(2 * n, dtype=np.int32)]
        self.L.orc_get_series_fx(self.v, *[a.ctypes.data for a in out])
        return [(out[2 * k], out[2 * k + 1]) for k in range(3)]

    def render_rows(self, rows, nthreads=0):
        rows = np.ascontiguousarray(rows, dtype=np.int32)
        out = np.zeros(len(rows) * self.nc, dtype=ESCAPE_DTYPE)
        self.L.orc_render_rows(self.v, rows.ctypes.data, len(rows), out.ctypes.data, nthreads)
        return out.reshape(len(rows), self.nc)

    def render_samples(self, rc, nthreads=0, pert=False):
        rc = np.ascontiguousarray(rc, dtype=np.int32).reshape(-1, 2)
        out = np.zeros(len(rc), dtype=ESCAPE_DTYPE)
        fn = self.L.orc_pert_samples if pert else self.L.orc_render_samples
        fn(self.v, rc.ctypes.data, len(rc), out.ctypes.data, nthreads)
        return out

    def pert_samples_ua(self, rc, ua, nthreads=0):
        """The GPU method with a given first unstable index per sample."""
        rc = np.ascontiguousarray(rc, dtype=np.int32).reshape(-1, 2)
        ua = np.ascontiguousarray(ua, dtype=np.int32)
        out = np.zeros(len(rc), dtype=ESCAPE_DTYPE)
        self.L.orc_pert_samples_ua(self.v, rc.ctypes.data, ua.ctypes.data, len(rc), out.ctypes.data, nthreads)
        return out

    def pert_samples_fx(self, rc, K, ua=None, nthreads=0):
        """The GPU method with FP32X deltas (kernel block size K); ua: given first
        unstable index per sample, or None for each sample's own scan."""
        rc = np.ascontiguousarray(rc, dtype=np.int32).reshape(-1, 2)
        out = np.zeros(len(rc), dtype=ESCAPE_DTYPE)
        u = None if ua is None else np.ascontiguousarray(ua, dtype=np.int32)
        self.L.orc_pert_samples_fx(self.v, rc.ctypes.data, None if u is None else u.ctypes.data, len(rc), K,
                                   out.ctypes.data, nthreads)
        return out

    def tile_probe_ua(self, T):
        ntc, ntr = (self.nc + T - 1) // T, (self.nr + T - 1) // T
        ua = np.zeros(ntc * ntr, dtype=np.int32)
        rc = np.zeros(2 * ntc * ntr, dtype=np.int32)
        n = self.L.orc_tile_probe_ua(self.v, T, ua.ctypes.data, rc.ctypes.data)
        assert n == ntc * ntr
        return ua.reshape(ntr, ntc), rc.reshape(ntr, ntc, 2)

    def sa_start(self, r, c):
        found = C.c_int()
        d = np.zeros(2)
        e = OEscape()
        ua = self.L.orc_sa_start(self.v, r, c, C.byref(found), d.ctypes.data, C.byref(e))
        return ua, found.value, d

    def pixel_eps(self, r, c):
        e = np.zeros(2)
        self.L.orc_pixel_eps(self.v, r, c, e.ctypes.data)
        return e

    def in_cardioid(self, r, c):
        return bool(self.L.orc_pixel_in_cardioid(self.v, r, c))

    def center(self):
        re = C.create_string_buffer(4096)
        im = C.create_string_buffer(4096)
        self.L.orc_get_center(self.v, re, im, 4096)
        return re.value.decode(), im.value.decode()


def mlp_forward(params, x):
    params = np.ascontiguousarray(params, dtype=np.float32)
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.zeros(len(x))
    lib().orc_mlp_forward(params.ctypes.data, x.ctypes.data, len(x), y.ctypes.data)
    return y


def mlp_grad(params, x, t):
    params = np.ascontiguousarray(params, dtype=np.float32)
    x = np.ascontiguousarray(x, dtype=np.float32)
    t = np.ascontiguousarray(t, dtype=np.float32)
    g = np.zeros(lib().orc_mlp_nparams())
    loss = lib().orc_mlp_grad(params.ctypes.data, x.ctypes.data, t.ctypes.data, len(x), g.ctypes.data)
    return loss, g


def point_hw(re, im, N=1024, bits=64):
    e = lib().orc_point_hw(_b(re), _b(im), N, bits)
    return e.iterations, e.smoothing


def point_in_cardioid(re, im, bits=64):
    return bool(lib().orc_point_in_cardioid(_b(re), _b(im), bits))


def downsample(fine, sc):
    fine = np.ascontiguousarray(fine)
    nr, nc = fine.shape
    out = np.zeros((nr // sc) * (nc // sc), dtype=ESCAPE_DTYPE)
    lib().orc_downsample(fine.ctypes.data, nr, nc, sc, out.ctypes.data)
    return out.reshape(nr // sc, nc // sc)


def upsample(coarse, sc):
    coarse = np.ascontiguousarray(coarse)
    nr, nc = coarse.shape
    out = np.zeros(nr * sc * nc * sc, dtype=ESCAPE_DTYPE)
    lib().orc_upsample(coarse.ctypes.data, nr, nc, sc, out.ctypes.data)
    return out.reshape(nr * sc, nc * sc)


def zoom_walk(seed, N, target_width, res=256, nthreads=0):
    re = C.create_string_buffer(8192)
    im = C.create_string_buffer(8192)
    steps = C.c_int()
    assert lib().orc_zoom_walk(seed, N, _b(target_width), res, nthreads, re, im, 8192, C.byref(steps)) == 0
    return re.value.decode(), im.value.decode(), steps.value


def center_render(cre, cim, width, res, N, nthreads=0):
    out = np.zeros(res * res, dtype=np.int32)
    lib().orc_center_render(_b(cre), _b(cim), _b(width), res, N, nthreads, out.ctypes.data)
    return out.reshape(res, res)


# ---- colouring (oracle/colour.c)
def pal_cache(text, N):
    """MultiWaveGenerator::cache(N) of the .pal text: (N, 3) uint8."""
    assert lib().orc_pal_parse(text.encode()) == 0
    out = np.zeros((N, 3), dtype=np.uint8)
    lib().orc_pal_cache(N, out.ctypes.data)
    return out


def hsl2rgb(h, s, l):
    out = np.zeros(3, dtype=np.uint8)
    lib().orc_hsl2rgb(h, s, l, out.ctypes.data)
    return out


def interp(a, b, t):
    a = np.ascontiguousarray(a, dtype=np.uint8)
    b = np.ascontiguousarray(b, dtype=np.uint8)
    out = np.zeros(3, dtype=np.uint8)
    lib().orc_interp(a.ctypes.data, b.ctypes.data, t, out.ctypes.data)
    return out


def colour(grid, N, pal, smooth, sc):
    """FractalViewer::getColor / colorLine of an ESCAPE_DTYPE grid."""
    its = np.ascontiguousarray(grid["iterations"], dtype=np.int32)
    sms = np.ascontiguousarray(grid["smoothing"], dtype=np.float32)
    p = np.ascontiguousarray(pal, dtype=np.uint8)
    nr, nc = its.shape
    out = np.zeros((nr // sc, nc // sc, 3), dtype=np.uint8)
    lib().orc_colour(its.ctypes.data, sms.ctypes.data, nr, nc, int(N), p.ctypes.data, 1 if smooth else 0, int(sc),
                     out.ctypes.data)
    return out


def video_frames(prev, nxt, nr, nc, rate):
    """VideoZoom::nextFrame's frames: (rate, nr, nc, 3) uint8."""
    p = np.ascontiguousarray(prev, dtype=np.uint8)
    q = np.ascontiguousarray(nxt, dtype=np.uint8)
    out = np.zeros((rate, nr, nc, 3), dtype=np.uint8)
    assert lib().orc_video_frames(p.ctypes.data, q.ctypes.data, p.shape[0], p.shape[1], nr, nc, rate,
                                  out.ctypes.data) == 0
    return out
