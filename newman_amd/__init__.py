"""newman_amd: MI355X-native per-pixel deep-zoom iteration for newman.

Python mirror of newman's render core class Mandelbrot (mandelbrot.h:7-53) over
the C ABI in include/newman_amd.h.  Every method forwards to the native library
(newman_amd/libnewman_amd.so, HIP kernels for gfx950); there is no CPU fallback:
rendering without a HIP device raises NewmanError.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libnewman_amd.so")

NM_OK = 0
NM_ERR_ARG = -1
NM_ERR_HIP = -2
NM_ERR_STATE = -3
NM_ERR_NODEVICE = -4
NM_ERR_RANGE = -5
NM_ERR_IO = -6
NM_ERR_NOMEM = -7
NM_DELTA_FP64 = 0
NM_DELTA_FP32X = 1
NM_SERIES_DOUBLE = 0
NM_SERIES_FLOATEXP = 1

_ERRNAMES = {NM_ERR_ARG: "NM_ERR_ARG", NM_ERR_HIP: "NM_ERR_HIP", NM_ERR_STATE: "NM_ERR_STATE",
             NM_ERR_NODEVICE: "NM_ERR_NODEVICE", NM_ERR_RANGE: "NM_ERR_RANGE", NM_ERR_IO: "NM_ERR_IO",
             NM_ERR_NOMEM: "NM_ERR_NOMEM"}

ESCAPE_DTYPE = np.dtype([("iterations", "<i4"), ("smoothing", "<f4")])  # grid.h:8-15


class NewmanError(RuntimeError):
    def __init__(self, code, what):
        self.code = code
        super().__init__("%s failed: %s (%d) %s" % (what, _ERRNAMES.get(code, "?"), code, last_error()))


class EscapeValue(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("smoothing", C.c_float)]


class RenderStats(C.Structure):
    _fields_ = [("samples", C.c_uint64), ("n_cardioid", C.c_uint64), ("n_sa_resolved", C.c_uint64),
                ("n_iterated", C.c_uint64), ("executed", C.c_uint64), ("skipped", C.c_uint64),
                ("n_rebased", C.c_uint64), ("credited", C.c_uint64), ("ms_start", C.c_double),
                ("ms_iterate", C.c_double), ("ms_total", C.c_double), ("block_k", C.c_int32),
                ("waves_per_simd", C.c_int32), ("executed_start", C.c_uint64), ("ms_compact", C.c_double),
                ("ms_finalize", C.c_double), ("graph", C.c_int32), ("tile", C.c_int32),
                ("fma2", C.c_int32), ("grouped", C.c_int32), ("chunk_run", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Launch(C.Structure):
    _fields_ = [("block_k", C.c_int32), ("waves_per_simd", C.c_int32), ("tile", C.c_int32), ("graph", C.c_int32),
                ("pre_iter", C.c_int32), ("chunk_run", C.c_int32), ("ms", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class TileStats(C.Structure):
    _fields_ = [("ntiles", C.c_int32), ("accepted", C.c_int32), ("rejected", C.c_int32), ("bounded", C.c_int32),
                ("evals", C.c_uint64), ("ms_search", C.c_double), ("ms_predict", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


SKIP_PIXEL, SKIP_TILE_SCAN, SKIP_TILE_TABLE, SKIP_TILE_LEARNED, SKIP_NONE = 0, 1, 2, 3, 4
PROBE_HOST, PROBE_GPU = 0, 1
SKIPNET_INPUTS, SKIPNET_PARAMS = 132, 12737


class Palette:
    """newman's multi-wave palette (multiwave.h): load/parse the .pal text, then
    cache(N) -> (N, 3) uint8 RGB."""

    def __init__(self, text=None, filename=None):
        h = _H()
        _check(lib().nm_palette_create(C.byref(h)), "nm_palette_create")
        self._h = h
        if text is not None:
            _check(lib().nm_palette_parse(self._h, _b(text)), "nm_palette_parse")
        if filename is not None:
            _check(lib().nm_palette_load(self._h, _b(str(filename))), "nm_palette_load")

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.nm_palette_destroy(self._h)
            self._h = None

    def save(self, filename):
        _check(lib().nm_palette_save(self._h, _b(str(filename))), "nm_palette_save")

    def cache(self, N):
        out = np.zeros((int(N), 3), dtype=np.uint8)
        _check(lib().nm_palette_cache(self._h, int(N), out.ctypes.data), "nm_palette_cache")
        return out


def colour_grid(grid, N, pal, smooth=True, sc=1):
    """FractalViewer::getColor / colorLine on the GPU for a caller's grid (an
    ESCAPE_DTYPE array, rows x cols); returns ((rows // sc, cols // sc, 3) uint8, kernel ms)."""
    g = np.ascontiguousarray(grid, dtype=ESCAPE_DTYPE)
    p = np.ascontiguousarray(pal, dtype=np.uint8)
    nr, nc = g.shape
    out = np.zeros((nr // sc, nc // sc, 3), dtype=np.uint8)
    ms = C.c_double(0.0)
    _check(lib().nm_colour_grid(g.ctypes.data, nr, nc, int(N), p.ctypes.data, p.shape[0], 1 if smooth else 0,
                                int(sc), out.ctypes.data, C.byref(ms)), "nm_colour_grid")
    return out, ms.value


def video_interpolate(prev, nxt, nr, nc, rate):
    """VideoZoom::nextFrame's in-between frames on the GPU: (rate, nr, nc, 3)
    uint8 between keyframes prev and nxt ((kr, kc, 3) uint8), and the kernel ms."""
    p = np.ascontiguousarray(prev, dtype=np.uint8)
    q = np.ascontiguousarray(nxt, dtype=np.uint8)
    assert p.shape == q.shape and p.ndim == 3 and p.shape[2] == 3
    out = np.zeros((int(rate), int(nr), int(nc), 3), dtype=np.uint8)
    ms = C.c_double(0.0)
    _check(lib().nm_video_interpolate(p.ctypes.data, q.ctypes.data, p.shape[0], p.shape[1], int(nr), int(nc),
                                      int(rate), out.ctypes.data, C.byref(ms)), "nm_video_interpolate")
    return out, ms.value


class SkipNet:
    """The learned series-skip predictor (132-64-64-1 ReLU MLP, trained on the GPU)."""

    def __init__(self):
        h = _H()
        _check(lib().nm_skipnet_create(C.byref(h)), "nm_skipnet_create")
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.nm_skipnet_destroy(self._h)
            self._h = None

    def get(self):
        p = np.zeros(SKIPNET_PARAMS, dtype=np.float32)
        mu = np.zeros(SKIPNET_INPUTS, dtype=np.float32)
        sc = np.zeros(SKIPNET_INPUTS, dtype=np.float32)
        _check(lib().nm_skipnet_get(self._h, p.ctypes.data, mu.ctypes.data, sc.ctypes.data), "nm_skipnet_get")
        return p, mu, sc

    def set(self, params, mean=None, scale=None):
        p = np.ascontiguousarray(params, dtype=np.float32)
        mu = None if mean is None else np.ascontiguousarray(mean, dtype=np.float32)
        sc = None if scale is None else np.ascontiguousarray(scale, dtype=np.float32)
        _check(lib().nm_skipnet_set(self._h, p.ctypes.data, None if mu is None else mu.ctypes.data,
                                    None if sc is None else sc.ctypes.data), "nm_skipnet_set")

    def train(self, x, t, epochs=200, batch=256, lr=3e-3, seed=1):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        curve = np.zeros(epochs, dtype=np.float32)
        _check(lib().nm_skipnet_train(self._h, x.ctypes.data, t.ctypes.data, len(t), epochs, batch, lr, seed,
                                      curve.ctypes.data), "nm_skipnet_train")
        return curve

    def predict(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.zeros(len(x), dtype=np.float32)
        _check(lib().nm_skipnet_predict(self._h, x.ctypes.data, len(x), y.ctypes.data), "nm_skipnet_predict")
        return y

    def gradient(self, xn, t):
        xn = np.ascontiguousarray(xn, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        g = np.zeros(SKIPNET_PARAMS, dtype=np.float32)
        loss = C.c_float()
        _check(lib().nm_skipnet_gradient(self._h, xn.ctypes.data, t.ctypes.data, len(t), g.ctypes.data,
                                         C.byref(loss)), "nm_skipnet_gradient")
        return loss.value, g


_lib = None

# (name, restype, argtypes) for every declaration of include/newman_amd.h
_H = C.c_void_p
_SIGS = [
    ("nm_create", C.c_int, [C.POINTER(_H), C.c_int, C.c_int]),
    ("nm_clone", C.c_int, [C.POINTER(_H), _H]),
    ("nm_destroy", None, [_H]),
    ("nm_rows", C.c_int, [_H]),
    ("nm_cols", C.c_int, [_H]),
    ("nm_set_N", C.c_int, [_H, C.c_int]),
    ("nm_get_N", C.c_int, [_H]),
    ("nm_set_error_tolerance", C.c_int, [_H, C.c_double]),
    ("nm_get_error_tolerance", C.c_double, [_H]),
    ("nm_set_center", C.c_int, [_H, C.c_char_p, C.c_char_p]),
    ("nm_get_center", C.c_int, [_H, C.c_char_p, C.c_char_p, C.c_size_t]),
    ("nm_set_size", C.c_int, [_H, C.c_char_p, C.c_char_p]),
    ("nm_get_size", C.c_int, [_H, C.c_char_p, C.c_char_p, C.c_size_t]),
    ("nm_get_size_d", C.c_int, [_H, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    ("nm_set_precision", C.c_int, [_H]),
    ("nm_precision_bits", C.c_long, [_H]),
    ("nm_use_hardware", C.c_int, [_H]),
    ("nm_precompute", C.c_int, [_H]),
    ("nm_compute_row", C.c_int, [_H, C.c_int]),
    ("nm_compute_all", C.c_int, [_H]),
    ("nm_compute_all_noalloc", C.c_int, [_H, C.c_void_p, C.c_size_t]),
    ("nm_at", C.c_int, [_H, C.c_int, C.c_int, C.POINTER(EscapeValue)]),
    ("nm_at_avg", C.c_int, [_H, C.c_int, C.c_int, C.c_int, C.POINTER(EscapeValue)]),
    ("nm_get_grid", C.c_int, [_H, C.c_void_p, C.c_size_t]),
    ("nm_point_at", C.c_int, [_H, C.c_int, C.c_int, C.c_int, C.c_char_p, C.c_char_p, C.c_size_t]),
    ("nm_translate", C.c_int, [_H, C.c_int, C.c_int, C.c_int]),
    ("nm_zoom", C.c_int, [_H, C.c_float]),
    ("nm_zoom_at", C.c_int, [_H, C.c_float, C.c_int, C.c_int, C.c_int]),
    ("nm_scale_up", C.c_int, [_H, C.c_int]),
    ("nm_scale_down", C.c_int, [_H, C.c_int]),
    ("nm_load_legacy", C.c_int, [_H, C.c_char_p]),
    ("nm_save_view", C.c_int, [_H, C.c_char_p]),
    ("nm_set_bailout", C.c_int, [_H, C.c_double]),
    ("nm_set_delta_mode", C.c_int, [_H, C.c_int]),
    ("nm_set_series_mode", C.c_int, [_H, C.c_int]),
    ("nm_set_view", C.c_int, [_H, C.c_int, C.c_char_p, C.c_char_p, C.c_char_p]),
    ("nm_precompute_at", C.c_int, [_H, C.c_char_p, C.c_char_p]),
    ("nm_reference_length", C.c_int, [_H]),
    ("nm_probe_winner", C.c_int, [_H]),
    ("nm_set_probe_mode", C.c_int, [_H, C.c_int]),
    ("nm_probe_candidates", C.c_int, [_H]),
    ("nm_get_reference", C.c_int, [_H, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("nm_render", C.c_int, [_H, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(RenderStats)]),
    ("nm_render_frames", C.c_int, [_H, C.POINTER(C.c_char_p), C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p,
                                   C.POINTER(RenderStats)]),
    ("nm_set_launch", C.c_int, [_H, C.c_void_p]),
    ("nm_get_launch", C.c_int, [_H, C.c_void_p]),
    ("nm_autotune_frames", C.c_int, [_H, C.POINTER(C.c_char_p), C.c_int, C.c_void_p, C.c_void_p, C.c_int,
                                     C.POINTER(C.c_int)]),
    ("nm_skipnet_create", C.c_int, [C.POINTER(_H)]),
    ("nm_skipnet_destroy", None, [_H]),
    ("nm_skipnet_get", C.c_int, [_H, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("nm_skipnet_set", C.c_int, [_H, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("nm_skipnet_train", C.c_int, [_H, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_float, C.c_uint64,
                                   C.c_void_p]),
    ("nm_skipnet_predict", C.c_int, [_H, C.c_void_p, C.c_int, C.c_void_p]),
    ("nm_skipnet_gradient", C.c_int, [_H, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]),
    ("nm_tile_features", C.c_int, [_H, C.c_int, C.c_void_p, C.c_int, C.POINTER(C.c_int)]),
    ("nm_tile_skips", C.c_int, [_H, C.c_int, C.c_int, _H, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]),
    ("nm_tile_targets", C.c_int, [_H, C.c_void_p, C.c_int, C.c_void_p]),
    ("nm_reference_point", C.c_int, [_H, C.c_char_p, C.c_char_p, C.c_size_t]),
    ("nm_set_skip_mode", C.c_int, [_H, C.c_int, C.c_int, _H]),
    ("nm_get_tile_stats", C.c_int, [_H, C.c_void_p]),
    ("nm_palette_create", C.c_int, [C.POINTER(_H)]),
    ("nm_palette_destroy", None, [_H]),
    ("nm_palette_load", C.c_int, [_H, C.c_char_p]),
    ("nm_palette_save", C.c_int, [_H, C.c_char_p]),
    ("nm_palette_parse", C.c_int, [_H, C.c_char_p]),
    ("nm_palette_cache", C.c_int, [_H, C.c_int, C.c_void_p]),
    ("nm_colour", C.c_int, [_H, C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p,
                            C.POINTER(C.c_double)]),
    ("nm_colour_grid", C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_int,
                                 C.c_void_p, C.POINTER(C.c_double)]),
    ("nm_video_interpolate", C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                       C.c_void_p, C.POINTER(C.c_double)]),
    ("nm_zoom_video_start", C.c_int, [_H, C.c_int]),
    ("nm_zoom_video_step", C.c_int, [_H, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                     C.POINTER(C.c_int), C.POINTER(C.c_double)]),
    ("nm_beauty_render", C.c_int, [_H, C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.POINTER(C.c_double)]),
    ("nm_last_error", C.c_char_p, []),
    ("nm_hp_op", C.c_int, [C.c_int, C.c_char_p, C.c_char_p, C.c_long, C.c_double, C.c_long,
                           C.POINTER(C.c_int), C.POINTER(C.c_long), C.c_void_p]),
]
EXPORTED = [s[0] for s in _SIGS]


def lib():
    """Load the native library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise NewmanError(NM_ERR_STATE, "loading %s (run python -m newman_amd.build)" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        for name, res, args in _SIGS:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def last_error():
    if _lib is None:
        return ""
    s = _lib.nm_last_error()
    return s.decode() if s else ""


def _check(code, what):
    if code < 0:
        raise NewmanError(code, what)
    return code


def _b(s):
    return s.encode() if isinstance(s, str) else s


def hp_op(op, a, b=None, iv=0, dv=0.0, bits=64):
    """Native HP arithmetic probe (see nm_hp_op): returns (sign, exp, words)."""
    L = lib()
    words = (C.c_uint32 * 256)()
    sign = C.c_int()
    exp = C.c_long()
    n = _check(L.nm_hp_op(op, _b(a), _b(b) if b is not None else None, iv, dv, bits, C.byref(sign), C.byref(exp),
                          words), "nm_hp_op")
    return sign.value, exp.value, list(words[:n])


class Mandelbrot:
    """newman's render core (mandelbrot.h:7-53) on the MI355X engine."""

    def __init__(self, nr=1, nc=1, _handle=None):
        L = lib()
        if _handle is not None:
            self._h = _handle
        else:
            h = _H()
            _check(L.nm_create(C.byref(h), nr, nc), "nm_create")
            self._h = h

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.nm_destroy(self._h)
            self._h = None

    def copy(self):
        h = _H()
        _check(lib().nm_clone(C.byref(h), self._h), "nm_clone")
        return Mandelbrot(_handle=h)

    # ---- public fields (mandelbrot.h:23-25)
    @property
    def N(self):
        return lib().nm_get_N(self._h)

    @N.setter
    def N(self, v):
        _check(lib().nm_set_N(self._h, int(v)), "nm_set_N")

    @property
    def error_tolerance(self):
        return lib().nm_get_error_tolerance(self._h)

    @error_tolerance.setter
    def error_tolerance(self, v):
        _check(lib().nm_set_error_tolerance(self._h, float(v)), "nm_set_error_tolerance")

    @property
    def center(self):
        re = C.create_string_buffer(4096)
        im = C.create_string_buffer(4096)
        _check(lib().nm_get_center(self._h, re, im, 4096), "nm_get_center")
        return re.value.decode(), im.value.decode()

    @center.setter
    def center(self, v):
        _check(lib().nm_set_center(self._h, _b(str(v[0])), _b(str(v[1]))), "nm_set_center")

    @property
    def sz(self):
        re = C.create_string_buffer(4096)
        im = C.create_string_buffer(4096)
        _check(lib().nm_get_size(self._h, re, im, 4096), "nm_get_size")
        return re.value.decode(), im.value.decode()

    @sz.setter
    def sz(self, v):
        _check(lib().nm_set_size(self._h, _b(str(v[0])), _b(str(v[1]))), "nm_set_size")

    def sz_d(self):
        re, im = C.c_double(), C.c_double()
        _check(lib().nm_get_size_d(self._h, C.byref(re), C.byref(im)), "nm_get_size_d")
        return re.value, im.value

    # ---- methods
    def rows(self):
        return lib().nm_rows(self._h)

    def cols(self):
        return lib().nm_cols(self._h)

    def setPrecision(self):
        _check(lib().nm_set_precision(self._h), "nm_set_precision")

    def precision_bits(self):
        return lib().nm_precision_bits(self._h)

    def useHardware(self):
        return bool(_check(lib().nm_use_hardware(self._h), "nm_use_hardware"))

    def precompute(self):
        _check(lib().nm_precompute(self._h), "nm_precompute")

    def computeRow(self, r):
        _check(lib().nm_compute_row(self._h, int(r)), "nm_compute_row")

    def computeAll(self):
        _check(lib().nm_compute_all(self._h), "nm_compute_all")

    def compute_all_noalloc(self, out=None):
        """nm_compute_all_noalloc: the plain-path grid into `out` (an ESCAPE_DTYPE
        array of rows x cols, allocated here when None) without device allocation."""
        if out is None:
            out = np.zeros((self.rows(), self.cols()), dtype=ESCAPE_DTYPE)
        assert out.dtype == ESCAPE_DTYPE and out.flags.c_contiguous
        _check(lib().nm_compute_all_noalloc(self._h, out.ctypes.data, out.size), "nm_compute_all_noalloc")
        return out

    def at(self, r, c, sc=None):
        e = EscapeValue()
        if sc is None:
            _check(lib().nm_at(self._h, r, c, C.byref(e)), "nm_at")
        else:
            _check(lib().nm_at_avg(self._h, r, c, sc, C.byref(e)), "nm_at_avg")
        return e.iterations, e.smoothing

    def colour(self, pal, smooth=True, sc=1, frame=-1):
        """The viewer's colouring of this grid (frame < 0) or of frame `frame` of the
        last render_frames, on the GPU; returns ((rows // sc, cols // sc, 3) uint8,
        kernel ms)."""
        p = np.ascontiguousarray(pal, dtype=np.uint8)
        out = np.zeros((self.rows() // sc, self.cols() // sc, 3), dtype=np.uint8)
        ms = C.c_double(0.0)
        _check(lib().nm_colour(self._h, int(frame), p.ctypes.data, p.shape[0], 1 if smooth else 0, int(sc),
                               out.ctypes.data, C.byref(ms)), "nm_colour")
        return out, ms.value

    def beauty_render(self, pal, smooth=True):
        """beautyRender (viewer.cpp:186-253): (1080, 1920, 3) uint8 of this view
        rendered 3 x 3-multisampled; this object is unchanged."""
        p = np.ascontiguousarray(pal, dtype=np.uint8)
        out = np.zeros((1080, 1920, 3), dtype=np.uint8)
        ms = C.c_double(0.0)
        _check(lib().nm_beauty_render(self._h, p.ctypes.data, p.shape[0], 1 if smooth else 0, out.ctypes.data,
                                      C.byref(ms)), "nm_beauty_render")
        return out

    def zoom_video_start(self, sc=1):
        """initAutoZoom (viewer.cpp:441-453)."""
        _check(lib().nm_zoom_video_start(self._h, int(sc)), "nm_zoom_video_start")

    def zoom_video_step(self, pal, smooth=True, rate=45, prev_key=None):
        """One keyframe of the viewer's zoom loop; returns (keyframe, frames or
        None, kernel ms)."""
        p = np.ascontiguousarray(pal, dtype=np.uint8)
        key = np.zeros((self.rows() // 2, self.cols() // 2, 3), dtype=np.uint8)
        frames = None
        if prev_key is not None:
            prev_key = np.ascontiguousarray(prev_key, dtype=np.uint8)
            frames = np.zeros((int(rate), self.rows() // 3, self.cols() // 3, 3), dtype=np.uint8)
        n = C.c_int(0)
        ms = C.c_double(0.0)
        _check(lib().nm_zoom_video_step(self._h, p.ctypes.data, p.shape[0], 1 if smooth else 0, int(rate),
                                        None if prev_key is None else prev_key.ctypes.data, key.ctypes.data,
                                        None if frames is None else frames.ctypes.data, C.byref(n), C.byref(ms)),
               "nm_zoom_video_step")
        return key, frames, ms.value

    def grid(self):
        out = np.zeros(self.rows() * self.cols(), dtype=ESCAPE_DTYPE)
        _check(lib().nm_get_grid(self._h, out.ctypes.data, out.size), "nm_get_grid")
        return out.reshape(self.rows(), self.cols())

    def pointAt(self, r, c, sc=1):
        re = C.create_string_buffer(4096)
        im = C.create_string_buffer(4096)
        _check(lib().nm_point_at(self._h, r, c, sc, re, im, 4096), "nm_point_at")
        return re.value.decode(), im.value.decode()

    def translate(self, dr, dc, sc=1):
        _check(lib().nm_translate(self._h, dr, dc, sc), "nm_translate")

    def zoom(self, scale):
        _check(lib().nm_zoom(self._h, scale), "nm_zoom")

    def zoomAt(self, scale, r, c, sc=1):
        _check(lib().nm_zoom_at(self._h, scale, r, c, sc), "nm_zoom_at")

    def scaleUp(self, sc):
        _check(lib().nm_scale_up(self._h, sc), "nm_scale_up")

    def scaleDown(self, sc):
        _check(lib().nm_scale_down(self._h, sc), "nm_scale_down")

    def loadLegacy(self, path):
        _check(lib().nm_load_legacy(self._h, _b(path)), "nm_load_legacy")

    def save_view(self, path):
        _check(lib().nm_save_view(self._h, _b(path)), "nm_save_view")

    # ---- extensions
    def set_bailout(self, b):
        _check(lib().nm_set_bailout(self._h, float(b)), "nm_set_bailout")

    def set_delta_mode(self, mode):
        _check(lib().nm_set_delta_mode(self._h, int(mode)), "nm_set_delta_mode")

    def set_series_mode(self, mode):
        """NM_SERIES_DOUBLE (the reference's doubles) or NM_SERIES_FLOATEXP (extension)."""
        _check(lib().nm_set_series_mode(self._h, int(mode)), "nm_set_series_mode")

    def set_view(self, N, cre, cim, width):
        _check(lib().nm_set_view(self._h, int(N), _b(str(cre)), _b(str(cim)), _b(str(width))), "nm_set_view")

    def precompute_at(self, re, im):
        _check(lib().nm_precompute_at(self._h, _b(str(re)), _b(str(im))), "nm_precompute_at")

    def reference_length(self):
        return lib().nm_reference_length(self._h)

    def probe_winner(self):
        return lib().nm_probe_winner(self._h)

    def set_probe_mode(self, mode):
        """PROBE_HOST: findProbe in HP for every probe; PROBE_GPU: probes screened
        by FP64 perturbation on the device, HP only for the near-best ones."""
        _check(lib().nm_set_probe_mode(self._h, int(mode)), "nm_set_probe_mode")

    def probe_candidates(self):
        return lib().nm_probe_candidates(self._h)

    def reference(self):
        n = self.reference_length()
        X = np.zeros(2 * n)
        A = np.zeros(2 * n)
        B = np.zeros(2 * n)
        Cc = np.zeros(2 * n)
        _check(lib().nm_get_reference(self._h, X.ctypes.data, A.ctypes.data, B.ctypes.data, Cc.ctypes.data),
               "nm_get_reference")
        return X, A, B, Cc

    def render(self, block_k=0, waves_per_simd=0, repeats=1, copy_back=True):
        st = RenderStats()
        _check(lib().nm_render(self._h, block_k, waves_per_simd, repeats, 1 if copy_back else 0, C.byref(st)),
               "nm_render")
        return st.as_dict()

    def tile_features(self, tile=64):
        n = C.c_int(0)
        _check(lib().nm_tile_features(self._h, tile, None, 0, C.byref(n)), "nm_tile_features")
        x = np.zeros((n.value, SKIPNET_INPUTS), dtype=np.float32)
        _check(lib().nm_tile_features(self._h, tile, x.ctypes.data, n.value, C.byref(n)), "nm_tile_features")
        return x

    def tile_skips(self, tile=64, mode=0, net=None):
        """First unstable index per tile probe: mode 0 scan, 1 tables, 2 learned."""
        nt = ((self.rows() + tile - 1) // tile) * ((self.cols() + tile - 1) // tile)
        ua = np.zeros(nt, dtype=np.int32)
        how = np.zeros(nt, dtype=np.int32)
        st = TileStats()
        _check(lib().nm_tile_skips(self._h, tile, mode, net._h if net else None, ua.ctypes.data, how.ctypes.data, nt,
                                   C.byref(st)), "nm_tile_skips")
        return ua, how, st.as_dict()

    def reference_point(self):
        re = C.create_string_buffer(8192)
        im = C.create_string_buffer(8192)
        _check(lib().nm_reference_point(self._h, re, im, 8192), "nm_reference_point")
        return re.value.decode(), im.value.decode()

    def tile_targets(self, ua):
        ua = np.ascontiguousarray(ua, dtype=np.int32)
        t = np.zeros(len(ua), dtype=np.float32)
        _check(lib().nm_tile_targets(self._h, ua.ctypes.data, len(ua), t.ctypes.data), "nm_tile_targets")
        return t

    def set_skip_mode(self, mode, tile=64, net=None):
        self._skipnet = net  # keeps the network alive while this view uses it
        _check(lib().nm_set_skip_mode(self._h, mode, tile, net._h if net else None), "nm_set_skip_mode")

    def tile_stats(self):
        st = TileStats()
        _check(lib().nm_get_tile_stats(self._h, C.byref(st)), "nm_get_tile_stats")
        return st.as_dict()

    def set_launch(self, block_k=0, waves_per_simd=0, tile=0, graph=False, pre_iter=0, chunk_run=0):
        l = Launch(block_k, waves_per_simd, tile, 1 if graph else 0, pre_iter, chunk_run, 0.0)
        _check(lib().nm_set_launch(self._h, C.byref(l)), "nm_set_launch")

    def get_launch(self):
        l = Launch()
        _check(lib().nm_get_launch(self._h, C.byref(l)), "nm_get_launch")
        return l.as_dict()

    def autotune_frames(self, widths):
        """Sweep the launch settings on the frame set; returns (best, table)."""
        n = len(widths)
        arr = (C.c_char_p * n)(*[_b(str(w)) for w in widths])
        best = Launch()
        table = (Launch * 64)()
        nt = C.c_int(0)
        _check(lib().nm_autotune_frames(self._h, arr, n, C.byref(best), table, 64, C.byref(nt)),
               "nm_autotune_frames")
        return best.as_dict(), [table[i].as_dict() for i in range(nt.value)]

    def render_frames(self, widths, block_k=0, waves_per_simd=0, repeats=1, copy_back=True):
        """Frames sharing this view's centre and reference; copy_back=False leaves
        the grids on the device and returns (None, stats)."""
        n = len(widths)
        arr = (C.c_char_p * n)(*[_b(str(w)) for w in widths])
        out = np.zeros(n * self.rows() * self.cols(), dtype=ESCAPE_DTYPE) if copy_back else None
        st = RenderStats()
        _check(lib().nm_render_frames(self._h, arr, n, block_k, waves_per_simd, repeats,
                                      out.ctypes.data if copy_back else None, C.byref(st)), "nm_render_frames")
        return (out.reshape(n, self.rows(), self.cols()) if copy_back else None), st.as_dict()
