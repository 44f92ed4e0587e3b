"""Every index the render kernels form, walked on the host: newman_amd/csrc/
index_check.cpp runs each thread of the start launches (plain and series), the
tile scan's two launches, the emit kernel and the iterate kernels' active-list
chunks through the kernels' own index code (csrc/sample_index.hpp) and checks
each index against the length of
the buffer it addresses, that the tiles' one-frame shortcut names each sample's
frame, that every pixel of every frame is written exactly once and that every
active-list entry is taken exactly once.  The shapes include the 2 x 2 frame of
the GPU tests' first launch, every sample order, frames that share tiles, the
configurations' own frame sets and a one-CU launch grid."""
import ctypes as C
import os

import pytest

LIB = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "newman_amd", "libnm_index_check.so")


def _lib():
    if not os.path.exists(LIB):
        from newman_amd import build
        build.build_index_check()
    L = C.CDLL(LIB)
    L.nm_index_check.restype = C.c_longlong
    L.nm_index_check.argtypes = [C.c_int] * 8 + [C.c_char_p, C.c_int, C.POINTER(C.c_ulonglong)]
    return L


def _check(nr, nc, nframes, tile, num_cu=256, mask=0, skip_T=0, plant=0):
    msg = C.create_string_buffer(400)
    n = C.c_ulonglong()
    v = _lib().nm_index_check(nr, nc, nframes, tile, num_cu, mask, skip_T, plant, msg, 400, C.byref(n))
    return v, msg.value.decode(), n.value


CASES = [
    (2, 2, 1, 0, 256, 0, 0),       # the GPU session fixture's first launch (2 x 2, N 4)
    (2, 2, 1, 8, 256, 1, 16),
    (1, 1, 1, 0, 256, 0, 0),
    (1, 1, 3, 0, 1, 0, 0),
    (3, 7, 5, 16, 7, 1, 8),        # frames sharing start tiles, padded tiles, an odd grid
    (33, 1000, 77, 32, 1, 1, 16),
    (480, 640, 4, 0, 256, 0, 0),
    (720, 1280, 3, 64, 256, 0, 64),
    (480, 640, 512, 0, 256, 0, 0),       # the plain batch (640x480 x 512)
    (4320, 7680, 1, 0, 256, 1, 64),      # deep zoom 7680x4320
    (1440, 2560, 1, 0, 256, 1, 64),      # 2560x1440
    (2160, 3840, 1, 0, 256, 1, 64),      # 3840x2160, 64x64 tile skips
    (720, 1280, 120, 64, 256, 0, 0),     # the 120-frame zoom-video batch in 64x64 tiles
    (720, 1280, 16, 128, 256, 0, 0),
    (1000, 1000, 2, 256, 256, 0, 0),
]


@pytest.mark.parametrize("case", CASES)
def test_every_kernel_index_is_in_bounds(case):
    v, msg, n = _check(*case)
    assert v == 0, msg
    assert n > 0


@pytest.mark.parametrize("plant", [1, 2, 3, 4, 5, 6])
def test_the_checker_finds_planted_mistakes(plant):
    """Negative controls of the checker itself: a column-table offset one entry
    on, frames one sample longer than their sample order, plain chunks
    interleaved with one chunk too few, scatter positions that count a lane's
    own bit, a tile scan launched with one group too few, and a sparse-tile
    scatter that selects the next set bit are each reported."""
    v, msg, _ = _check(3, 7, 5, 16, 7, 1, 8, plant=plant)
    assert v > 0 and msg
