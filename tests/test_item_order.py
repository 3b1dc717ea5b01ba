"""The camera-ray refill's item order (dev_order.hpp, host build of the
device header): every (frame, pixel) of a task is issued exactly once, the
task's first frame comes first, and the order keeps the locality it exists
for. Samples are seeded by frame and pixel (renderer.frag:129-142,
rng.glsl:23-29), so these properties are what the GPU accumulation parity
relies on; the GPU tests (test_gpu_tiles.py, test_gpu_jobs.py) render the
same mapping against the oracle."""
import ctypes

import numpy as np
import pytest

from test_devhost_bsdf import devhost  # noqa: F401  (fixture)

GROUP = 16

SIZES = [(1, 1, 1), (1, 1, 40), (8, 8, 17), (13, 7, 1), (13, 7, 2), (13, 7, 18), (64, 3, 33), (100, 37, 50),
         (9, 64, 16), (8, 8, 2), (3, 130, 21)]


def _items(lib, rw, rh, nf):
    n = rw * rh * nf
    f, x, y = (np.zeros(n, np.int32) for _ in range(3))
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    assert lib.devhost_task_items(rw, rh, nf, p(f), p(x), p(y)) == 0
    return f, x, y


@pytest.mark.parametrize("rw,rh,nf", SIZES)
def test_every_frame_pixel_once(devhost, rw, rh, nf):  # noqa: F811
    f, x, y = _items(devhost, rw, rh, nf)
    npix = rw * rh
    assert f.min() >= 0 and f.max() < nf
    assert x.min() >= 0 and x.max() < rw and y.min() >= 0 and y.max() < rh
    key = (f.astype(np.int64) * rh + y) * rw + x
    assert np.array_equal(np.sort(key), np.arange(npix * nf))
    # the first frame first, every pixel of it before any other frame
    assert (f[:npix] == 0).all()
    # later frames in groups of GROUP: group index never decreases
    g = (f[npix:] - 1) // GROUP
    assert (np.diff(g) >= 0).all()


@pytest.mark.parametrize("rw,rh,nf", [(16, 16, 40), (24, 8, 18), (13, 7, 18), (64, 3, 33)])
def test_chunks_of_one_frame_and_one_tile(devhost, rw, rh, nf):  # noqa: F811
    """Inside a frame group, each run of 64 items of a full 64-pixel chunk is
    one frame and one chunk of pixels, and the group's frames walk the same
    chunk back to back."""
    f, x, y = _items(devhost, rw, rh, nf)
    npix = rw * rh
    C = npix // 64
    first = x[:npix] + rw * y[:npix]  # region pixel order of frame 0 (every frame uses it)
    pos = np.empty(npix, np.int64)
    pos[first] = np.arange(npix)
    for g0 in range(1, nf, GROUP):
        gf = min(GROUP, nf - g0)
        base = npix + (g0 - 1) * npix
        for c in range(C):
            for k in range(gf):
                s = slice(base + (c * gf + k) * 64, base + (c * gf + k + 1) * 64)
                assert (f[s] == g0 + k).all()
                assert np.array_equal(np.sort(pos[x[s] + rw * y[s]]), np.arange(c * 64, c * 64 + 64))


def test_tiles_of_eight(devhost):  # noqa: F811
    """A region whose sides are multiples of 8: the first 64 items are the
    8 x 8 square at the corner, the next 64 the square to its right."""
    f, x, y = _items(devhost, 32, 16, 1)
    assert set(zip(x[:64], y[:64])) == {(i, j) for i in range(8) for j in range(8)}
    assert set(zip(x[64:128], y[64:128])) == {(i, j) for i in range(8, 16) for j in range(8)}
    assert set(zip(x[256:320], y[256:320])) == {(i, j) for i in range(8) for j in range(8, 16)}
