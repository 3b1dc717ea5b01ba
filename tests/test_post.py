"""CPU pins of the display-chain oracle (oracle/post_oracle.hpp) and the
PNG writer/reader.

Pins, chosen so a dropped term, a wrong index or a swapped operand fails:
  * addressing: every display/tonemap-pass NEAREST fetch at equal size is the
    identity map (all widths 1..4096), so mode -1 (no operator), gamma off,
    FXAA off returns the accumulation bit for bit, alpha included; a 2x canvas
    replicates 2x2 blocks; a 1/2 canvas picks the odd texels;
  * operators against their closed forms in float64 (within f32 rounding):
    exposure (mode 0), Reinhard x/(1+x), Cineon with its zero below 0.004 and
    limit 1, ACES (Hill) with its fixed points 0 -> 0 and large -> 1,
    Uncharted2's white point 1 -> 1; gamma 1/2.2; operators act per channel
    (a permutation of the input channels permutes the output) except ACES,
    whose matrices mix channels (a grey input stays grey only up to the fit);
  * FXAA: a constant image and axis-aligned step edges are unchanged (worked
    out by hand from fxaa.frag: the search direction runs along the edge and
    every tap stays on one side); every output lies within the range of the
    input; a diagonal edge is smoothed (values strictly between the sides);
  * 8-bit conversion: k/255 -> k, round half up, NaN/negative -> 0, >1 -> 255.
"""
import os

import numpy as np
import pytest

import oracle_ref as O


def _rand(h, w, seed=0, scale=4.0):
    rng = np.random.default_rng(seed)
    a = rng.uniform(0, scale, (h, w, 4)).astype(np.float32)
    a[..., 3] = rng.uniform(0.5, 1.0, (h, w)).astype(np.float32)
    return a


def test_identity_addressing():
    for w in list(range(1, 40)) + [127, 640, 1080, 1920, 2160, 3840, 4096]:
        a = _rand(2, w, seed=w)
        f, _ = O.postprocess(a, -1, gamma=False)
        np.testing.assert_array_equal(f, a)
    a = _rand(37, 53, seed=1)
    f, _ = O.postprocess(a, -1, gamma=False)
    np.testing.assert_array_equal(f, a)


def test_display_resample():
    a = _rand(12, 16, seed=2)
    f2, _ = O.postprocess(a, -1, gamma=False, out_size=(32, 24))
    np.testing.assert_array_equal(f2, np.repeat(np.repeat(a, 2, 0), 2, 1))
    fh, _ = O.postprocess(a, -1, gamma=False, out_size=(8, 6))
    np.testing.assert_array_equal(fh, a[1::2, 1::2])


def _tm(x, mode, exposure=1.0, gamma=False):
    a = np.zeros((1, len(x), 4), np.float32)
    a[0, :, 0] = x
    a[0, :, 1] = x
    a[0, :, 2] = x
    a[0, :, 3] = 1
    return O.postprocess(a, mode, exposure=exposure, gamma=gamma)[0][0, :, :3]


def test_operator_closed_forms():
    x = np.float32(np.concatenate([[0.0, 1e-4, 0.004, 0.0041, 0.1, 0.5, 1.0, 2.0, 10.0, 1e3, 1e6]]))
    e = np.float32(0.7)
    # exposure only: one f32 multiply
    np.testing.assert_array_equal(_tm(x, 0, e)[:, 0], x * e)
    # Reinhard
    xe = (x * e).astype(np.float64)
    np.testing.assert_allclose(_tm(x, 1, e)[:, 0], xe / (1 + xe), rtol=3e-7, atol=1e-30)
    # Cineon: x' = max(0, x-0.004), ((x'(6.2x'+.5))/(x'(6.2x'+1.7)+.06))^2.2
    c = _tm(x, 2, 1.0)[:, 0].astype(np.float64)
    xp = np.maximum(0.0, x.astype(np.float64) - 0.004)
    ref = ((xp * (6.2 * xp + 0.5)) / (xp * (6.2 * xp + 1.7) + 0.06)) ** 2.2
    np.testing.assert_allclose(c, ref, rtol=2e-5, atol=1e-9)
    assert c[0] == 0 and c[1] == 0 and c[2] == 0 and abs(c[-1] - 1) < 1e-5
    # ACES (Hill): grey in -> matrices -> fit -> matrices, clamped
    M_in = np.array([[0.59719, 0.35458, 0.04823], [0.07600, 0.90834, 0.01566], [0.02840, 0.13383, 0.83777]])
    M_out = np.array([[1.60475, -0.53108, -0.07367], [-0.10208, 1.10813, -0.00605], [-0.00327, -0.07276, 1.07602]])
    g = _tm(x, 3, 0.6)  # exposure/0.6 = 1
    v = M_in @ np.stack([x, x, x]).astype(np.float64)
    fit = (v * (v + 0.0245786) - 0.000090537) / (v * (0.983729 * v + 0.4329510) + 0.238081)
    ref = np.clip((M_out @ fit).T, 0, 1)
    np.testing.assert_allclose(g, ref, rtol=1e-5, atol=2e-6)
    assert np.all(g[0] == 0) and np.all(g[-1] == 1)
    # Uncharted2 white point
    np.testing.assert_array_equal(_tm(np.float32([1.0]), 4, 1.0), np.ones((1, 3), np.float32))
    u = _tm(x, 4, 1.0)[:, 0]
    assert np.all(np.diff(u) >= 0) and u[0] < 1e-6  # 0 in exact arithmetic; f32 leaves ~3e-8
    # gamma
    gg = _tm(x, 0, 1.0, gamma=True)[:, 0].astype(np.float64)
    np.testing.assert_allclose(gg, x.astype(np.float64) ** (1 / 2.2), rtol=3e-7)


def test_operators_are_per_channel():
    a = _rand(4, 5, seed=3)
    perm = a[..., [2, 0, 1, 3]]
    for mode in (0, 1, 2, 4):
        f, _ = O.postprocess(a, mode, exposure=1.3, gamma=True)
        fp, _ = O.postprocess(perm, mode, exposure=1.3, gamma=True)
        np.testing.assert_array_equal(fp, f[..., [2, 0, 1, 3]])
    for mode in range(-1, 5):
        assert np.all(np.isfinite(O.postprocess(a, mode)[0]))


def test_fxaa_constant_and_axis_edges():
    c = np.full((9, 11, 4), 0.37, np.float32)
    f, _ = O.postprocess(c, 1, fxaa=True)
    g, _ = O.postprocess(c, 1, fxaa=False)
    np.testing.assert_array_equal(f, g)
    for axis in (0, 1):
        a = np.zeros((16, 16, 4), np.float32)
        a[..., 3] = 1
        if axis == 0:
            a[7:] = [0.9, 0.9, 0.9, 1]
        else:
            a[:, 7:] = [0.9, 0.9, 0.9, 1]
        f, _ = O.postprocess(a, -1, gamma=False, fxaa=True)
        np.testing.assert_array_equal(f, a)


def test_fxaa_bounds_and_diagonal():
    a = _rand(24, 24, seed=4, scale=1.0)
    f, _ = O.postprocess(a, -1, gamma=False, fxaa=True)
    assert f[..., :3].min() >= a[..., :3].min() and f[..., :3].max() <= a[..., :3].max()
    np.testing.assert_array_equal(f[..., 3], a[..., 3])
    # a staircase edge of slope 1/3: the taps cross steps, so edge pixels
    # blend (0.25 / 0.5 / 0.75 mixes of the two sides); pixels away from the
    # edge keep their value (a 45-degree staircase is left alone: its search
    # direction runs along the steps)
    y, x = np.mgrid[0:24, 0:24]
    d = np.zeros((24, 24, 4), np.float32)
    d[..., 3] = 1
    d[3 * y < x] = [1, 1, 1, 1]
    f, _ = O.postprocess(d, -1, gamma=False, fxaa=True)
    changed = np.abs(f[..., 0] - d[..., 0]) > 0
    assert changed.sum() > 20
    v = f[..., 0][changed]
    assert np.all((v > 0) & (v < 1))
    assert set(np.unique(v)) <= {0.25, 0.5, 0.75}
    far = np.abs(x - 3 * y) > 12
    np.testing.assert_array_equal(f[far], d[far])


def test_unorm8():
    k = np.arange(256)
    a = np.zeros((1, 256, 4), np.float32)
    a[0, :, 0] = (k / 255.0).astype(np.float32)
    a[0, :, 1] = ((k + 0.5) / 255.0).astype(np.float32)
    a[0, :3, 2] = [np.nan, -1.0, 2.0]
    a[0, :, 3] = 1
    _, u8 = O.postprocess(a, -1, gamma=False)
    np.testing.assert_array_equal(u8[0, :, 0], k)
    assert np.all(u8[0, :255, 1] - k[:255] <= 1) and np.all(u8[0, :255, 1] >= k[:255])
    assert list(u8[0, :3, 2]) == [0, 0, 255]


@pytest.mark.parametrize("channels", [1, 2, 3, 4])
def test_png_round_trip(tmp_path, channels):
    from dspt.image import read_png, write_png
    rng = np.random.default_rng(channels)
    img = rng.integers(0, 256, (13, 17, channels), dtype=np.uint8)
    p = os.path.join(tmp_path, "x.png")
    write_png(p, img)
    back = read_png(p)
    exp = np.full((13, 17, 4), 255, np.uint8)
    if channels in (1, 2):
        exp[..., :3] = img[..., :1]
        if channels == 2:
            exp[..., 3] = img[..., 1]
    else:
        exp[..., :channels] = img
    np.testing.assert_array_equal(back, exp)


def test_png_filters():
    """Rows written with each PNG filter (PNG spec 9.2, encoded here from the
    spec's definitions) decode to the same pixels."""
    import struct
    import zlib
    from dspt.image import read_png
    rng = np.random.default_rng(9)
    h, w, bpp = 6, 9, 4
    img = rng.integers(0, 256, (h, w * bpp), dtype=np.int32)
    rows = []
    prev = np.zeros(w * bpp, np.int32)
    for y in range(h):
        ft = y % 5
        cur = img[y]
        left = np.concatenate([np.zeros(bpp, np.int32), cur[:-bpp]])
        ul = np.concatenate([np.zeros(bpp, np.int32), prev[:-bpp]])
        if ft == 0:
            enc = cur
        elif ft == 1:
            enc = cur - left
        elif ft == 2:
            enc = cur - prev
        elif ft == 3:
            enc = cur - (left + prev) // 2
        else:
            p = left + prev - ul
            pa, pb, pc = np.abs(p - left), np.abs(p - prev), np.abs(p - ul)
            pred = np.where((pa <= pb) & (pa <= pc), left, np.where(pb <= pc, prev, ul))
            enc = cur - pred
        rows.append(bytes([ft]) + (enc & 0xFF).astype(np.uint8).tobytes())
        prev = cur

    def chunk(t, d):
        return struct.pack(">I", len(d)) + t + d + struct.pack(">I", zlib.crc32(t + d) & 0xFFFFFFFF)
    data = b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, 6, 0, 0, 0))
    data += chunk(b"IDAT", zlib.compress(b"".join(rows))) + chunk(b"IEND", b"")
    np.testing.assert_array_equal(read_png(data), img.reshape(h, w, 4).astype(np.uint8))
