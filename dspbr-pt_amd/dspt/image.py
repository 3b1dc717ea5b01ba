"""PNG files (8-bit, zlib; no other dependency).

write_png   the display chain's RGBA8 canvas (or any u8 image) to a PNG file
read_png    8-bit PNGs (grey, grey+alpha, RGB, RGBA, palette; non-interlaced
            and Adam7) -> u8[h, w, 4], for material textures

Rows are in file order (top row first); the engine's canvas and
accumulation are bottom row first (GL order), so callers flip with [::-1].
"""
import struct
import zlib

import numpy as np

_PNG_SIG = b"\x89PNG\r\n\x1a\n"


def _chunk(tag, data):
    return struct.pack(">I", len(data)) + tag + data + struct.pack(">I", zlib.crc32(tag + data) & 0xFFFFFFFF)


def write_png(path, img):
    """img: u8[h, w, 4|3|1] (or [h, w]), top row first."""
    a = np.asarray(img, np.uint8)
    if a.ndim == 2:
        a = a[..., None]
    h, w, c = a.shape
    ctype = {1: 0, 2: 4, 3: 2, 4: 6}[c]
    raw = np.concatenate([np.zeros((h, 1), np.uint8), a.reshape(h, w * c)], axis=1).tobytes()  # filter 0 per row
    data = _PNG_SIG + _chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, ctype, 0, 0, 0))
    data += _chunk(b"IDAT", zlib.compress(raw, 6)) + _chunk(b"IEND", b"")
    with open(path, "wb") as f:
        f.write(data)


def _unfilter(raw, h, stride, bpp):
    """PNG filters 0-4 (PNG spec 9.2) over h rows of `stride` bytes."""
    out = np.zeros((h, stride), np.uint8)
    src = np.frombuffer(raw, np.uint8)
    prev = np.zeros(stride, np.int32)
    pos = 0
    for y in range(h):
        ft = src[pos]
        line = src[pos + 1:pos + 1 + stride].astype(np.int32)
        pos += 1 + stride
        if ft == 0:
            cur = line
        elif ft == 2:
            cur = (line + prev) & 0xFF
        elif ft == 1:  # Sub: a running sum along each byte lane of a pixel
            cur = (np.cumsum(line.reshape(-1, bpp), axis=0) & 0xFF).reshape(-1) if stride % bpp == 0 else None
            if cur is None:
                raise ValueError("PNG: row length is not a whole number of pixels")
        elif ft in (3, 4):
            cur = np.zeros(stride, np.int32)
            for x in range(stride):
                a = cur[x - bpp] if x >= bpp else 0
                if ft == 3:
                    p = (a + prev[x]) >> 1
                else:
                    b, cc = prev[x], (prev[x - bpp] if x >= bpp else 0)
                    pa, pb, pc = abs(b - cc), abs(a - cc), abs(a + b - 2 * cc)
                    p = a if (pa <= pb and pa <= pc) else (b if pb <= pc else cc)
                cur[x] = (line[x] + p) & 0xFF
        else:
            raise ValueError(f"PNG: bad filter type {ft}")
        out[y] = cur
        prev = cur
    return out, pos


def read_png(path_or_bytes):
    """8-bit PNG -> u8[h, w, 4] RGBA, top row first."""
    data = path_or_bytes if isinstance(path_or_bytes, (bytes, bytearray)) else open(path_or_bytes, "rb").read()
    if data[:8] != _PNG_SIG:
        raise ValueError("not a PNG file")
    pos, idat, plte, trns = 8, [], None, None
    w = h = depth = ctype = interlace = None
    while pos < len(data):
        n, tag = struct.unpack(">I4s", data[pos:pos + 8])
        body = data[pos + 8:pos + 8 + n]
        pos += 12 + n
        if tag == b"IHDR":
            w, h, depth, ctype, _, _, interlace = struct.unpack(">IIBBBBB", body)
        elif tag == b"PLTE":
            plte = np.frombuffer(body, np.uint8).reshape(-1, 3)
        elif tag == b"tRNS":
            trns = body
        elif tag == b"IDAT":
            idat.append(body)
        elif tag == b"IEND":
            break
    if depth != 8:
        raise ValueError(f"PNG: only 8-bit channels are supported (got {depth})")
    chans = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}[ctype]
    raw = zlib.decompress(b"".join(idat))
    if interlace:
        px = np.zeros((h, w, chans), np.uint8)
        off = 0
        for x0, y0, dx, dy in ((0, 0, 8, 8), (4, 0, 8, 8), (0, 4, 4, 8), (2, 0, 4, 4), (0, 2, 2, 4), (1, 0, 2, 2),
                               (0, 1, 1, 2)):
            pw, ph = (w - x0 + dx - 1) // dx, (h - y0 + dy - 1) // dy
            if pw <= 0 or ph <= 0:
                continue
            rows, used = _unfilter(raw[off:], ph, pw * chans, chans)
            off += used
            px[y0::dy, x0::dx] = rows.reshape(ph, pw, chans)
    else:
        rows, _ = _unfilter(raw, h, w * chans, chans)
        px = rows.reshape(h, w, chans)
    out = np.full((h, w, 4), 255, np.uint8)
    if ctype == 3:
        out[..., :3] = plte[px[..., 0]]
        if trns is not None:
            alpha = np.full(256, 255, np.uint8)
            alpha[:len(trns)] = np.frombuffer(trns, np.uint8)
            out[..., 3] = alpha[px[..., 0]]
    elif ctype in (0, 4):
        out[..., :3] = px[..., :1]
        if ctype == 4:
            out[..., 3] = px[..., 1]
    else:
        out[..., :chans] = px
    return out
