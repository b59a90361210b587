"""Film masks from PNG images, as the reference reads them.

helper_functions.cpp:173-219 (load_mask, FreeImage): grey = (uint8)(.3 r +
.59 g + .11 b) per pixel, scanline y = 0 is the BOTTOM row of the image
(FreeImage stores bitmaps bottom-up); main.cpp:190-193 makes a cell magnetic
where grey == 0.  This reader decodes non-interlaced 8-bit PNGs (grey, grey +
alpha, RGB, RGBA, palette) with zlib and the five PNG row filters.
"""
import struct
import zlib

import numpy as np


def _paeth(a, b, c):
    p = a + b - c
    pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
    if pa <= pb and pa <= pc:
        return a
    return b if pb <= pc else c


def read_png(path):
    """Return an (height, width, channels) uint8 array in file row order (top first)."""
    data = open(path, "rb").read()
    if data[:8] != b"\x89PNG\r\n\x1a\n":
        raise ValueError(f"{path}: not a PNG file")
    pos, idat, palette = 8, b"", None
    w = h = depth = ctype = interlace = None
    while pos < len(data):
        ln = struct.unpack(">I", data[pos:pos + 4])[0]
        typ = data[pos + 4:pos + 8]
        body = data[pos + 8:pos + 8 + ln]
        pos += 12 + ln
        if typ == b"IHDR":
            w, h, depth, ctype, _, _, interlace = struct.unpack(">IIBBBBB", body)
        elif typ == b"PLTE":
            palette = np.frombuffer(body, np.uint8).reshape(-1, 3)
        elif typ == b"IDAT":
            idat += body
        elif typ == b"IEND":
            break
    if depth != 8 or interlace != 0:
        raise ValueError(f"{path}: only 8-bit non-interlaced PNGs are supported")
    ch = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}[ctype]
    raw = zlib.decompress(idat)
    stride = w * ch
    out = np.zeros((h, stride), np.uint8)
    prev = np.zeros(stride, np.int32)
    for r in range(h):
        f = raw[r * (stride + 1)]
        line = np.frombuffer(raw[r * (stride + 1) + 1:(r + 1) * (stride + 1)], np.uint8).astype(np.int32)
        cur = np.zeros(stride, np.int32)
        for i in range(stride):
            a = cur[i - ch] if i >= ch else 0
            b = prev[i]
            c = prev[i - ch] if i >= ch else 0
            x = line[i]
            if f == 0:
                v = x
            elif f == 1:
                v = x + a
            elif f == 2:
                v = x + b
            elif f == 3:
                v = x + (a + b) // 2
            elif f == 4:
                v = x + _paeth(a, b, c)
            else:
                raise ValueError(f"{path}: bad filter {f}")
            cur[i] = v & 0xFF
        out[r] = cur
        prev = cur
    img = out.reshape(h, w, ch)
    if ctype == 3:
        img = palette[img[:, :, 0]]
    return img


def load_mask(path):
    """(xdim, ydim, grey[y][x]) with y = 0 the bottom image row, as load_mask."""
    img = read_png(path)
    h, w = img.shape[:2]
    if img.shape[2] >= 3:
        r, g, b = (img[:, :, i].astype(np.float64) for i in range(3))
        grey = (.3 * r + .59 * g + .11 * b).astype(np.uint8)
    else:
        grey = img[:, :, 0]
    return w, h, grey[::-1].copy()


def material_from_png(path):
    """2D material flags (1 = magnetic where grey == 0), reference orientation."""
    w, h, grey = load_mask(path)
    return (grey == 0).astype(np.uint8)
