"""Radiance RGBE (.hdr) reader for environment maps.

The reference loads environment maps with stb_image's float loader
(loaders/hdr_loader.cpp:10-26: `stbi_loadf(name, &w, &h, &ch, 3)`, rows from the
top, then EnvironmentMap(data, w, h)).  stb is a submodule absent here; this
reader follows the published Radiance picture format (Greg Ward, "Real Pixels",
Graphics Gems II, 1991) and stb's documented behaviour for it:

- header: first line `#?RADIANCE` or `#?RGBE`, lines up to an empty line, which
  must name `FORMAT=32-bit_rle_rgbe`; then the resolution line `-Y <h> +X <w>`
  (the only orientation stb accepts);
- scanlines: flat 4-byte RGBE pixels, or, for 8 <= width < 32768 and a scanline
  that opens with the bytes 2, 2 and a width below 32768, new-style run-length
  encoding: each of the four channels in turn, as runs (count > 128: one byte
  repeated count - 128 times) and literals (count <= 128: count bytes);
- conversion: a pixel with exponent byte e != 0 is (r, g, b) * 2^(e - 136) in
  float32 (exact: an 8-bit mantissa times a power of two), e = 0 is black.
"""
from __future__ import annotations

import numpy as np


class HdrError(ValueError):
    pass


def _header(buf: bytes):
    pos = 0

    def line():
        nonlocal pos
        end = buf.find(b"\n", pos)
        if end < 0:
            raise HdrError("truncated header")
        s = buf[pos:end]
        pos = end + 1
        return s.decode("latin-1")

    magic = line()
    if magic not in ("#?RADIANCE", "#?RGBE"):
        raise HdrError("not a Radiance HDR file")
    fmt_ok = False
    while True:
        s = line()
        if s == "":
            break
        if s == "FORMAT=32-bit_rle_rgbe":
            fmt_ok = True
    if not fmt_ok:
        raise HdrError("unsupported format (only 32-bit_rle_rgbe)")
    tok = line().split()
    if len(tok) != 4 or tok[0] != "-Y" or tok[2] != "+X":
        raise HdrError("unsupported orientation (only -Y h +X w)")
    h, w = int(tok[1]), int(tok[3])
    if h <= 0 or w <= 0:
        raise HdrError("bad resolution")
    return w, h, pos


def _rle_scanline(buf, pos, w):
    out = np.empty((4, w), np.uint8)
    for c in range(4):
        i = 0
        while i < w:
            if pos >= len(buf):
                raise HdrError("truncated scanline")
            count = buf[pos]
            pos += 1
            if count > 128:
                count -= 128
                if i + count > w or pos >= len(buf):
                    raise HdrError("bad run length")
                out[c, i:i + count] = buf[pos]
                pos += 1
            else:
                if count == 0 or i + count > w or pos + count > len(buf):
                    raise HdrError("bad literal length")
                out[c, i:i + count] = np.frombuffer(buf, np.uint8, count, pos)
                pos += count
            i += count
    return out.T, pos


def rgbe_to_float(rgbe: np.ndarray) -> np.ndarray:
    """(..., 4) uint8 RGBE -> (..., 3) float32 linear radiance."""
    e = rgbe[..., 3].astype(np.int32)
    f = np.where(e != 0, np.ldexp(np.float32(1.0), e - 136), np.float32(0.0)).astype(np.float32)
    return (rgbe[..., :3].astype(np.float32) * f[..., None]).astype(np.float32)


def read_hdr(path: str) -> np.ndarray:
    """(height, width, 3) float32, rows from the top of the picture."""
    with open(path, "rb") as fh:
        buf = fh.read()
    w, h, pos = _header(buf)
    rgbe = np.empty((h, w, 4), np.uint8)
    if w < 8 or w >= 32768:
        need = pos + 4 * w * h
        if need > len(buf):
            raise HdrError("truncated pixel data")
        rgbe[:] = np.frombuffer(buf, np.uint8, 4 * w * h, pos).reshape(h, w, 4)
        return rgbe_to_float(rgbe)
    for y in range(h):
        if pos + 4 > len(buf):
            raise HdrError("truncated scanline")
        c1, c2, len_hi, len_lo = buf[pos], buf[pos + 1], buf[pos + 2], buf[pos + 3]
        if c1 != 2 or c2 != 2 or (len_hi & 0x80):
            # not run-length encoded: a picture whose first scanline does not open with a
            # run-length marker is flat RGBE throughout; a marker missing further down is
            # a corrupt file
            if y != 0:
                raise HdrError("corrupt run-length scanline")
            n = 4 * w * h
            if pos + n > len(buf):
                raise HdrError("truncated pixel data")
            rgbe[:] = np.frombuffer(buf, np.uint8, n, pos).reshape(h, w, 4)
            break
        if (len_hi << 8 | len_lo) != w:
            raise HdrError("scanline width mismatch")
        rgbe[y], pos = _rle_scanline(buf, pos + 4, w)
    return rgbe_to_float(rgbe)
