"""PNG decoding to RGBA8, as the reference's scene parser loads image textures.

The reference decodes a texture file with lodepng (lib/lodepng.cpp, vendored in
its tree) into 8-bit RGBA (pbrt_parser.cpp:736-766, ``LCT_RGBA``).  This module
restates that decode for the PNG forms lodepng reads:

* chunk walk with CRC checks (IHDR, PLTE, tRNS, IDAT, IEND; unknown ancillary
  chunks skipped, unknown critical chunks an error);
* zlib inflate of the concatenated IDAT data (Adler-32 checked);
* the five scanline filters (PNG spec, section 9), with Adam7 interlacing;
* lodepng's conversion to RGBA8 (getPixelColorsRGBA8): 16-bit channels keep
  their high byte, grey of 1/2/4 bits becomes value * 255 / (2^bits - 1)
  (integer division), palette indices past the palette become opaque black,
  tRNS colour keys make matching pixels transparent.
"""
from __future__ import annotations

import struct
import zlib
from typing import List, Tuple

import numpy as np

_SIGNATURE = b"\x89PNG\r\n\x1a\n"
# colour type -> (channels, allowed bit depths)
_COLOR_TYPES = {0: (1, (1, 2, 4, 8, 16)), 2: (3, (8, 16)), 3: (1, (1, 2, 4, 8)), 4: (2, (8, 16)), 6: (4, (8, 16))}
# Adam7 passes: (x start, y start, x step, y step)
_ADAM7 = ((0, 0, 8, 8), (4, 0, 8, 8), (0, 4, 4, 8), (2, 0, 4, 4), (0, 2, 2, 4), (1, 0, 2, 2), (0, 1, 1, 2))


class PngError(ValueError):
    """A file lodepng would refuse (or a form this decoder does not read)."""


def _unfilter(data: memoryview, pos: int, width: int, height: int, bpp_bits: int) -> Tuple[np.ndarray, int]:
    """Reverses the scanline filters of one (sub)image starting at data[pos];
    returns (height, stride) raw bytes and the position after it."""
    stride = (width * bpp_bits + 7) // 8
    bpp = max(1, bpp_bits // 8)
    out = np.zeros((height, stride), np.uint8)
    prior = bytearray(stride)
    for y in range(height):
        if pos + 1 + stride > len(data):
            raise PngError("image data is shorter than the image")
        ftype = data[pos]
        line = bytearray(data[pos + 1: pos + 1 + stride])
        pos += 1 + stride
        if ftype == 0:
            pass
        elif ftype == 1:                                        # Sub
            for i in range(bpp, stride):
                line[i] = (line[i] + line[i - bpp]) & 255
        elif ftype == 2:                                        # Up
            a = np.frombuffer(bytes(line), np.uint8).astype(np.uint16) + np.frombuffer(bytes(prior), np.uint8)
            line = bytearray((a & 255).astype(np.uint8).tobytes())
        elif ftype == 3:                                        # Average
            for i in range(stride):
                left = line[i - bpp] if i >= bpp else 0
                line[i] = (line[i] + ((left + prior[i]) >> 1)) & 255
        elif ftype == 4:                                        # Paeth
            for i in range(stride):
                if i >= bpp:
                    a, c = line[i - bpp], prior[i - bpp]
                else:
                    a = c = 0
                b = prior[i]
                p = a + b - c
                pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
                if pa <= pb and pa <= pc:
                    pr = a
                elif pb <= pc:
                    pr = b
                else:
                    pr = c
                line[i] = (line[i] + pr) & 255
        else:
            raise PngError(f"unknown filter type {ftype}")
        out[y] = np.frombuffer(bytes(line), np.uint8)
        prior = line
    return out, pos


def _samples(raw: np.ndarray, width: int, bits: int, per_pixel: int) -> np.ndarray:
    """Unpacks (h, stride) rows into (h, width * per_pixel) samples of `bits`."""
    if bits == 8:
        return raw[:, : width * per_pixel].astype(np.uint32)
    if bits == 16:
        w = raw[:, : 2 * width * per_pixel].astype(np.uint32)
        return (w[:, 0::2] << 8) | w[:, 1::2]
    bitsarr = np.unpackbits(raw, axis=1)[:, : width * per_pixel * bits]
    bitsarr = bitsarr.reshape(raw.shape[0], width * per_pixel, bits).astype(np.uint32)
    weights = (1 << np.arange(bits - 1, -1, -1)).astype(np.uint32)
    return (bitsarr * weights).sum(axis=2).astype(np.uint32)


def _to_rgba8(s: np.ndarray, ctype: int, bits: int, palette: np.ndarray, key) -> np.ndarray:
    """lodepng getPixelColorsRGBA8 for one (h, w * channels) block of samples."""
    h = s.shape[0]
    hi = (lambda v: (v >> 8).astype(np.uint8)) if bits == 16 else (lambda v: v.astype(np.uint8))
    if ctype == 0:
        if bits < 8:
            g = ((s * 255) // ((1 << bits) - 1)).astype(np.uint8)
        else:
            g = hi(s)
        a = np.full(s.shape, 255, np.uint8)
        if key is not None:
            a[s == key[0]] = 0
        return np.stack([g, g, g, a], -1).reshape(h, -1, 4)
    if ctype == 2:
        px = s.reshape(h, -1, 3)
        rgb = hi(px)
        a = np.full(px.shape[:2], 255, np.uint8)
        if key is not None:
            a[(px[..., 0] == key[0]) & (px[..., 1] == key[1]) & (px[..., 2] == key[2])] = 0
        return np.concatenate([rgb, a[..., None]], -1)
    if ctype == 3:
        idx = s
        table = np.zeros((256, 4), np.uint8)
        table[:, 3] = 255                                       # past the palette: opaque black
        table[: len(palette)] = palette
        return table[np.minimum(idx, 255)].reshape(h, -1, 4)
    if ctype == 4:
        px = hi(s.reshape(h, -1, 2))
        return np.stack([px[..., 0], px[..., 0], px[..., 0], px[..., 1]], -1)
    return hi(s.reshape(h, -1, 4))


def decode_png(blob: bytes) -> np.ndarray:
    """PNG bytes -> (H, W, 4) uint8 RGBA, rows top to bottom."""
    if blob[:8] != _SIGNATURE:
        raise PngError("not a PNG file (bad signature)")
    pos = 8
    ihdr = None
    palette = np.zeros((0, 4), np.uint8)
    trns = None
    idat: List[bytes] = []
    ended = False
    while pos + 12 <= len(blob):
        (n,) = struct.unpack(">I", blob[pos: pos + 4])
        ctype = blob[pos + 4: pos + 8]
        body = blob[pos + 8: pos + 8 + n]
        if len(body) != n or pos + 12 + n > len(blob):
            raise PngError("chunk runs past the end of the file")
        (crc,) = struct.unpack(">I", blob[pos + 8 + n: pos + 12 + n])
        if zlib.crc32(ctype + body) & 0xFFFFFFFF != crc:
            raise PngError(f"CRC mismatch in chunk {ctype!r}")
        pos += 12 + n
        if ctype == b"IHDR":
            if n != 13:
                raise PngError("bad IHDR")
            ihdr = struct.unpack(">IIBBBBB", body)
        elif ctype == b"PLTE":
            if n % 3 or n == 0 or n > 768:
                raise PngError("bad palette")
            p = np.frombuffer(body, np.uint8).reshape(-1, 3)
            palette = np.concatenate([p, np.full((len(p), 1), 255, np.uint8)], 1)
        elif ctype == b"tRNS":
            trns = body
        elif ctype == b"IDAT":
            idat.append(body)
        elif ctype == b"IEND":
            ended = True
            break
        elif not (ctype[0] & 0x20):
            raise PngError(f"unknown critical chunk {ctype!r}")
    if ihdr is None or not ended:
        raise PngError("missing IHDR or IEND")
    width, height, bits, color, comp, filt, interlace = ihdr
    if width == 0 or height == 0 or comp != 0 or filt != 0 or interlace > 1:
        raise PngError("unsupported IHDR fields")
    if color not in _COLOR_TYPES or bits not in _COLOR_TYPES[color][1]:
        raise PngError(f"invalid colour type {color} with bit depth {bits}")
    if color == 3 and len(palette) == 0:
        raise PngError("palette image without PLTE")
    key = None
    if trns is not None:
        if color == 3:
            if len(trns) > len(palette):
                raise PngError("tRNS longer than the palette")
            palette = palette.copy()
            palette[: len(trns), 3] = np.frombuffer(trns, np.uint8)
        elif color == 0 and len(trns) == 2:
            key = struct.unpack(">H", trns)
        elif color == 2 and len(trns) == 6:
            key = struct.unpack(">HHH", trns)
        else:
            raise PngError("bad tRNS chunk")
    try:
        data = memoryview(zlib.decompress(b"".join(idat)))
    except zlib.error as e:
        raise PngError(f"image data does not inflate: {e}") from e
    channels = _COLOR_TYPES[color][0]
    bpp_bits = channels * bits
    out = np.zeros((height, width, 4), np.uint8)
    if interlace == 0:
        raw, _ = _unfilter(data, 0, width, height, bpp_bits)
        out[:] = _to_rgba8(_samples(raw, width, bits, channels), color, bits, palette, key)
    else:
        pos = 0
        for x0, y0, dx, dy in _ADAM7:
            pw = (width - x0 + dx - 1) // dx if width > x0 else 0
            ph = (height - y0 + dy - 1) // dy if height > y0 else 0
            if pw == 0 or ph == 0:
                continue
            raw, pos = _unfilter(data, pos, pw, ph, bpp_bits)
            out[y0::dy, x0::dx] = _to_rgba8(_samples(raw, pw, bits, channels), color, bits, palette, key)
    return out


def read_png(path: str) -> np.ndarray:
    """The texture file at `path` as (H, W, 4) uint8 RGBA."""
    with open(path, "rb") as fh:
        return decode_png(fh.read())


def encode_png(rgba: np.ndarray, filters=(0,)) -> bytes:
    """(H, W, 4) uint8 -> 8-bit RGBA PNG, rows filtered with the given filter
    types in turn (0-4).  For writing texture files."""
    a = np.ascontiguousarray(rgba, np.uint8)
    h, w, _ = a.shape
    rows = []
    prior = np.zeros(w * 4, np.int32)
    for y in range(h):
        cur = a[y].reshape(-1).astype(np.int32)
        ft = filters[y % len(filters)]
        left = np.concatenate([np.zeros(4, np.int32), cur[:-4]])
        upleft = np.concatenate([np.zeros(4, np.int32), prior[:-4]])
        if ft == 0:
            f = cur
        elif ft == 1:
            f = cur - left
        elif ft == 2:
            f = cur - prior
        elif ft == 3:
            f = cur - ((left + prior) >> 1)
        else:
            p = left + prior - upleft
            pa, pb, pc = np.abs(p - left), np.abs(p - prior), np.abs(p - upleft)
            pred = np.where((pa <= pb) & (pa <= pc), left, np.where(pb <= pc, prior, upleft))
            f = cur - pred
        rows.append(bytes([ft]) + (f & 255).astype(np.uint8).tobytes())
        prior = cur

    def chunk(t: bytes, body: bytes) -> bytes:
        return struct.pack(">I", len(body)) + t + body + struct.pack(">I", zlib.crc32(t + body) & 0xFFFFFFFF)

    return (_SIGNATURE + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, 6, 0, 0, 0)) +
            chunk(b"IDAT", zlib.compress(b"".join(rows), 6)) + chunk(b"IEND", b""))
