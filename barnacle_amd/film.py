"""Film output: Film.PostProcess and Film.Save (Base/Film.fs:21-30, 55-66).

The tone curve, clamp and byte conversion run in the library
(bn_film_postprocess); this module writes the results.  Film.Save goes
through ImageSharp, which picks the format from the file name; here `.png`
gets 8-bit RGBA (what ImageSharp writes for an Image<Rgba32>) and `.pfm` the
linear float film, lossless, for comparisons.  Rows are written in
Film.Pixels order, top row first (the film stores camera row y at H - y - 1).
"""
from __future__ import annotations

import ctypes as C
import struct
import zlib

import numpy as np

from . import native

TONE = {"identity": 0, "aces": 1, "gamma": 2}


def postprocess(film, tone_mapping: str):
    """(tone-mapped, clamped float RGB (n, 3), 8-bit RGBA (n, 4)) of a film
    in Film.Pixels order."""
    lib = native.load()
    f = np.ascontiguousarray(film, np.float32).reshape(-1, 3)
    rgb = np.zeros_like(f)
    rgba = np.zeros((len(f), 4), np.uint8)
    if tone_mapping not in TONE:
        raise ValueError(f"Unknown tone mapping type: {tone_mapping}")
    native.check(lib.bn_film_postprocess(f.ctypes.data_as(C.POINTER(C.c_float)), len(f), TONE[tone_mapping],
                                         rgb.ctypes.data_as(C.POINTER(C.c_float)),
                                         rgba.ctypes.data_as(C.POINTER(C.c_uint8))))
    return rgb, rgba


def _chunk(tag: bytes, data: bytes) -> bytes:
    return struct.pack(">I", len(data)) + tag + data + struct.pack(">I", zlib.crc32(tag + data) & 0xFFFFFFFF)


def write_png(path: str, rgba, width: int, height: int):
    """8-bit RGBA PNG, rows top first, filter type 0."""
    px = np.ascontiguousarray(rgba, np.uint8).reshape(height, width * 4)
    raw = b"".join(b"\x00" + px[y].tobytes() for y in range(height))
    with open(path, "wb") as fh:
        fh.write(b"\x89PNG\r\n\x1a\n")
        fh.write(_chunk(b"IHDR", struct.pack(">IIBBBBB", width, height, 8, 6, 0, 0, 0)))
        fh.write(_chunk(b"IDAT", zlib.compress(raw, 6)))
        fh.write(_chunk(b"IEND", b""))


def read_png_rgba(path: str):
    """Decode a PNG written by write_png (8-bit RGBA, filter 0): (H, W, 4)."""
    with open(path, "rb") as fh:
        data = fh.read()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, idat = 8, b""
    w = h = 0
    while pos < len(data):
        n = struct.unpack(">I", data[pos:pos + 4])[0]
        tag, body = data[pos + 4:pos + 8], data[pos + 8:pos + 8 + n]
        if tag == b"IHDR":
            w, h = struct.unpack(">II", body[:8])
        elif tag == b"IDAT":
            idat += body
        pos += 12 + n
    raw = np.frombuffer(zlib.decompress(idat), np.uint8).reshape(h, 1 + 4 * w)
    assert np.all(raw[:, 0] == 0)
    return raw[:, 1:].reshape(h, w, 4).copy()


def write_pfm(path: str, film, width: int, height: int):
    """Linear float RGB (Portable Float Map, little-endian).  PFM stores the
    bottom row first, so Film.Pixels rows are written in reverse."""
    f = np.ascontiguousarray(film, np.float32).reshape(height, width, 3)[::-1]
    with open(path, "wb") as fh:
        fh.write(f"PF\n{width} {height}\n-1.0\n".encode())
        fh.write(np.ascontiguousarray(f, "<f4").tobytes())


def read_pfm(path: str):
    """(H*W, 3) float32 in Film.Pixels order from a file of write_pfm."""
    with open(path, "rb") as fh:
        assert fh.readline().strip() == b"PF"
        w, h = (int(v) for v in fh.readline().split())
        scale = float(fh.readline())
        a = np.frombuffer(fh.read(), "<f4" if scale < 0 else ">f4").reshape(h, w, 3)
    return np.ascontiguousarray(a[::-1]).reshape(-1, 3).astype(np.float32)


def save(path: str, film, width: int, height: int, tone_mapping: str):
    """Film.Save: `.pfm` -> the linear film; anything else -> tone-mapped PNG."""
    if path.lower().endswith(".pfm"):
        write_pfm(path, film, width, height)
    else:
        _, rgba = postprocess(film, tone_mapping)
        write_png(path, rgba, width, height)
