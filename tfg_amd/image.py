"""Frame output: GPU resolve of the FP64 accumulation buffer (mean over the
accumulated frames, ACES filmic curve, 8-bit quantisation; tfg_image_resolve)
and an 8-bit RGB PNG writer.

Replaces the reference's host loop (main.cpp:1070-1071, 1109-1115),
Imagen::filmic (imagen.h:194-215) and guardarPNG (imagen_utils.h:22-74).
The PNG is written with the standard library's zlib (one IDAT, filter type 0
on every row); decoders see the same pixels libpng would have written.
"""
from __future__ import annotations

import struct
import zlib

import numpy as np

from ._lib import check, lib


def resolve(torch, accum, samples: int, tonemap: bool = True, mean: bool = False, stream=None):
    """accum: (H*W, 3) float64 CUDA tensor.  Returns the (H*W, 3) uint8 bytes
    (CUDA tensor) and, with mean=True, also the FP32 mean image."""
    n = accum.shape[0]
    assert accum.dtype == torch.float64 and accum.is_cuda and accum.shape[1] == 3 and accum.is_contiguous()
    rgb8 = torch.empty((n, 3), dtype=torch.uint8, device=accum.device)
    m = torch.empty((n, 3), dtype=torch.float32, device=accum.device) if mean else None
    check(lib().tfg_image_resolve(accum.data_ptr(), n, int(samples), 1 if tonemap else 0, rgb8.data_ptr(),
                                  m.data_ptr() if m is not None else None, stream), "tfg_image_resolve")
    return (rgb8, m) if mean else rgb8


def _chunk(tag: bytes, data: bytes) -> bytes:
    return struct.pack(">I", len(data)) + tag + data + struct.pack(">I", zlib.crc32(tag + data) & 0xFFFFFFFF)


def png_bytes(rgb8: np.ndarray, width: int, height: int) -> bytes:
    """8-bit RGB PNG (colour type 2, no interlace) of a (height*width, 3) or
    (height, width, 3) uint8 array, rows top to bottom."""
    px = np.ascontiguousarray(rgb8, dtype=np.uint8).reshape(height, width * 3)
    raw = np.concatenate([np.zeros((height, 1), np.uint8), px], axis=1).tobytes()   # filter 0 per row
    ihdr = struct.pack(">IIBBBBB", width, height, 8, 2, 0, 0, 0)
    return b"\x89PNG\r\n\x1a\n" + _chunk(b"IHDR", ihdr) + _chunk(b"IDAT", zlib.compress(raw, 6)) + _chunk(b"IEND", b"")


def write_png(path, rgb8: np.ndarray, width: int, height: int) -> None:
    with open(path, "wb") as f:
        f.write(png_bytes(rgb8, width, height))


def read_png_rgb8(data: bytes):
    """Decoder for the PNGs png_bytes writes (8-bit RGB, filter 0): returns
    (width, height, (height*width, 3) uint8).  For tests and tools."""
    assert data[:8] == b"\x89PNG\r\n\x1a\n", "not a PNG"
    pos, idat = 8, b""
    w = h = None
    while pos < len(data):
        (ln,) = struct.unpack(">I", data[pos:pos + 4])
        tag = data[pos + 4:pos + 8]
        body = data[pos + 8:pos + 8 + ln]
        (crc,) = struct.unpack(">I", data[pos + 8 + ln:pos + 12 + ln])
        assert crc == zlib.crc32(tag + body) & 0xFFFFFFFF, f"bad CRC in {tag!r}"
        if tag == b"IHDR":
            w, h, depth, ctype, _, _, interlace = struct.unpack(">IIBBBBB", body)
            assert depth == 8 and ctype == 2 and interlace == 0
        elif tag == b"IDAT":
            idat += body
        pos += 12 + ln
    raw = np.frombuffer(zlib.decompress(idat), np.uint8).reshape(h, 1 + 3 * w)
    assert (raw[:, 0] == 0).all(), "only filter type 0 is supported"
    return w, h, raw[:, 1:].reshape(h * w, 3).copy()
