"""Frame export: RGBA8 frames of the thickness field and PNG files.

``export_frame`` restates image_export.rs:42-80 (validation and error kinds
of ``ExportError``) and writes the PNG with the standard library (zlib);
``render_frame`` produces the frame on the GPU from a drainage context
(``sf_render_lut_rgba8``).
"""
from __future__ import annotations

import ctypes as C
import struct
import zlib
from typing import Optional, Sequence

import numpy as np

from . import _lib as L


LUT_SHAPE = (64, 256, 4)          # angle-major RGBA8 runtime LUT (interference_lut.rs:9-13)
LUT_BYTES = 64 * 256 * 4


class ExportError(Exception):
    """Kinds: "BufferCreation", "SaveError", "InvalidDimensions" (image_export.rs:6-14)."""

    def __init__(self, kind: str, message: str):
        super().__init__("%s: %s" % (kind, message))
        self.kind = kind


def _chunk(tag: bytes, data: bytes) -> bytes:
    return struct.pack(">I", len(data)) + tag + data + struct.pack(">I", zlib.crc32(tag + data) & 0xFFFFFFFF)


def export_frame(path: str, width: int, height: int, data) -> None:
    """Write RGBA8 pixel data (row-major, width*height*4 bytes) as a PNG.
    ``data`` is bytes or a uint8 array (the reference takes ``&[u8]``); other
    element types are refused rather than cast."""
    if width == 0 or height == 0:
        raise ExportError("InvalidDimensions", "Invalid dimensions: %dx%d" % (width, height))
    if isinstance(data, (bytes, bytearray)):
        buf = bytes(data)
    else:
        arr = np.asarray(data)
        if arr.dtype != np.uint8:
            raise TypeError("export_frame takes RGBA8 bytes (uint8), got %s" % arr.dtype)
        buf = np.ascontiguousarray(arr).tobytes()
    expected = width * height * 4
    if len(buf) != expected:
        raise ExportError("BufferCreation", "Data length %d doesn't match expected %d (%dx%dx4)"
                          % (len(buf), expected, width, height))
    raw = b"".join(b"\x00" + buf[r * width * 4:(r + 1) * width * 4] for r in range(height))
    png = (b"\x89PNG\r\n\x1a\n"
           + _chunk(b"IHDR", struct.pack(">IIBBBBB", width, height, 8, 6, 0, 0, 0))
           + _chunk(b"IDAT", zlib.compress(raw, 6))
           + _chunk(b"IEND", b""))
    try:
        with open(path, "wb") as f:
            f.write(png)
    except OSError as e:
        raise ExportError("SaveError", str(e))


def read_png_rgba(path: str) -> np.ndarray:
    """Decode an 8-bit RGBA PNG written by export_frame (filter type 0 rows)."""
    blob = open(path, "rb").read()
    assert blob[:8] == b"\x89PNG\r\n\x1a\n"
    pos = 8
    width = height = 0
    idat = b""
    while pos < len(blob):
        n = struct.unpack(">I", blob[pos:pos + 4])[0]
        tag = blob[pos + 4:pos + 8]
        data = blob[pos + 8:pos + 8 + n]
        if tag == b"IHDR":
            width, height = struct.unpack(">II", data[:8])
        elif tag == b"IDAT":
            idat += data
        pos += 12 + n
    raw = zlib.decompress(idat)
    rows = [raw[r * (width * 4 + 1) + 1:(r + 1) * (width * 4 + 1)] for r in range(height)]
    return np.frombuffer(b"".join(rows), np.uint8).reshape(height, width, 4)


def render_frame(ctx, camera: Optional[Sequence[float]] = None, refractive_index: float = 1.33,
                 intensity: float = 4.0, lut: Optional[np.ndarray] = None) -> np.ndarray:
    """RGBA8 frame [n_theta][n_phi][4] of a DrainageContext's thickness."""
    cam = None
    if camera is not None:
        cam = L.Camera()
        for k in range(3):
            cam.position[k] = camera[k]
    out = np.empty((ctx.n_theta, ctx.n_phi, 4), np.uint8)
    lu = None
    if lut is not None:
        # sf_render_lut_rgba8 reads the renderer's whole 64 x 256 RGBA8 table
        lu = np.asarray(lut)
        if lu.dtype != np.uint8 or lu.size != LUT_BYTES or (lu.ndim > 1 and lu.shape != LUT_SHAPE):
            raise ValueError("lut must be the 64 x 256 RGBA8 table (uint8, shape %s or %d bytes), got %s %s"
                             % (LUT_SHAPE, LUT_BYTES, lu.dtype, lu.shape))
        lu = np.ascontiguousarray(lu)
    L.check(L.lib().sf_render_lut_rgba8(ctx.handle, L.u8ptr(lu), float(refractive_index), float(intensity),
                                        None if cam is None else C.byref(cam), L.u8ptr(out)))
    return out
