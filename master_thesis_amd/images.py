"""Image files and the image-difference metric of the reference's tooling.

- to_rgb8: clamp to [0, 1], then (char)(v * 255.0f) (software_renderer/
  graphic/graphics_engine.cpp:90-95).
- PPM P6 (GraphicsEngine::saveImage, graphics_engine.cpp:32-40) and PNG
  (RGB, 8 bit, no interlace; what master.py converts the renderers' PPMs to,
  master.py:251-256,301-315), written with the standard library only.
- compare: master.py's compare() (master.py:278-299): per pixel the mean over
  the channels of ((|a - b| / 255)^2 * 255), summed over the image and
  divided by width * height; the difference image holds (3 e, 3 e div 2, 3 e)
  per pixel with e = int(per-pixel error), clipped to 255 as PIL's putdata
  stores it.
"""
from __future__ import annotations

import struct
import zlib
from pathlib import Path

import numpy as np


def to_rgb8(rgb: np.ndarray) -> np.ndarray:
    v = np.clip(np.asarray(rgb, np.float32), np.float32(0.0), np.float32(1.0))
    return (v * np.float32(255.0)).astype(np.uint8)


def write_ppm(path, img8: np.ndarray) -> None:
    img8 = np.ascontiguousarray(img8, np.uint8)
    h, w = img8.shape[:2]
    Path(path).write_bytes(b"P6\n%d\n%d\n255\n" % (w, h) + img8.tobytes())


def read_ppm(path) -> np.ndarray:
    """Binary PPM (P6, maxval 255; whitespace-separated header, comments
    allowed) -> (H, W, 3) uint8."""
    data = Path(path).read_bytes()
    fields, pos = [], 0
    while len(fields) < 4:
        while data[pos:pos + 1].isspace():
            pos += 1
        if data[pos:pos + 1] == b"#":
            pos = data.index(b"\n", pos) + 1
            continue
        start = pos
        while not data[pos:pos + 1].isspace():
            pos += 1
        fields.append(data[start:pos])
    if fields[0] != b"P6" or int(fields[3]) != 255:
        raise ValueError(f"{path}: not a binary 8-bit PPM")
    w, h = int(fields[1]), int(fields[2])
    pos += 1  # the single whitespace after maxval
    return np.frombuffer(data[pos:pos + w * h * 3], np.uint8).reshape(h, w, 3).copy()


def _chunk(tag: bytes, body: bytes) -> bytes:
    return struct.pack(">I", len(body)) + tag + body + struct.pack(">I", zlib.crc32(tag + body) & 0xFFFFFFFF)


def write_png(path, img8: np.ndarray) -> None:
    """RGB 8-bit PNG: IHDR, one zlib IDAT of filter-0 scanlines, IEND."""
    img8 = np.ascontiguousarray(img8, np.uint8)
    h, w = img8.shape[:2]
    raw = np.concatenate([np.zeros((h, 1), np.uint8), img8.reshape(h, w * 3)], axis=1).tobytes()
    png = (b"\x89PNG\r\n\x1a\n" + _chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, 2, 0, 0, 0)) +
           _chunk(b"IDAT", zlib.compress(raw, 6)) + _chunk(b"IEND", b""))
    Path(path).write_bytes(png)


def ppm_to_png(ppm_path, png_path) -> None:
    write_png(png_path, read_ppm(ppm_path))


def compare(img1: np.ndarray, img2: np.ndarray, diff_path=None, size=None):
    """master.py's compare(): returns (metric, difference image (H, W, 3)
    uint8); writes the difference image as PNG when diff_path is given.  size
    = (width, height) of the normalisation (the images' own by default; the
    reference passes 1920x1080)."""
    a = np.asarray(img1)[..., :3].astype(np.int64)
    b = np.asarray(img2)[..., :3].astype(np.int64)
    h = min(a.shape[0], b.shape[0])
    w = min(a.shape[1], b.shape[1])  # zip() stops at the shorter row / image
    a, b = a[:h, :w], b[:h, :w]
    e = np.abs(a - b).astype(np.float64) / 255.0
    e = e * e
    e = e * 255.0
    # float(sum of three) / 3.0, channels added in order
    err = ((e[..., 0] + e[..., 1]) + e[..., 2]) / 3.0
    # total_error += error, pixel by pixel in row-major order
    total = float(np.cumsum(err.reshape(-1))[-1]) if err.size else 0.0
    v = err.astype(np.int64) * 3
    diff = np.clip(np.stack([v, v // 2, v], axis=-1), 0, 255).astype(np.uint8)
    W, H = size if size is not None else (w, h)
    if diff_path is not None:
        write_png(diff_path, diff)
    return total / float(W * H), diff
