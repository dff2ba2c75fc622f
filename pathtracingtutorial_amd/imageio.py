"""Image files: write_image / read_image go through libptamd (ImageIO::write
and ImageIO::load of Part_2/src/io/image_io.cpp: .hdr .pfm .png .jpg .jpeg
.bmp .tga out, .hdr .png .jpg .jpeg in); rows follow the film convention
(row 0 = bottom image row, flipped on write).  write_hdr / write_pfm /
read_hdr are small numpy codecs of the float formats.
"""
from __future__ import annotations

import ctypes as C

import numpy as np


def _rgbe(rgb: np.ndarray) -> np.ndarray:
    rgb = np.maximum(rgb.astype(np.float32), 0.0)
    m = rgb.max(axis=-1)
    out = np.zeros(rgb.shape[:-1] + (4,), np.uint8)
    ok = m >= 1e-32
    mant, ex = np.frexp(m[ok])
    scale = mant * 256.0 / m[ok]
    out[ok, 0:3] = (rgb[ok] * scale[:, None]).astype(np.uint8)
    out[ok, 3] = (ex + 128).astype(np.uint8)
    return out


def write_hdr(path: str, img: np.ndarray, flip_y: bool = True) -> None:
    h, w = img.shape[:2]
    rows = img[::-1] if flip_y else img
    with open(path, "wb") as f:
        f.write(f"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y {h} +X {w}\n".encode())
        f.write(_rgbe(rows[..., :3]).tobytes())


def write_pfm(path: str, img: np.ndarray) -> None:
    h, w = img.shape[:2]
    with open(path, "wb") as f:
        f.write(f"PF\n{w} {h}\n-1.0\n".encode())
        f.write(np.ascontiguousarray(img[..., :3], dtype="<f4").tobytes())


def write_image(path: str, img: np.ndarray) -> None:
    """ImageIO::write of an (H, W, 3|4) float image whose row 0 is the bottom row."""
    from . import _capi, lib
    h, w = img.shape[:2]
    a = np.ones((h, w, 4), np.float32)
    a[..., :3] = img[..., :3]
    _capi.check(lib().pt_image_write(path.encode(), a.ctypes.data_as(C.POINTER(C.c_float)), w, h))


def read_image(path: str) -> np.ndarray:
    """ImageIO::load: (H, W, 4) float32, linear, alpha 1, row 0 = bottom row."""
    from . import _capi, lib
    w, h = C.c_int32(0), C.c_int32(0)
    _capi.check(lib().pt_image_read(path.encode(), None, C.byref(w), C.byref(h)))
    out = np.empty((h.value, w.value, 4), np.float32)
    _capi.check(lib().pt_image_read(path.encode(), out.ctypes.data_as(C.POINTER(C.c_float)), C.byref(w), C.byref(h)))
    return out


def read_hdr(path: str) -> np.ndarray:
    """RGBE decode (flat or new-style RLE), value = m * 2^(e-136); returns (H, W, 4) float32, alpha 1."""
    data = open(path, "rb").read()
    p = 0

    def line():
        nonlocal p
        e = data.index(b"\n", p)
        s = data[p:e].decode("latin-1")
        p = e + 1
        return s

    magic = line()
    if magic not in ("#?RADIANCE", "#?RGBE"):
        raise ValueError("not a Radiance file")
    while line() != "":
        pass
    parts = line().split()
    h, w = int(parts[1]), int(parts[3])
    px = np.zeros((h, w, 4), np.uint8)
    buf = np.frombuffer(data, np.uint8)
    for y in range(h):
        if 8 <= w < 32768 and buf[p] == 2 and buf[p + 1] == 2 and not (buf[p + 2] & 0x80):
            p += 4
            for ch in range(4):
                x = 0
                while x < w:
                    c = int(buf[p]); p += 1
                    if c > 128:
                        c -= 128
                        px[y, x:x + c, ch] = buf[p]; p += 1
                    else:
                        px[y, x:x + c, ch] = buf[p:p + c]; p += c
                    x += c
        else:
            n = (h - y) * w
            px[y:].reshape(-1, 4)[:] = buf[p:p + 4 * n].reshape(-1, 4)
            p += 4 * n
            break
    e = px[..., 3].astype(np.int32)
    f = np.where(e > 0, np.ldexp(np.float32(1.0), e - 136), 0.0).astype(np.float32)
    out = np.empty((h, w, 4), np.float32)
    out[..., :3] = px[..., :3].astype(np.float32) * f[..., None]
    out[..., 3] = 1.0
    return out
