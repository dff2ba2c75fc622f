"""Image files through libptamd (ImageIO::write / ImageIO::load of
Part_2/src/io/image_io.cpp) against PIL as an independent codec.

Writers: PIL reads our .png/.bmp/.tga back to exactly the bytes of the
reference's mapping clamp(pow(v, 1/2.2), 0, 1) * 255 + 0.5 (image_io.cpp:
119-127), rows flipped (the film's row 0 is the bottom row); our .jpg at
quality 90 decodes in PIL close to those bytes.  Readers: PNGs of every
colour type / bit depth PIL writes (and an Adam7-interlaced one built here)
decode to PIL's pixels exactly; JPEGs (4:4:4, 4:2:2, 4:2:0, grey,
progressive, restart markers) within a few levels (inverse DCT and chroma
upsampling are not bit-specified by the standard)."""
import os
import struct
import zlib

import numpy as np
import pytest

from pathtracingtutorial_amd import imageio

PIL = pytest.importorskip("PIL.Image")

REF_TEX = "/root/reference/Part_2/models/sponza/textures"


def _film(h=37, w=53, seed=0):
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:h, 0:w]
    img = np.zeros((h, w, 4), np.float32)
    img[..., 0] = (x / w) ** 1.5
    img[..., 1] = (y / h) * 1.2  # above 1: clamped
    img[..., 2] = 0.5 + 0.5 * np.sin(x * 0.3) * np.cos(y * 0.2)
    img[..., :3] += rng.normal(0.0, 0.01, (h, w, 3)).astype(np.float32)
    img[0, 0, :3] = [-0.5, np.inf, 0.0]  # negative (NaN power) -> 0, inf -> 255
    img[..., 3] = 1.0
    return img


def _expected_bytes(film):
    with np.errstate(invalid="ignore"):
        g = np.power(film[..., :3].astype(np.float32), np.float32(1.0 / 2.2))
    g = np.where(g >= 0, g, 0.0)
    g = np.minimum(g, 1.0)
    return (g * np.float32(255.0) + np.float32(0.5)).astype(np.uint8)[::-1]  # file rows: top first


@pytest.mark.parametrize("ext", [".png", ".bmp", ".tga"])
def test_lossless_writers_read_back_in_pil(tmp_path, ext):
    film = _film()
    p = str(tmp_path / ("img" + ext))
    imageio.write_image(p, film)
    got = np.asarray(PIL.open(p).convert("RGB"))
    assert got.shape == (37, 53, 3)
    exp = _expected_bytes(film)
    assert np.array_equal(got, exp)
    assert tuple(got[-1, 0]) == (0, 255, 0)  # the film's pixel (0, 0) is the file's bottom-left


def test_jpeg_writer_quality90_in_pil(tmp_path):
    film = _film(96, 128, seed=1)
    p = str(tmp_path / "img.jpg")
    imageio.write_image(p, film)
    im = PIL.open(p)
    assert im.format == "JPEG" and im.size == (128, 96)
    got = np.asarray(im.convert("RGB")).astype(np.float64)
    exp = _expected_bytes(film).astype(np.float64)
    mse = np.mean((got - exp) ** 2)
    psnr = 10 * np.log10(255.0 ** 2 / mse)
    assert psnr > 35.0, psnr


def _to_bytes(lin):
    """Linear values of ImageIO::load back to the file's 8-bit values (rows top first)."""
    return np.round(np.power(np.clip(lin[::-1, :, :3], 0, 1), 1.0 / 2.2) * 255.0)


@pytest.mark.parametrize("mode", ["RGB", "RGBA", "L", "LA", "P", "I;16", "1"])
def test_png_reader_matches_pil(tmp_path, mode):
    rng = np.random.default_rng(3)
    rgb = rng.integers(0, 256, (29, 41, 3), dtype=np.uint8)
    base = PIL.fromarray(rgb, "RGB")
    if mode == "P":
        im = base.quantize(colors=50)
        im.info["transparency"] = 3
    elif mode == "I;16":
        im = PIL.fromarray((rng.integers(0, 65536, (29, 41))).astype(np.uint16))
    elif mode == "1":
        im = base.convert("L").point(lambda v: 255 if v > 128 else 0).convert("1")
    else:
        im = base.convert(mode)
    p = str(tmp_path / "x.png")
    im.save(p)
    ours = imageio.read_image(p)
    assert ours.shape == (29, 41, 4) and np.all(ours[..., 3] == 1.0)
    ref = np.asarray(PIL.open(p).convert("RGB")).astype(np.float64)
    if mode == "I;16":  # PIL's RGB view of 16-bit grey saturates; compare with the high byte
        ref = np.repeat((np.asarray(PIL.open(p)).astype(np.uint32) >> 8)[..., None], 3, axis=2).astype(np.float64)
    assert np.array_equal(_to_bytes(ours), ref)
    exp_lin = np.power((ref[::-1] / 255.0).astype(np.float32), np.float32(2.2))
    assert np.allclose(ours[..., :3], exp_lin, rtol=2e-6, atol=1e-7)


def _adam7_png(rgb):
    """An interlaced RGB PNG (filter 'Sub' on every row) built here from the spec."""
    h, w, _ = rgb.shape
    passes = [(0, 0, 8, 8), (4, 0, 8, 8), (0, 4, 4, 8), (2, 0, 4, 4), (0, 2, 2, 4), (1, 0, 2, 2), (0, 1, 1, 2)]
    raw = bytearray()
    for x0, y0, dx, dy in passes:
        sub = rgb[y0::dy, x0::dx]
        if sub.size == 0:
            continue
        for row in sub:
            flat = row.reshape(-1).astype(np.int32)
            filt = flat.copy()
            filt[3:] = (flat[3:] - flat[:-3]) & 0xFF
            raw += bytes([1]) + bytes(filt.astype(np.uint8))

    def chunk(t, d):
        return struct.pack(">I", len(d)) + t + d + struct.pack(">I", zlib.crc32(t + d) & 0xFFFFFFFF)

    return (b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, 2, 0, 0, 1)) +
            chunk(b"IDAT", zlib.compress(bytes(raw), 9)) + chunk(b"IEND", b""))


def test_png_reader_adam7(tmp_path):
    rng = np.random.default_rng(4)
    rgb = rng.integers(0, 256, (23, 19, 3), dtype=np.uint8)
    p = str(tmp_path / "i.png")
    open(p, "wb").write(_adam7_png(rgb))
    assert np.array_equal(np.asarray(PIL.open(p).convert("RGB")), rgb)  # the file is valid
    assert np.array_equal(_to_bytes(imageio.read_image(p)), rgb)


def test_png_reader_rejects_corruption(tmp_path):
    p = str(tmp_path / "c.png")
    PIL.fromarray(np.zeros((8, 8, 3), np.uint8)).save(p)
    data = bytearray(open(p, "rb").read())
    data[40] ^= 0x55  # inside IDAT: the chunk CRC no longer matches
    open(p, "wb").write(bytes(data))
    import pathtracingtutorial_amd as pt
    with pytest.raises(pt.PtError):
        imageio.read_image(p)


JPEG_CASES = [
    dict(quality=90, subsampling=0),
    dict(quality=75, subsampling=1),
    dict(quality=85, subsampling=2),
    dict(quality=80, subsampling=2, progressive=True),
    dict(quality=95, subsampling=0, progressive=True),
    dict(quality=80, subsampling=2, restart_marker_blocks=3),
    dict(quality=85, grey=True),
]


@pytest.mark.parametrize("case", JPEG_CASES, ids=lambda c: "-".join(f"{k}{v}" for k, v in c.items()))
def test_jpeg_reader_close_to_pil(tmp_path, case):
    case = dict(case)
    film = _film(70, 93, seed=5)
    rgb = _expected_bytes(film)
    im = PIL.fromarray(rgb, "RGB")
    if case.pop("grey", False):
        im = im.convert("L")
    p = str(tmp_path / "x.jpg")
    try:
        im.save(p, **case)
    except (TypeError, ValueError, OSError) as e:
        pytest.skip(f"PIL cannot write this JPEG variant: {e}")
    ours = _to_bytes(imageio.read_image(p))
    ref = np.asarray(PIL.open(p).convert("RGB")).astype(np.float64)
    d = np.abs(ours - ref)
    assert d.max() <= 6 and d.mean() < 0.6, (d.max(), d.mean())


@pytest.mark.skipif(not os.path.isdir(REF_TEX), reason="reference textures not present")
def test_png_reader_on_reference_textures():
    names = sorted(f for f in os.listdir(REF_TEX) if f.endswith(".png"))[:6]
    assert names
    for n in names:
        p = os.path.join(REF_TEX, n)
        ours = _to_bytes(imageio.read_image(p))
        ref = np.asarray(PIL.open(p).convert("RGB")).astype(np.float64)
        assert np.array_equal(ours, ref), n


def test_unsupported_extension(tmp_path):
    import pathtracingtutorial_amd as pt
    with pytest.raises(pt.PtError):
        imageio.write_image(str(tmp_path / "x.webp"), _film())
