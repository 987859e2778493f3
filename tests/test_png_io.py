"""The reference's on-disk image formats (util/ImageHelpers.java), host only.

* cu_palette_load_or_default reads the first row of a PNG the way
  loadColorPaletteFromFile does (getRGB then fromBGRAtoRGBA): the reference's own
  palette files (tests/golden/palettes/, first rows decoded by PIL into
  palettes_argb.json) must load exactly; PNGs of every colour type PIL writes,
  and an Adam7-interlaced one written here, must load as PIL decodes them.
  Grey and 16-bit images are compared with PIL's conversion (grey replicated,
  high byte kept); Java's colour management for those is parity unpinned.
* a missing or empty path, or a file that is not a readable PNG (bad signature,
  truncated image data, CRC mismatch, unknown critical chunk), gives the
  1536-entry default palette; a missing IEND is accepted, as ImageIO does.
* cu_image_save_png writes an 8-bit RGB PNG that PIL reads back as the low three
  bytes of each RGBA word, creating missing parent directories.
"""
import json
import os
import struct
import zlib

import numpy as np
import pytest
from PIL import Image

from conftest import GOLDEN


def pil_first_row_argb(path):
    im = Image.open(path).convert("RGBA")
    px = [im.getpixel((x, 0)) for x in range(im.size[0])]
    return np.array([(a << 24) | (r << 16) | (g << 8) | b for r, g, b, a in px], dtype=np.uint32)


def test_reference_palettes_load_exactly(cu):
    with open(os.path.join(GOLDEN, "palettes_argb.json")) as f:
        want = json.load(f)
    assert len(want) == 8
    for name, argb in want.items():
        pal, loaded = cu.load_palette(os.path.join(GOLDEN, "palettes", name))
        assert loaded, name
        assert np.array_equal(pal, cu.palette_from_argb(np.array(argb, dtype=np.uint32))), name


def test_default_palette_png_equals_generated_default(cu):
    pal, loaded = cu.load_palette(os.path.join(GOLDEN, "palettes", "defaultPalette.png"))
    assert loaded and pal.size == 1536
    assert np.array_equal(pal, cu.default_palette())


@pytest.mark.parametrize("mode", ["RGB", "RGBA", "P", "PA", "L", "LA", "1", "I;16"])
def test_colour_types_match_pil(cu, tmp_path, mode):
    rng = np.random.default_rng(len(mode))
    w, h = 37, 3
    path = str(tmp_path / f"p_{mode.replace(';', '')}.png")
    if mode in ("RGB", "RGBA"):
        Image.fromarray(rng.integers(0, 256, (h, w, len(mode)), dtype=np.uint8), mode).save(path)
    elif mode == "P":
        im = Image.fromarray(rng.integers(0, 40, (h, w), dtype=np.uint8), "P")
        im.putpalette(list(rng.integers(0, 256, 40 * 3, dtype=np.uint8)))
        im.save(path, transparency=5)
    elif mode == "PA":
        im = Image.fromarray(rng.integers(0, 16, (h, w), dtype=np.uint8), "P")
        im.putpalette(list(rng.integers(0, 256, 16 * 3, dtype=np.uint8)))
        im.save(path, bits=4)
    elif mode in ("L", "LA"):
        Image.fromarray(rng.integers(0, 256, (h, w, len(mode)), dtype=np.uint8).squeeze(), mode).save(path)
    elif mode == "1":
        Image.fromarray(rng.integers(0, 2, (h, w), dtype=np.uint8) * 255, "L").convert("1").save(path)
    else:
        grey16 = rng.integers(0, 65536, (h, w), dtype=np.uint16)
        Image.fromarray(grey16).save(path)
    pal, loaded = cu.load_palette(path)
    assert loaded
    if mode == "I;16":  # PIL clips 16-bit grey when converting; the loader keeps the high byte
        g = (grey16[0] >> 8).astype(np.uint32)
        want = 0xFF000000 | (g << 16) | (g << 8) | g
    else:
        want = pil_first_row_argb(path)
    assert np.array_equal(pal, cu.palette_from_argb(want))


def png_chunk(kind, data):
    return struct.pack(">I", len(data)) + kind + data + struct.pack(">I", zlib.crc32(kind + data))


@pytest.mark.parametrize("w,h", [(21, 11), (1, 1), (3, 1), (9, 2)])
def test_adam7_first_row(cu, tmp_path, w, h):
    """An interlaced RGB PNG (passes written here, sub filter on every line)."""
    rng = np.random.default_rng(7)
    img = rng.integers(0, 256, (h, w, 3), dtype=np.uint8)
    passes = [(0, 0, 8, 8), (4, 0, 8, 8), (0, 4, 4, 8), (2, 0, 4, 4), (0, 2, 2, 4), (1, 0, 2, 2),
              (0, 1, 1, 2)]
    raw = b""
    for x0, y0, dx, dy in passes:
        sub = img[y0::dy, x0::dx]
        if sub.size == 0:
            continue
        for line in sub:
            flat = line.reshape(-1).astype(np.int32)
            filt = flat.copy()
            filt[3:] = (flat[3:] - flat[:-3]) % 256  # filter 1 (sub), bpp 3
            raw += b"\x01" + filt.astype(np.uint8).tobytes()
    png = (b"\x89PNG\r\n\x1a\n" + png_chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, 2, 0, 0, 1))
           + png_chunk(b"IDAT", zlib.compress(raw)) + png_chunk(b"IEND", b""))
    path = str(tmp_path / "adam7.png")
    with open(path, "wb") as f:
        f.write(png)
    want = pil_first_row_argb(path)
    row = img[0].astype(np.uint32)
    assert np.array_equal(want, 0xFF000000 | (row[:, 0] << 16) | (row[:, 1] << 8) | row[:, 2])
    pal, loaded = cu.load_palette(path)
    assert loaded and np.array_equal(pal, cu.palette_from_argb(want))


def test_fallback_to_default(cu, tmp_path):
    default = cu.default_palette()
    bad = tmp_path / "bad.png"
    good = open(os.path.join(GOLDEN, "palettes", "green.png"), "rb").read()
    for content in (b"not a png", good[:40], good[:45] + bytes([good[45] ^ 0xFF]) + good[46:],
                    good.replace(b"IHDR", b"IHDX")):
        bad.write_bytes(content)
        pal, loaded = cu.load_palette(str(bad))
        assert not loaded and np.array_equal(pal, default)
    bad.write_bytes(good[:-12])  # no IEND: the image data is complete, as ImageIO reads it
    assert cu.load_palette(str(bad))[1]
    for path in (None, "", str(tmp_path / "missing.png")):
        pal, loaded = cu.load_palette(path)
        assert not loaded and np.array_equal(pal, default)


def test_palette_capacity_is_checked(cu):
    import ctypes as C
    out = np.zeros(100, dtype=np.uint32)
    n = C.c_size_t()
    with pytest.raises(cu.IllegalArgumentError, match="512 entries"):
        cu.abi.call("cu_palette_load_or_default", os.path.join(GOLDEN, "palettes", "green.png").encode(),
                    out.ctypes.data_as(C.POINTER(C.c_uint32)), C.c_size_t(100), C.byref(n), None)


def test_save_png_round_trip(cu, tmp_path):
    rng = np.random.default_rng(3)
    rgba = rng.integers(0, 2 ** 32, (23, 41), dtype=np.uint64).astype(np.uint32)
    path = str(tmp_path / "a" / "b" / "frame.png")
    cu.save_png(path, rgba)
    im = Image.open(path)
    assert im.mode == "RGB" and im.size == (41, 23)
    got = np.asarray(im).astype(np.uint32)
    want = np.stack([rgba & 0xFF, (rgba >> 8) & 0xFF, (rgba >> 16) & 0xFF], axis=-1)
    assert np.array_equal(got, want)
    # and the loader reads its first row back (alpha 255)
    pal, loaded = cu.load_palette(path)
    assert loaded and np.array_equal(pal, (rgba[0] & 0xFFFFFF) | 0xFF000000)
    with pytest.raises(cu.IllegalArgumentError):
        cu.save_png(str(tmp_path / "x.png"), np.zeros((0, 4), dtype=np.uint32))


@pytest.mark.parametrize("w,h,interlace", [(1 << 25, 1, 0), (64, (1 << 31) - 1, 1)])
def test_oversized_headers_fall_back(cu, tmp_path, w, h, interlace):
    """A header asking for a huge first row, or an interlaced image whose first
    row lies past a huge inflate, is refused (default palette) before anything
    of that size is allocated."""
    png = (b"\x89PNG\r\n\x1a\n" + png_chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, 2, 0, 0, interlace))
           + png_chunk(b"IDAT", zlib.compress(b"\x00" * 64)) + png_chunk(b"IEND", b""))
    path = tmp_path / "big.png"
    path.write_bytes(png)
    pal, loaded = cu.load_palette(str(path))
    assert not loaded and np.array_equal(pal, cu.default_palette())
