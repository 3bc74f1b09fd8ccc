"""Film output on the host (no GPU): sRGB resolve, PNG and PFM writers.

* resolve: rgb / weight clamped to [0, 1] (NaN -> 1, as CUDA's fminf/fmaxf
  ignore a NaN operand), the sRGB transfer curve (linear segment 12.92 v below
  0.0031308, 1.055 v^(1/2.4) - 0.055 above), times 255 truncated; checked
  against the curve evaluated independently in binary64 away from integer
  boundaries, plus the hand-computed single-precision end point (white -> 254);
* PNG: the file decodes with Python's zlib (stored deflate blocks, Adler-32),
  every chunk CRC verifies, and the pixels come back unchanged, across more
  than one 64 KiB stored block;
* PFM: round trip of rgb / weight with the bottom-to-top row order.
"""
import struct
import zlib

import numpy as np

from cavalry_amd.render import Film, read_pfm


def srgb64(v):
    v = np.clip(v, 0.0, 1.0)
    return np.where(v <= 0.0031308, 12.92 * v, 1.055 * np.power(v, 1 / 2.4) - 0.055)


def test_srgb8_resolve():
    rng = np.random.default_rng(0)
    h, w = 7, 9
    val = rng.uniform(-0.2, 1.3, size=(h, w, 3))
    val[0, 0] = [0.0, 0.5, 0.002]
    wt = rng.integers(1, 20, size=(h, w)).astype(np.float32)
    rgb = (val * wt[..., None]).astype(np.float32)
    out = Film(rgb, wt).to_srgb8()
    exact = srgb64((rgb / wt[..., None]).astype(np.float64)) * 255
    safe = np.abs(exact - np.round(exact)) > 1e-3
    np.testing.assert_array_equal(out[safe], np.floor(exact[safe]).astype(np.uint8))
    assert list(out[0, 0]) == [0, 187, int(12.92 * 0.002 * 255)]


def test_srgb8_special_values():
    rgb = np.array([[[1.0, np.nan, -3.0], [np.inf, 2.0, 0.0]]], np.float32)
    wt = np.array([[1.0, 1.0]], np.float32)
    out = Film(rgb, wt).to_srgb8()
    # white: 1.055f - 0.055f rounds to 0.99999994f, so 255 * it truncates to 254
    assert list(out[0, 0]) == [254, 254, 0]
    assert list(out[0, 1]) == [254, 254, 0]
    # a pixel without samples: 0 / 0 = NaN -> clamped to 1
    z = Film(np.zeros((1, 1, 3), np.float32), np.zeros((1, 1), np.float32)).to_srgb8()
    assert list(z[0, 0]) == [254, 254, 254]


def _read_png(path):
    data = open(path, "rb").read()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, chunks = 8, []
    while pos < len(data):
        n, = struct.unpack(">I", data[pos:pos + 4])
        typ = data[pos + 4:pos + 8]
        body = data[pos + 8:pos + 8 + n]
        crc, = struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])
        assert crc == zlib.crc32(typ + body) & 0xffffffff, typ
        chunks.append((typ, body))
        pos += 12 + n
    assert [c[0] for c in chunks] == [b"IHDR", b"IDAT", b"IEND"]
    w, h, depth, ctype, comp, filt, inter = struct.unpack(">IIBBBBB", chunks[0][1])
    assert (depth, ctype, comp, filt, inter) == (8, 2, 0, 0, 0)
    raw = zlib.decompress(chunks[1][1])
    rows = np.frombuffer(raw, np.uint8).reshape(h, 1 + 3 * w)
    assert (rows[:, 0] == 0).all()
    return rows[:, 1:].reshape(h, w, 3)


def test_png_round_trip(tmp_path):
    rng = np.random.default_rng(1)
    h, w = 150, 200  # 90 KiB of rows: two stored blocks
    val = rng.uniform(0, 1, size=(h, w, 3)).astype(np.float32)
    film = Film(val, np.ones((h, w), np.float32))
    p = str(tmp_path / "f.png")
    film.save_png(p)
    np.testing.assert_array_equal(_read_png(p), film.to_srgb8())


def test_pfm_round_trip(tmp_path):
    rng = np.random.default_rng(2)
    h, w = 5, 4
    rgb = rng.uniform(0, 10, size=(h, w, 3)).astype(np.float32)
    wt = rng.integers(1, 9, size=(h, w)).astype(np.float32)
    p = str(tmp_path / "f.pfm")
    Film(rgb, wt).save_pfm(p)
    back = read_pfm(p)
    np.testing.assert_array_equal(back, (rgb / wt[..., None]).astype(np.float32))
    with open(p, "rb") as f:
        assert f.readline() == b"PF\n" and f.readline() == b"4 5\n"
