"""Output files (rtwca_amd/output.py) -- CPU only.

  part                                   pinned by
  -------------------------------------  ------------------------------------------------
  spectral_radiance (color_theory.jl:    the CIE trapezoid weights of calc_xyY applied to
  67-76, ray_to_I_lambda)                it give the oracle's XYZ of the same rays
  write_png16                            a decoder written here (chunk CRCs, zlib, filter 0,
                                         big-endian samples) returns the image bit for bit;
                                         Pillow opens it as a 16-bit RGB PNG of that size
  write_fits                             a reader written here: 2880-byte blocks, mandatory
                                         keywords in order, big-endian float64 planes equal
                                         to the ray-order (lambda, y, x) cubes, the TAU extension
"""
import os
import struct
import zlib

import numpy as np
import pytest

import oracle as O
from rtwca_amd import output

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden", "oracle_cfg1_16px_64l.npz")


def test_spectral_radiance_gives_the_xyz():
    g = np.load(GOLD)
    lam = g["lam"]
    I = output.spectral_radiance(g["rays"], lam)
    W = O.cie_weights(lam)
    xyz = I @ W.T
    assert np.allclose(xyz, g["xyz"], rtol=1e-12, atol=1e-300)
    assert np.array_equal(output.optical_depth(g["rays"], len(lam)), g["rays"][:, 9::2])
    with pytest.raises(ValueError):
        output.spectral_radiance(g["rays"][:, :100], lam)


def _read_png(path):
    data = open(path, "rb").read()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    p, chunks = 8, []
    while p < len(data):
        n, = struct.unpack(">I", data[p:p + 4])
        tag, body = data[p + 4:p + 8], data[p + 8:p + 8 + n]
        crc, = struct.unpack(">I", data[p + 8 + n:p + 12 + n])
        assert crc == zlib.crc32(tag + body) & 0xFFFFFFFF
        chunks.append((tag, body))
        p += 12 + n
    assert chunks[0][0] == b"IHDR" and chunks[-1][0] == b"IEND"
    w, h, depth, ctype, comp, filt, inter = struct.unpack(">IIBBBBB", chunks[0][1])
    assert (depth, ctype, comp, filt, inter) == (16, 2, 0, 0, 0)
    raw = zlib.decompress(b"".join(b for t, b in chunks if t == b"IDAT"))
    rows = np.frombuffer(raw, np.uint8).reshape(h, 1 + 6 * w)
    assert np.all(rows[:, 0] == 0)
    return rows[:, 1:].copy().view(">u2").reshape(h, w, 3).astype(np.uint16)


def test_png16_round_trip(tmp_path):
    rng = np.random.default_rng(3)
    img = rng.integers(0, 65536, size=(7, 11, 3), dtype=np.uint16)
    img[0, 0] = (0, 65535, 256)
    path = str(tmp_path / "img.png")
    output.write_png16(path, img)
    assert np.array_equal(_read_png(path), img)
    PIL = pytest.importorskip("PIL.Image")
    with PIL.open(path) as im:
        assert im.size == (11, 7)
        im.load()
    with pytest.raises(ValueError):
        output.write_png16(path, img.astype(np.uint8))


def _read_fits(path):
    data = open(path, "rb").read()
    assert len(data) % 2880 == 0
    hdus, p = [], 0
    while p < len(data):
        cards = {}
        order = []
        while True:
            block = data[p:p + 2880].decode("ascii")
            p += 2880
            done = False
            for k in range(36):
                c = block[80 * k:80 * (k + 1)]
                key = c[:8].strip()
                if key == "END":
                    done = True
                    break
                if c[8:10] == "= ":
                    v = c[10:].split(" / ")[0].strip()
                    if v.startswith("'"):
                        v = v.strip("'").strip()
                    elif v in ("T", "F"):
                        v = v == "T"
                    elif "." in v or "E" in v:
                        v = float(v)
                    else:
                        v = int(v)
                    cards[key] = v
                    order.append(key)
            if done:
                break
        shape = tuple(cards["NAXIS%d" % (k + 1)] for k in range(cards["NAXIS"]))[::-1]
        nbytes = int(np.prod(shape)) * 8
        arr = np.frombuffer(data[p:p + nbytes], dtype=">f8").reshape(shape)
        p += nbytes + (-nbytes % 2880)
        hdus.append((order, cards, arr))
    return hdus


def test_fits_cube(tmp_path):
    g = np.load(GOLD)
    lam = g["lam"]
    path = str(tmp_path / "cube.fits")
    output.write_fits(path, g["rays"], lam, 16, 16, header=[("SCENE", "cfg1"), ("TOL", 1e-8)])
    (o1, c1, I), (o2, c2, tau) = _read_fits(path)
    assert o1[:6] == ["SIMPLE", "BITPIX", "NAXIS", "NAXIS1", "NAXIS2", "NAXIS3"]
    assert c1["SIMPLE"] is True and c1["BITPIX"] == -64 and c1["EXTEND"] is True
    assert o2[:8] == ["XTENSION", "BITPIX", "NAXIS", "NAXIS1", "NAXIS2", "NAXIS3", "PCOUNT", "GCOUNT"]
    assert c2["XTENSION"] == "IMAGE" and c2["EXTNAME"] == "TAU"
    assert c1["SCENE"] == "cfg1" and c1["TOL"] == 1e-8
    assert c1["CRVAL3"] == lam[0] and c1["CDELT3"] == pytest.approx(lam[1] - lam[0], rel=1e-12)
    assert [c1["LAM%d" % (k + 1)] for k in range(len(lam))] == list(lam)
    assert I.shape == (len(lam), 16, 16) and tau.shape == I.shape
    want = output.spectral_radiance(g["rays"], lam).reshape(16, 16, -1)
    # FITS row 1 is the bottom image row, ray row j = 0
    assert np.array_equal(I[:, 0, :], want[0].T) and np.array_equal(I[:, 15, :], want[15].T)
    assert np.array_equal(tau[:, 0, :], g["rays"][:16, 9::2].T)
