"""Output files of a render: the 16-bit sRGB image and the raw spectral cube.

* write_png16 writes what the reference saves for its RGB{N0f16} images
  (main/initialization.jl:374 `save(... .png, img)`): a 16-bit-per-channel RGB PNG of
  the rtw_xyz_to_srgb16 image, rows top to bottom as returned.
* spectral_radiance is ray_to_I_lambda (main/color_theory.jl:67-76): the spectral
  radiance I_lambda = (I_nu / nu^3) f^3 c / lambda^2 in W sr^-1 m^-2 nm^-1 of each
  pixel at each wavelength, from the final rays (rtw_render's out_rays);
* write_fits writes that cube and the optical depths as FITS images (the reference's
  planned "raw spectral brightness data (i.e. something like FITS files)", README.md:21):
  primary HDU I_lambda (NAXIS1 = x, NAXIS2 = y, NAXIS3 = wavelength), image extension
  TAU, float64.  FITS rows run bottom to top, as the camera's pixel rows do (ray
  i + j x_pix, j = 0 the bottom row: the PNG's top row is j = y_pix - 1 after the
  reference's transpose(reverse(img, dims=2)), LttM.jl:204), so both display alike.

These are host-side formatting of results already copied back from the GPU.
"""
import math
import struct
import zlib

import numpy as np

C_LIGHT = 299792458.0  # main/64BitConstants.jl:17


def spectral_radiance(rays, lam_nm):
    """I_lambda of each ray and wavelength (color_theory.jl:67-76).

    rays: (..., n) final ray states [t r theta phi k^t k^r k^theta k^phi (I/nu^3, tau) x N
    (bundle slots)]; lam_nm: the N wavelengths.  Returns (..., N)."""
    lam = np.asarray(lam_nm, dtype=np.float64)
    rays = np.asarray(rays, dtype=np.float64)
    n = len(lam)
    if rays.shape[-1] < 8 + 2 * n:
        raise ValueError("rays hold %d components, %d colours need %d" % (rays.shape[-1], n, 8 + 2 * n))
    f = 1.0 / lam
    return rays[..., 8:8 + 2 * n:2] * f ** 3 * C_LIGHT / lam ** 2


def optical_depth(rays, n_lambda):
    """tau of each ray and colour, (..., N)."""
    rays = np.asarray(rays, dtype=np.float64)
    return rays[..., 9:9 + 2 * n_lambda:2]


def _png_chunk(tag, data):
    body = tag + data
    return struct.pack(">I", len(data)) + body + struct.pack(">I", zlib.crc32(body) & 0xFFFFFFFF)


def write_png16(path, rgb16):
    """A (height, width, 3) uint16 image as a 16-bit RGB PNG (colour type 2, no interlace)."""
    img = np.asarray(rgb16)
    if img.ndim != 3 or img.shape[2] != 3 or img.dtype != np.uint16:
        raise ValueError("expected a (height, width, 3) uint16 image")
    h, w, _ = img.shape
    if h < 1 or w < 1:
        raise ValueError("empty image")
    rows = img.astype(">u2").reshape(h, w * 3).view(np.uint8).reshape(h, w * 6)
    raw = np.concatenate([np.zeros((h, 1), np.uint8), rows], axis=1).tobytes()  # filter type 0 per row
    with open(path, "wb") as fh:
        fh.write(b"\x89PNG\r\n\x1a\n")
        fh.write(_png_chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 16, 2, 0, 0, 0)))
        fh.write(_png_chunk(b"IDAT", zlib.compress(raw, 6)))
        fh.write(_png_chunk(b"IEND", b""))


def _card(key, value=None, comment=""):
    key = key.upper()
    if len(key) > 8:
        raise ValueError("FITS keyword longer than 8 characters: %r" % key)
    if value is None:
        s = key.ljust(8)
    else:
        if isinstance(value, bool):
            v = ("T" if value else "F").rjust(20)
        elif isinstance(value, (int, np.integer)):
            v = str(int(value)).rjust(20)
        elif isinstance(value, (float, np.floating)):
            if not math.isfinite(value):
                raise ValueError("FITS header values must be finite")
            v = ("%.17G" % value).rjust(20)
            if "." not in v and "E" not in v:
                v = (v.strip() + ".").rjust(20)
        else:
            t = str(value).replace("'", "''")
            v = ("'" + t.ljust(8) + "'").ljust(20)
        s = key.ljust(8) + "= " + v
        if comment:
            s += " / " + comment
    if len(s) > 80:
        s = s[:80]
    return s.ljust(80)


def _hdu(cards, data):
    head = "".join(cards) + "END".ljust(80)
    head += " " * (-len(head) % 2880)
    body = np.ascontiguousarray(data, dtype=">f8").tobytes()
    body += b"\0" * (-len(body) % 2880)
    return head.encode("ascii") + body


def write_fits(path, rays, lam_nm, x_pix, y_pix, header=None):
    """The spectral radiance and optical depth cubes of one pose as a FITS file.

    rays: (x_pix * y_pix, n) final rays in ray order (pixel i + j x_pix, j = 0 the bottom
    row, as rtw_render returns them for one pose); header: extra
    (keyword, value) pairs for the primary HDU (keywords of at most 8 characters)."""
    lam = np.asarray(lam_nm, dtype=np.float64)
    n = len(lam)
    rays = np.asarray(rays, dtype=np.float64)
    if rays.shape[0] != x_pix * y_pix:
        raise ValueError("expected %d rays, got %d" % (x_pix * y_pix, rays.shape[0]))
    # ray order (j, i, lambda) with j = 0 the bottom row -> FITS order (lambda, y, x)
    I = spectral_radiance(rays, lam).reshape(y_pix, x_pix, n).transpose(2, 0, 1)
    tau = optical_depth(rays, n).reshape(y_pix, x_pix, n).transpose(2, 0, 1)
    axes = [("NAXIS1", x_pix), ("NAXIS2", y_pix), ("NAXIS3", n)]
    wcs = [_card("CTYPE3", "WAVE", "wavelength"), _card("CUNIT3", "nm")]
    step = np.diff(lam)
    if n >= 2 and np.allclose(step, step[0], rtol=1e-12, atol=0):  # a linear axis; otherwise LAMk only
        wcs += [_card("CRPIX3", 1.0), _card("CRVAL3", float(lam[0])), _card("CDELT3", float(step[0]))]
    lam_cards = [_card("LAM%d" % (k + 1), float(v), "wavelength of plane %d (nm)" % (k + 1))
                 for k, v in enumerate(lam)] if n <= 99999 else []
    prim = [_card("SIMPLE", True), _card("BITPIX", -64), _card("NAXIS", 3)] + [_card(k, v) for k, v in axes]
    prim += [_card("EXTEND", True), _card("BUNIT", "W sr-1 m-2 nm-1", "spectral radiance I_lambda")]
    prim += wcs + lam_cards
    for k, v in (header or []):
        prim.append(_card(k, v))
    ext = [_card("XTENSION", "IMAGE"), _card("BITPIX", -64), _card("NAXIS", 3)] + [_card(k, v) for k, v in axes]
    ext += [_card("PCOUNT", 0), _card("GCOUNT", 1), _card("EXTNAME", "TAU"), _card("BUNIT", "", "optical depth")]
    ext += wcs
    with open(path, "wb") as fh:
        fh.write(_hdu(prim, I))
        fh.write(_hdu(ext, tau))
