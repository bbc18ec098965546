"""Output files from a GPU render: the spectral cube of the returned rays weighted by the
CIE trapezoid weights of calc_xyY (color_theory.jl:79-108; the oracle's weights) equals
the XYZ the kernels accumulated (relative 1e-12), and the PNG / FITS writers take the
library's outputs.
"""
import os

import numpy as np
import pytest

import oracle as O
from rtwca_amd import host, output, scenes

pytestmark = pytest.mark.gpu


def test_spectrum_png_and_fits_of_a_render(tmp_path):
    w = scenes.config(1, x_pix=32, n_lambda=14)
    ctx = host.Context(0)
    r = ctx.render(w.scene, w.cameras, w.integ, w.lam, want_rays=True)
    I = output.spectral_radiance(r["rays"], w.lam)
    W = O.cie_weights(w.lam)
    assert np.allclose(I @ W.T, r["xyz"], rtol=1e-12, atol=1e-300)
    img = ctx.xyz_to_srgb16(r["xyz"], r["x_pix"], r["y_pix"])
    output.write_png16(str(tmp_path / "a.png"), img)
    output.write_fits(str(tmp_path / "a.fits"), r["rays"], w.lam, r["x_pix"], r["y_pix"])
    assert os.path.getsize(str(tmp_path / "a.fits")) % 2880 == 0
