"""Pins of the oracle's wormhole scene (main/Wormhole.source) and skyboxes -- CPU only.

  part                                      pinned by
  ----------------------------------------  -------------------------------------------
  r(l) (source_geometry.jl:16-30)           the Dneg wormhole of James et al. (2015,
                                            eqs. 5-6) evaluated independently; throat
                                            radius 1 inside |l| <= a; even in l
  dr_dl (:33-44), Christoffel (:129-146)    symbols without r' equal 1/2 g^-1 dg of the
                                            metric; those with r' use the reference's
                                            dr_dl = numerical dr/dl + 2 sign(l)(x-1)/
                                            (pi (1+x^2)) (the extra term is the quirk)
  calc_skybox_coord (source_model.jl:175)   the six axis directions land on faces 1..6 at
                                            the centre texel; u, v in [-1, 1]
  cubemap_coord_calc (gen_geometry.jl:347)  a texture whose bytes encode (face, i, j)
  skybox_I_nu_calc (source_model.jl:150)    returns them through the 3N + 1 - j row flip;
                                            Normal_Gauss integrates to 1
  skybox_handling! (:393-443)               a uniform sky gives every escaping ray
                                            I = 1e-12 sum(w_c NG_c(f)) / f^3, f = (1/lambda) k^t,
                                            on the side given by the sign of l
"""
import math

import numpy as np
import pytest

import oracle as O
from rtwca_amd import abi, scenes, skybox

PI = math.pi


def _dneg(l):
    q = PI / (2 * math.sqrt(2))
    M = 0.05 / (-math.log(1 / math.cos(q)) + q * math.tan(q))
    a = 0.005
    if abs(l) <= a:
        return 1.0, 0.0, M
    x = 2 * (abs(l) - a) / (PI * M)
    return 1.0 + M * (x * math.atan(x) - 0.5 * math.log(1 + x * x)), x, M


def test_r_of_l_is_the_dneg_wormhole():
    for l in (0.0, 0.004, -0.004, 0.005, 0.0051, 0.05, -0.05, 0.5, 3.0, -3.0, 80.0):
        r, _ = O.wormhole_r(l)
        assert r == pytest.approx(_dneg(l)[0], rel=1e-14, abs=1e-15)
        assert r == O.wormhole_r(-l)[0]
    assert O.wormhole_r(0.0)[0] == 1.0


@pytest.mark.parametrize("l", [0.02, -0.3, 2.0, -40.0])
def test_christoffel_symbols_and_the_dr_dl_quirk(l):
    sc = scenes.wormhole()
    th = 1.1
    x = np.array([0.0, l, th, 0.7])
    G = O.christoffel(sc, x)
    # 1/2 g^-1 dg for the symbols without r'
    assert G[2, 3, 3] == pytest.approx(-math.sin(th) * math.cos(th), rel=1e-14)
    assert G[3, 2, 3] == pytest.approx(1 / math.tan(th), rel=1e-14)
    e = 1e-6
    drdl_num = (O.wormhole_r(l + e)[0] - O.wormhole_r(l - e)[0]) / (2 * e)
    r, xx, M = _dneg(l)
    quirk = 2 * math.copysign(1, l) * (xx - 1) / (PI * (1 + xx * xx))
    drdl_ref = O.wormhole_r(l)[1]
    assert drdl_num == pytest.approx(2 * math.copysign(1, l) * math.atan(xx) / PI, rel=1e-7)
    assert drdl_ref == pytest.approx(drdl_num + quirk, rel=1e-7, abs=1e-9)
    assert abs(quirk) > 1e-4  # the reference's dr_dl differs from the derivative of its r(l)
    assert G[1, 2, 2] == pytest.approx(-r * drdl_ref, rel=1e-13)
    assert G[1, 3, 3] == pytest.approx(-r * drdl_ref * math.sin(th) ** 2, rel=1e-13)
    assert G[2, 1, 2] == pytest.approx(drdl_ref / r, rel=1e-13) and G[2, 2, 1] == G[2, 1, 2]
    assert G[3, 1, 3] == pytest.approx(drdl_ref / r, rel=1e-13) and G[3, 3, 1] == G[3, 1, 3]


def _ray_dir(d, l=60.0, th=1.0, ph=0.4):
    """A ray state at (l, th, ph) whose tetrad direction is the Cartesian vector d."""
    r = _dneg(l)[0]
    st, ct, sp, cp = math.sin(th), math.cos(th), math.sin(ph), math.cos(ph)
    er = np.array([st * cp, st * sp, ct])
    et = np.array([ct * cp, ct * sp, -st])
    ep = np.array([-sp, cp, 0.0])
    d = np.asarray(d, float) / np.linalg.norm(d)
    return np.array([0, l, th, ph, 1.0, d @ er, (d @ et) / r, (d @ ep) / (r * st)])


def test_cube_faces_and_texel_addressing():
    N = 8
    for d, face in (((1, 0, 0), 1), ((-1, 0, 0), 3), ((0, 1, 0), 4), ((0, -1, 0), 2), ((0, 0, 1), 6),
                    ((0, 0, -1), 5)):
        i, j, u, v, f, ia, ja = O.skybox_coord(_ray_dir(d), N)
        assert f == face and abs(u) < 1e-12 and abs(v) < 1e-12
        assert (i, j) == (N // 2 + 1, N // 2 + 1)
    # a texture whose bytes encode (face, i, j): red = face * 40, green = i, blue = j
    img = np.zeros((3 * N, 4 * N, 3), dtype=np.uint8)
    for face, (r0, c0) in skybox.cross_faces(N).items():
        for i in range(1, N + 1):
            for j in range(1, N + 1):
                img[r0 + (N - j), c0 + i - 1] = (face * 40, i, j)  # row 3N + 1 - j_abs (1-based)
    O.load_skybox(1, img)
    rng = np.random.default_rng(0)
    f0 = 0.001587  # red line centre: the red channel dominates
    for _ in range(200):
        d = rng.standard_normal(3)
        i, j, u, v, face, ia, ja = O.skybox_coord(_ray_dir(d), N)
        assert -1 <= u <= 1 and -1 <= v <= 1 and 1 <= i <= N and 1 <= j <= N
        ng = lambda x, s, m: math.exp(-(x - m) ** 2 / (2 * s * s)) / (s * math.sqrt(2 * PI))  # noqa: E731
        want = 1e-12 * ((face * 40 / 255) * ng(f0, 5e-5, 0.001587) + 0.843 * (i / 255) * ng(f0, 7e-5, 0.001879)
                        + 0.914 * (j / 255) * ng(f0, 9e-5, 0.002141))
        assert O.skybox_I(1, ia, ja, f0) == pytest.approx(want, rel=1e-12)


def test_normal_gauss_lines_are_normalised():
    O.load_skybox(1, np.full((6, 8, 3), 255, dtype=np.uint8))
    f = np.linspace(0.0012, 0.0026, 20001)
    I = np.array([O.skybox_I(1, 1, 2, x) for x in f])
    assert np.trapezoid(I, f) == pytest.approx(1e-12 * (1.0 + 0.843 + 0.914), rel=1e-6)


def test_uniform_sky_end_to_end():
    sc, pos, d, p, beta = scenes.initialize_world("wormhole")
    cam = abi.make_camera(pos, d, beta, p, 12, 85, 60)
    lam = [400.0, 550.0, 650.0]
    it = abi.make_integrator(1e-4, 1e-4, 0.1, 1e4)
    O.load_skybox(1, np.full((24, 32, 3), (200, 100, 50), dtype=np.uint8))
    O.load_skybox(2, np.full((24, 32, 3), (10, 20, 250), dtype=np.uint8))
    r = O.render(sc, cam, it, lam, nthreads=4, want_rays=True)
    ng = lambda x, s, m: math.exp(-(x - m) ** 2 / (2 * s * s)) / (s * math.sqrt(2 * PI))  # noqa: E731
    sides = set()
    for ray in r["rays"]:
        near = ray[1] > 0
        sides.add(near)
        rgb = (200, 100, 50) if near else (10, 20, 250)
        w = (rgb[0] / 255, 0.843 * rgb[1] / 255, 0.914 * rgb[2] / 255)
        for c, lm in enumerate(lam):
            f = ray[4] / lm
            want = 1e-12 * (w[0] * ng(f, 5e-5, 0.001587) + w[1] * ng(f, 7e-5, 0.001879)
                            + w[2] * ng(f, 9e-5, 0.002141)) / f ** 3
            assert ray[8 + 2 * c] == pytest.approx(want, rel=1e-12)
    assert sides == {True, False}
