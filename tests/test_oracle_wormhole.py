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


# ---------------------------------------------------------------- ray bundles
def _face_dir(face, i, j, N):
    """Direction of the centre of texel (i, j) of a face: the inverse of calc_skybox_coord."""
    u, v = 2 * (i - 0.5) / N - 1, 2 * (j - 0.5) / N - 1
    d = {1: (1, -u, v), 2: (-u, -1, v), 3: (-1, u, v), 4: (u, 1, v), 5: (-u, -v, -1), 6: (-u, v, 1)}[face]
    d = np.array(d, float)
    return d / np.linalg.norm(d)


def test_cube_face_inverse_projection():
    N = 8
    for face in range(1, 7):
        for i in range(1, N + 1):
            for j in range(1, N + 1):
                d = _face_dir(face, i, j, N)
                got = O.skybox_coord(_ray_dir(d), N)
                assert got[4] == face and got[0] == i and got[1] == j


def test_cube_wrap_follows_the_cube():
    """Shifts by one texel across every face edge land on the geometrically adjacent texel
    (the centres are one texel apart), and shifts inside a face move (i, j)."""
    N = 8
    step = np.arctan(2 / N)  # largest centre-to-centre angle of neighbouring texels
    for face in range(1, 7):
        assert O.cube_wrap(4, 5, face, 0, 0, N) == (4, 5, face)
        assert O.cube_wrap(4, 5, face, 2, 0, N) == (6, 5, face)
        assert O.cube_wrap(4, 5, face, 0, -3, N) == (4, 2, face)
        for i in range(1, N + 1):
            for j in range(1, N + 1):
                if 1 < i < N and 1 < j < N:
                    continue
                for x, y in ((1, 0), (-1, 0), (0, 1), (0, -1)):
                    i2, j2, f2 = O.cube_wrap(i, j, face, x, y, N)
                    assert 1 <= i2 <= N and 1 <= j2 <= N and 1 <= f2 <= 6
                    ang = math.acos(min(1.0, _face_dir(face, i, j, N) @ _face_dir(f2, i2, j2, N)))
                    assert ang <= 1.01 * step, (face, i, j, x, y, (i2, j2, f2), ang)


def test_distort_gauss_is_the_covariance_transform():
    rng = np.random.default_rng(3)
    for _ in range(50):
        sx, sy, th = rng.uniform(0.1, 3), rng.uniform(0.1, 3), rng.uniform(-PI, PI)
        a, b = rng.uniform(0.5, 5), rng.uniform(0.5, 5)
        Rm = np.array([[math.cos(th), -math.sin(th)], [math.sin(th), math.cos(th)]])
        S = np.diag([a, b]) @ Rm @ np.diag([sx * sx, sy * sy]) @ Rm.T @ np.diag([a, b])
        sxn, syn, thn = O.distort_gauss(sx, sy, th, a, b)
        Rn = np.array([[math.cos(thn), -math.sin(thn)], [math.sin(thn), math.cos(thn)]])
        assert np.allclose(Rn @ np.diag([sxn * sxn, syn * syn]) @ Rn.T, S, rtol=1e-10, atol=1e-12 * S.max())
    # D2_Gauss is that normal density: it integrates to 1
    sxn, syn, thn = O.distort_gauss(1.2, 0.7, 0.4, 1.5, 2.0)
    st = min(sxn, syn) / 4
    ext = 8 * max(sxn, syn)
    xs = np.arange(-ext, ext + st / 2, st)
    tot = sum(O.d2_gauss(x, y, sxn, syn, thn) for x in xs for y in xs) * st * st
    assert tot == pytest.approx(1.0, rel=1e-6)


def test_bundle_params_of_the_camera():
    sc, pos, d, p, beta = scenes.initialize_world("wormhole")
    cam = abi.make_camera(pos, d, beta, p, 9, 85, 60)
    Mc = O.camera_matrix(sc, cam)
    r = _dneg(pos[1])[0]
    B3 = Mc[1:, 1:] / np.array([1.0, 1 / r, 1 / (r * math.sin(pos[2]) + 1e-24)])[:, None]
    W, H = math.tan(PI * 85 / 360), math.tan(PI * 60 / 360)
    ny = O.camera_y_pix(cam)
    rh, rv = 2 * W / 10, 2 * H / (ny + 1)
    for i, j in ((0, 0), (4, 3), (8, ny - 1), (2, 5)):
        w, h = rh * (i + 0.5) - W, rv * (j + 0.5) - H
        th = math.atan(math.hypot(w, h))
        ph = math.atan2(h, w)
        n = np.array([-math.sin(th) * math.cos(ph), -math.sin(th) * math.sin(ph), -math.cos(th)])
        E = 1 / (abs(n[2]) + 1e-24)
        J = np.cross(n, [0, 0, -1.0])
        if J @ J < 1e-3:
            J = np.array([0, -1.0, 0])
        J, nr = B3 @ J, B3 @ n
        J, nr = J / np.linalg.norm(J), nr / np.linalg.norm(nr)
        a = np.array([nr[1], 1 - nr[1] ** 2 / (1 - nr[0]), -nr[1] * nr[2] / (1 - nr[0])])
        b = np.array([nr[2], -nr[1] * nr[2] / (1 - nr[0]), 1 - nr[2] ** 2 / (1 - nr[0])])
        assert abs(a @ nr) < 1e-12 and abs(b @ nr) < 1e-12  # the bundle plane is transverse to the ray
        a, b = a / (a @ a), b / (b @ b)  # the reference's squared-norm division
        want = [0, 0, 0, 0, 0, (a @ J) * (1 + E) / 2, (b @ J) * (1 + E) / 2, (a @ J) * (1 - E) / 2,
                (b @ J) * (1 - E) / 2, 0, abs(math.atan(rh) / 2 * n[2])]
        assert np.allclose(O.bundle_params(cam, i, j), want, rtol=1e-10, atol=1e-14)


def test_beam_skybox_of_a_uniform_sky_is_the_weight_sum():
    sc, pos, d, p, beta = scenes.initialize_world("wormhole")
    cam = abi.make_camera(pos, d, beta, p, 10, 85, 60)
    lam = [450.0, 630.0]
    O.load_skybox(1, np.full((24, 32, 3), 180, dtype=np.uint8))
    O.load_skybox(2, np.full((24, 32, 3), 90, dtype=np.uint8))
    single = O.render(sc, cam, abi.make_integrator(1e-4, 1e-4, 0.1, 1e4), lam, nthreads=4, want_rays=True)
    beam = O.render(sc, cam, abi.make_integrator(1e-4, 1e-4, 0.1, 1e4, ray_bundles=True), lam, nthreads=4,
                    want_rays=True)
    assert beam["rays"].shape[1] == single["rays"].shape[1] + 11
    assert np.array_equal(beam["steps"], single["steps"])  # bundle slots are not integrated
    ratio = beam["rays"][:, 8] / single["rays"][:, 8]
    assert np.all(ratio > 0.5) and np.all(ratio < 1.0 + 1e-12)
    # the ratio is the sum of the D2_Gauss weights over the (2h+1)^2 texels
    k = 7
    ray, b = beam["rays"][k], beam["rays"][k][-11:]
    N = 8
    dp = b[10] * ray[4]
    m1, m2 = math.hypot(b[5], b[6]), math.hypot(b[7], b[8])
    mu = b[4] + math.atan2(b[5] * b[8] + b[6] * b[7], b[5] * b[7] - b[6] * b[8]) / 2
    sbs = math.atan(2 / (N + 1)) / 2
    sx, sy, th = O.distort_gauss(dp * (m1 + m2), dp * (m1 - m2), mu, 1 / (sbs + 1e-24), 1 / (sbs + 1e-24))
    h = min(math.ceil(max(sx, sy)), 10)
    wsum = sum(O.d2_gauss(x, y, sx, sy, th) for x in range(-h, h + 1) for y in range(-h, h + 1))
    assert ratio[k] == pytest.approx(wsum, rel=1e-12)
