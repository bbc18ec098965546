"""Pins of the oracle's two flat Cartesian scenes -- CPU only.

main/Minkowski_Sphere_Cartesian.source (the reference's "cart_sph" test scene) and
main/Synchro_Test.source (a power-law synchrotron cylinder).

  part                                        pinned by
  ------------------------------------------  -------------------------------------------
  metric, Christoffel symbols, vierbein       diag(-1, 1, 1, 1), zero symbols: every ray
  (source_geometry.jl:8-82)                   stays straight, k unchanged to the last bit
  pad_max_dt (Cart :130-132, Synchro :229)    max_dt_scale (1 + |x| + |y| (+ |z|)) at points
  keepinbounds!, near_singularity             no-ops (a negative "theta" slot is kept)
  calc_terminate (Cart :158-172)              rays stop beyond 100 from the origin, not at
                                              z > 8 pi (no winding test), class escape
  sphere blackbody (Cart :51-89)              the optical depth of each ray is a_nu times
                                              its chord through the unit sphere (closed form
                                              from the camera direction), converging with
                                              the tolerance; I / nu^3 = B_nu / nu^3 (1 - e^-tau)
  skybox_handling! (Cart :406-431)            every ray, through the sphere or not, adds the
                                              texel of skybox 1 in the direction of its k
                                              (face and texel from a cube-map projection
                                              written here), attenuated by e^-tau
  Synchro_Test field (Synchro :50-171)        b = (0, 0, 0, B) so cos^2 = (k^z / k^t)^2, nu_c =
                                              9.34e-11 B; no floor: j = a = 0 along the field
  Synchro_Test transfer                       rays from the origin cross the unit cylinder
                                              over 1 / sqrt(n_x^2 + n_y^2): tau = a s / k_t^2,
                                              I / nu^3 = j / (a nu^3) (1 - e^-tau) with the
                                              point's (constant) coefficients; j / a ~ nu^(5/2)
"""
import math

import numpy as np
import pytest

import oracle as O
from rtwca_amd import abi, scenes, skybox

PI = math.pi
F32 = lambda x: float(np.float32(x))  # noqa: E731
LAM = [400.0, 550.0, 700.0]


def _ng(x, s, m):
    return math.exp(-(x - m) ** 2 / (2 * s * s)) / (s * math.sqrt(2 * PI))


def _sky_I(rgb, f):
    w = (rgb[0] / 255, 0.843 * rgb[1] / 255, 0.914 * rgb[2] / 255)
    return 1e-12 * (w[0] * _ng(f, 5e-5, 0.001587) + w[1] * _ng(f, 7e-5, 0.001879) + w[2] * _ng(f, 9e-5, 0.002141))


def _cube(d, N):
    """Face, 1-based in-face texel of direction d: the standard major-axis cube map."""
    x, y, z = np.asarray(d, float) / np.linalg.norm(d)
    ax, ay, az = abs(x), abs(y), abs(z)
    if ax >= ay and ax >= az:
        face, u, v = (1, -y / ax, z / ax) if x > 0 else (3, y / ax, z / ax)
    elif ay >= az:
        face, u, v = (4, x / ay, z / ay) if y > 0 else (2, -x / ay, z / ay)
    else:
        face, u, v = (6, -x / az, y / az) if z > 0 else (5, -x / az, -y / az)
    return face, min(int(math.floor(N * (u + 1) / 2)) + 1, N), min(int(math.floor(N * (v + 1) / 2)) + 1, N)


def _chord_sphere(x0, k):
    m = -np.asarray(k) / np.linalg.norm(k)  # rays are traced backwards, along -k
    t0 = -(m @ x0)
    d2 = float(np.sum((x0 + t0 * m) ** 2))
    return 2 * math.sqrt(1 - d2) if (d2 < 1 and t0 > 0) else 0.0


@pytest.fixture(scope="module")
def cart():
    return scenes.initialize_world("minkowski_t,x,y,z_sphere")


@pytest.fixture(scope="module")
def synch():
    return scenes.initialize_world("synch")


def test_flat_geometry(cart, synch):
    for sc in (cart[0], synch[0]):
        x = [0.3, 2.0, -1.5, 7.0]
        assert np.array_equal(O.metric(sc, x), np.diag([-1.0, 1.0, 1.0, 1.0]))
        assert not np.any(O.christoffel(sc, x))
        ray = np.array([0.0, -0.5, -2.0, 30.0, 1.0, 0.6, -0.8, 0.0])
        out = O.keepinbounds(sc, ray)
        assert np.array_equal(out, ray)  # no theta reflection: slot 2 is y here
        flag, h = O.near_singularity(sc, ray, -0.25, 1e-10)
        assert not flag and h == -0.25
    ray = np.array([0.0, -0.5, -2.0, 30.0, 1.0, 0.6, -0.8, 0.0])
    assert O.pad_max_dt(cart[0], ray, 0.1) == pytest.approx(0.1 * (1 + 0.5 + 2.0 + 30.0), rel=1e-15)
    assert O.pad_max_dt(synch[0], ray, 0.1) == pytest.approx(0.1 * (1 + 0.5 + 2.0), rel=1e-15)


def test_stop_beyond_100_without_winding_test(cart):
    sc = cart[0]
    it = abi.make_integrator(1e-8, 1e-10, 0.1, 1e4)
    assert not O.terminate(sc, it, 1, [0, 0, 0, 30.0, 1, 0, 0, 1, 0, 0], 5)  # z = 30 > 8 pi
    assert O.terminate(sc, it, 1, [0, 60.0, 60.0, 60.0, 1, 0, 0, 1, 0, 0], 5)  # |x| = 104 > 100
    assert not O.terminate(sc, it, 1, [0, 57.0, 57.0, 57.0, 1, 0, 0, 1, 0, 0], 5)  # |x| = 98.7
    assert O.classify(sc, it, 1, [0, 60.0, 60.0, 60.0, 1, 0, 0, 1, 0, 0]) == 2


def _render_cart(cart, x_pix, rtol, atol, sky, lam=LAM, fov=(85, 60), **kw):
    sc, pos, d, p, beta = cart
    cam = abi.make_camera(pos, d, beta, p, x_pix, *fov)
    O.load_skybox(1, sky)
    r = O.render(sc, cam, abi.make_integrator(rtol, atol, 0.1, 1e5, **kw), lam, nthreads=4, want_rays=True)
    return cam, O.camera_rays(sc, cam), r


def test_sphere_chords_and_blackbody(cart):
    x0 = np.array(cart[1][1:], float)
    errs, ierrs = [], []
    for rtol, atol in ((1e-6, 1e-8), (1e-8, 1e-10), (1e-10, 1e-12)):
        cam, R0, r = _render_cart(cart, 16, rtol, atol, np.zeros((24, 32, 3), np.uint8))
        rays = r["rays"]
        assert np.array_equal(rays[:, 4:8], R0[:, 4:8])  # straight lines: k never changes
        assert np.all(r["cls"] == 2)
        L = np.array([_chord_sphere(x0, k) for k in R0[:, 5:8]])
        assert (L > 0).sum() > 10 and (L == 0).sum() > 10
        tau = L / R0[:, 4] ** 2  # d tau / d lambda = a / k^t, |dx / d lambda| = k^t
        errs.append(np.max(np.abs(rays[:, 9::2] - tau[:, None])))
        B3 = np.array([O.planck(5778.0, 1 / lm) * lm ** 3 for lm in LAM])
        ierrs.append(np.max(np.abs(rays[:, 8::2] / B3[None, :] - (1 - np.exp(-tau[:, None])))))
        # the end point lies on the straight line from the camera, beyond 100
        s = (rays[:, 1:4] - x0) @ (-R0[:, 5:8].T)
        perp = rays[:, 1:4] - x0 - np.diag(s)[:, None] * (-R0[:, 5:8])
        assert np.max(np.abs(perp)) < 1e-9 and np.all(np.linalg.norm(rays[:, 1:4], axis=1) > 100)
    # both converge to the closed forms as the tolerance falls (the error comes from the
    # steps that straddle the sphere's surface; I converges more slowly, about as the
    # square root of the tolerance, since the averaged slope that starts the step after a
    # crossing (the reference's FSAL by average) carries the other side's emission)
    assert errs[2] < 1e-7 and errs[1] < errs[0] / 30 and errs[2] < errs[1] / 30, errs
    assert ierrs[2] < 5e-5 and ierrs[1] < ierrs[0] / 5 and ierrs[2] < ierrs[1] / 5, ierrs


def test_skybox_texels_in_the_ray_direction(cart):
    N = 8
    img = np.zeros((3 * N, 4 * N, 3), dtype=np.uint8)
    for face, (r0, c0) in skybox.cross_faces(N).items():
        for i in range(1, N + 1):
            for j in range(1, N + 1):
                img[r0 + (N - j), c0 + i - 1] = (face * 40, i * 20, j * 20)
    cam, R0, r = _render_cart(cart, 14, 1e-6, 1e-8, img, fov=(160, 140))  # wide: several faces
    _, _, dark = _render_cart(cart, 14, 1e-6, 1e-8, np.zeros_like(img), fov=(160, 140))
    faces = set()
    for k, ray in enumerate(r["rays"]):
        face, i, j = _cube(R0[k, 5:8], N)
        faces.add(face)
        for c, lm in enumerate(LAM):
            f = ray[4] / lm
            want = dark["rays"][k, 8 + 2 * c] + _sky_I((face * 40, i * 20, j * 20), f) * math.exp(-ray[9 + 2 * c]) / f ** 3
            assert ray[8 + 2 * c] == pytest.approx(want, rel=1e-12)
    assert len(faces) == 5  # all but the face behind the camera


def test_bundled_beam_of_a_uniform_sky(cart):
    sky = np.full((24, 32, 3), 160, dtype=np.uint8)
    _, R0, single = _render_cart(cart, 10, 1e-4, 1e-4, sky, lam=[450.0])
    _, _, beam = _render_cart(cart, 10, 1e-4, 1e-4, sky, lam=[450.0], ray_bundles=True)
    assert beam["rays"].shape[1] == single["rays"].shape[1] + 11
    # a pure-sky ray: the beam's weight sum times the texel intensity
    k = int(np.argmin(single["rays"][:, 9]))
    assert single["rays"][k, 9] == 0.0
    ray, b = beam["rays"][k], beam["rays"][k][-11:]
    dp = b[10] * ray[4]
    m1, m2 = math.hypot(b[5], b[6]), math.hypot(b[7], b[8])
    mu = b[4] + math.atan2(b[5] * b[8] + b[6] * b[7], b[5] * b[7] - b[6] * b[8]) / 2
    sbs = math.atan(2 / (8 + 1)) / 2
    sx, sy, th = O.distort_gauss(dp * (m1 + m2), dp * (m1 - m2), mu, 1 / (sbs + 1e-24), 1 / (sbs + 1e-24))
    h = min(math.ceil(max(sx, sy)), 10)
    wsum = sum(O.d2_gauss(x, y, sx, sy, th) for x in range(-h, h + 1) for y in range(-h, h + 1))
    assert beam["rays"][k, 8] == pytest.approx(wsum * single["rays"][k, 8], rel=1e-12)


# ---------------------------------------------------------------- Synchro_Test
def test_synchro_field_and_pitch_angle(synch):
    sc = synch[0]
    for k in ([1.0, 0.6, 0.0, 0.8], [1.0, 0.0, 0.28, 0.96], [2.0, 2.0, 0.0, 0.0]):
        ray = np.array([0.0, 0.3, -0.2, 5.0] + k)
        fp, u = O.fluid(sc, ray)
        assert np.array_equal(u, [1.0, 0.0, 0.0, 0.0])
        assert fp[0] == 2.5 and fp[1] == 30.0 and fp[2] == 1e9
        assert fp[3] == pytest.approx(F32(1e7) * 100.0, rel=1e-15)  # n_e = ne_0 b^2
        assert fp[4] == pytest.approx((k[3] / k[0]) ** 2, rel=1e-15, abs=1e-300)
        assert fp[5] == pytest.approx(F32(9.34e-11) * 10.0, rel=1e-15)
    # along the field: sin(theta) = 0 exactly, no floor, no emission or absorption
    a, j = O.radiative(sc, [0, 0.3, 0.2, 1.0, 1.0, 0.0, 0.0, 1.0], 1 / 550.0)
    assert a == 0.0 and j == 0.0
    # outside the cylinder nothing
    a, j = O.radiative(sc, [0, 0.8, 0.7, 1.0, 1.0, 0.6, 0.0, 0.8], 1 / 550.0)
    assert a == 0.0 and j == 0.0


def test_synchro_power_law(synch):
    sc = synch[0]
    ray = np.array([0.0, 0.3, -0.2, 5.0, 1.0, 0.6, 0.0, 0.8])
    nus = np.array([1 / 900.0, 1 / 600.0, 1 / 400.0])
    aj = np.array([O.radiative(sc, ray, nu) for nu in nus])
    lr = np.log(nus[1:] / nus[:-1])
    assert np.allclose(np.log(aj[1:, 1] / aj[:-1, 1]) / lr, -0.75, rtol=0, atol=1e-10)
    assert np.allclose(np.log(aj[1:, 0] / aj[:-1, 0]) / lr, -3.25, rtol=0, atol=1e-10)
    fp, _ = O.fluid(sc, ray)
    c2 = fp[4]
    # sin theta dependence: j ~ sin^((p+1)/2) Phi, a ~ sin^((p+2)/2) Phi (Rybicki & Lightman)
    ray2 = np.array([0.0, 0.3, -0.2, 5.0, 1.0, 0.28, 0.0, 0.96])
    c2b = O.fluid(sc, ray2)[0][4]
    a2, j2 = O.radiative(sc, ray2, nus[1])
    phi = lambda c: (1 - F32(0.99) * c) ** (-1.25)  # noqa: E731
    sr = math.sqrt((1 - c2b) / (1 - c2))
    assert j2 / aj[1, 1] == pytest.approx(sr ** 1.75 * phi(c2b) / phi(c2), rel=1e-12)
    assert a2 / aj[1, 0] == pytest.approx(sr ** 2.25 * phi(c2b) / phi(c2), rel=1e-12)


def test_synchro_cylinder_transfer(synch):
    sc, pos, d, p, beta = synch
    cam = abi.make_camera(pos, d, beta, p, 12, 85, 60)
    lam = [450.0, 650.0]
    r = O.render(sc, cam, abi.make_integrator(1e-10, 1e-30, 0.1, 1e5), lam, nthreads=4, want_rays=True)
    R0 = O.camera_rays(sc, cam)
    rays = r["rays"]
    assert np.array_equal(rays[:, 4:8], R0[:, 4:8])
    assert np.all(r["cls"] == 2)
    n = -R0[:, 5:8] / np.linalg.norm(R0[:, 5:8], axis=1)[:, None]
    s = 1 / np.hypot(n[:, 0], n[:, 1])  # from the origin to the cylinder wall
    for k in range(len(rays)):
        for c, lm in enumerate(lam):
            fsh = R0[k, 4]
            nu = fsh / lm
            a, j = O.radiative(sc, np.r_[0.0, 0.0, 0.0, 0.0, R0[k, 4:8]], nu)
            tau = a * s[k] / fsh ** 2
            assert rays[k, 9 + 2 * c] == pytest.approx(tau, rel=1e-6)
            assert rays[k, 8 + 2 * c] == pytest.approx(-j / (a * nu ** 3) * math.expm1(-tau), rel=1e-6)


def test_committed_fixtures_are_the_oracles():
    """The GPU parity fixtures of these scenes are what the oracle computes now."""
    import sys as _sys
    import os as _os
    _sys.path.insert(0, _os.path.join(_os.path.dirname(_os.path.abspath(__file__)), "golden"))
    from make_golden import CART_FIXTURE, cart_fixture_name, cart_workload, synchro_fixture_name, synchro_workload
    for bundles in (False, True):
        sc, cam, it, lam, sky = cart_workload(*CART_FIXTURE, bundles=bundles)
        O.load_skybox(1, sky)
        r = O.render(sc, cam, it, lam, nthreads=4, want_rays=True)
        gold = np.load(cart_fixture_name(bundles))
        for k in ("xyz", "cls", "steps", "rays"):
            assert np.array_equal(r[k], gold[k]), k
    sc, cam, it, lam = synchro_workload()
    r = O.render(sc, cam, it, lam, nthreads=4, want_rays=True)
    gold = np.load(synchro_fixture_name())
    for k in ("xyz", "cls", "steps", "rays"):
        assert np.array_equal(r[k], gold[k]), k
