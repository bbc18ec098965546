"""Pins of the oracle's thermal-synchrotron torus + jet scene -- CPU only.

The scene has no reference counterpart (the nearest is main/Powehi.source, with
power-law synchrotron), so every part is pinned by physics that a transcription
mistake would break:

  part                                pinned by
  ----------------------------------  -------------------------------------------------
  torus angular momentum l            a = 0 closed form sqrt(M r) / (1 - 3M/r); for a != 0
                                      u_phi u^t of the circular orbit built from the metric
  torus potential ln h                zero at the inner edge, stationary at the pressure
                                      maximum, and relativistic hydrostatic equilibrium
                                      d ln h = 1/2 (d g_ab) u^a u^b at interior points
  torus / jet / FIDO 4-velocities     unit timelike; u_phi u^t = l; jet: u_phi = 0,
                                      Lorentz factor 2 and speed sqrt(3)/2 in the ZAMO frame
  densities                           normalisation at the pressure maximum; jet profile
                                      values on the axis, at the wall and at the base
  emission / absorption               j / a = Planck (Kirchhoff) and a against Leung,
                                      Gammie & Noble (2011) eq. 31 evaluated with
                                      scipy.constants and scipy.special.kv
"""
import math

import numpy as np
import pytest

import oracle as O
from rtwca_amd import abi, scenes

PI = math.pi


def _scene(spin=0.94, **kw):
    return scenes.kerr_synch(spin=spin, **kw)


def _lnh(sc, r, th):
    return O.synch_point(sc, [0.0, r, th, 0.0])[2]


def test_torus_angular_momentum_schwarzschild_closed_form():
    for rmax in (8.0, 12.0, 20.0):
        sc = _scene(spin=0.0, r_in=5.0, r_max=rmax)
        ell = O.synch_derived(sc)["ell"]
        assert ell == pytest.approx(math.sqrt(rmax) / (1 - 3 / rmax), rel=1e-13)


def test_torus_angular_momentum_from_the_metric():
    sc = _scene()
    r = sc.torus_r_max
    g = O.metric(sc, [0, r, PI / 2, 0])
    om = 1.0 / (r ** 1.5 + sc.spin)  # prograde circular orbit, Bardeen-Press-Teukolsky
    ut = 1.0 / math.sqrt(-(g[0, 0] + 2 * g[0, 3] * om + g[3, 3] * om * om))
    uphi_low = (g[0, 3] + g[3, 3] * om) * ut
    assert O.synch_derived(sc)["ell"] == pytest.approx(uphi_low * ut, rel=1e-12)


def test_torus_potential_edges_and_maximum():
    sc = _scene()
    d = O.synch_derived(sc)
    assert abs(_lnh(sc, 6.0, PI / 2)) < 1e-14
    assert _lnh(sc, 12.0, PI / 2) == pytest.approx(math.log1p(d["hm1max"]), rel=1e-13)
    e = 1e-4
    slope = (_lnh(sc, 12.0 + e, PI / 2) - _lnh(sc, 12.0 - e, PI / 2)) / (2 * e)
    assert abs(slope) < 1e-9
    # bound torus: ln h < 0 far out and at the inner edge's inside
    assert _lnh(sc, 50.0, PI / 2) < 0 and _lnh(sc, 5.9, PI / 2) < 0
    # the density is (h - 1)^3 normalised at the maximum
    reg, dens, _ = O.synch_point(sc, [0, 12.0, PI / 2, 0])
    assert reg == 1 and dens == pytest.approx(1.0, rel=1e-12)


@pytest.mark.parametrize("r,th", [(8.0, 1.45), (12.0, 1.3), (12.0, 1.1), (20.0, 1.2), (30.0, 1.5)])
def test_torus_hydrostatic_equilibrium(r, th):
    """For a stationary circular flow a_mu = -1/2 d_mu g_ab u^a u^b; the Euler equation of
    the isentropic torus gives d_mu ln h = -a_mu."""
    sc = _scene()
    x = np.array([0.0, r, th, 0.0])
    assert O.synch_point(sc, x)[0] == 1
    u = O.source_velocity(sc, x)
    for mu, e in ((1, 1e-5), (2, 1e-6)):
        xp, xm = x.copy(), x.copy()
        xp[mu] += e
        xm[mu] -= e
        dlnh = (_lnh(sc, xp[1], xp[2]) - _lnh(sc, xm[1], xm[2])) / (2 * e)
        dg = (O.metric(sc, xp) - O.metric(sc, xm)) / (2 * e)
        rhs = 0.5 * u @ dg @ u
        assert dlnh == pytest.approx(rhs, rel=1e-6, abs=1e-11)


def test_four_velocities():
    sc = _scene()
    ell = O.synch_derived(sc)["ell"]
    # torus
    for x in ([0, 12.0, 1.3, 0.4], [0, 8.0, 1.5, 2.0]):
        g = O.metric(sc, x)
        u = O.source_velocity(sc, x)
        assert u @ g @ u == pytest.approx(-1.0, abs=1e-12)
        assert u[1] == 0 and u[2] == 0
        assert (g @ u)[3] * u[0] == pytest.approx(ell, rel=1e-12)
    # jet: radial outflow with Lorentz factor 2 in the ZAMO frame
    x = [0, 30.0, 0.05, 1.0]
    assert O.synch_point(sc, x)[0] == 2
    g = O.metric(sc, x)
    u = O.source_velocity(sc, x)
    assert u @ g @ u == pytest.approx(-1.0, abs=1e-12)
    assert abs((g @ u)[3]) < 1e-12 * abs(u[0])
    lapse = math.sqrt(-1.0 / np.linalg.inv(g)[0, 0])
    assert lapse * u[0] == pytest.approx(2.0, rel=1e-12)
    assert math.sqrt(g[1, 1]) * u[1] / (lapse * u[0]) == pytest.approx(math.sqrt(3) / 2, rel=1e-12)
    # vacuum: the FIDO
    x = [0, 30.0, 0.8, 0.0]
    assert O.synch_point(sc, x)[0] == 0
    g = O.metric(sc, x)
    u = O.source_velocity(sc, x)
    assert u @ g @ u == pytest.approx(-1.0, abs=1e-12)
    assert abs((g @ u)[3]) < 1e-12 * abs(u[0])


def test_jet_density_profile():
    sc = _scene(jet_ratio=0.3, r_jet=10.0, jet_half_angle_deg=10.0)
    reg, d, _ = O.synch_point(sc, [0, 20.0, 0.0, 0])
    assert reg == 2 and d == pytest.approx(0.3 * 0.25 * 0.5, rel=1e-14)
    reg, d, _ = O.synch_point(sc, [0, 20.0, PI, 0])  # the counter-jet
    assert reg == 2 and d == pytest.approx(0.3 * 0.25 * 0.5, rel=1e-14)
    reg, d, _ = O.synch_point(sc, [0, 10.0, 0.01, 0])  # base
    assert reg == 2 and d == 0
    th_w = math.radians(10.0)
    reg, d, _ = O.synch_point(sc, [0, 20.0, th_w * (1 - 1e-9), 0])  # just inside the wall
    assert reg == 2 and d < 1e-15
    assert O.synch_point(sc, [0, 20.0, th_w * 1.01, 0])[0] == 0
    assert O.synch_point(sc, [0, 9.0, 0.01, 0])[0] == 0


def test_coefficients_kirchhoff_and_leung_formula():
    sp = pytest.importorskip("scipy.special")
    const = pytest.importorskip("scipy.constants")
    sc = _scene()
    d = O.synch_derived(sc)
    x = [0, 12.0, PI / 2, 0]
    e_cgs = const.e * const.c * 10  # statcoulomb
    me, c, h, k = const.m_e * 1e3, const.c * 1e2, const.h * 1e7, const.k * 1e7
    Te = sc.theta_e * me * c * c / k
    assert d["T_e"] == pytest.approx(Te, rel=1e-8)
    nus = 2.0 / 9.0 * e_cgs * sc.bfield / (2 * PI * me * c) * sc.theta_e ** 2 * math.sin(PI / 3)
    assert d["nu_s"] == pytest.approx(nus, rel=1e-8)
    ne = sc.density  # at the pressure maximum n_e = density
    for lam in (380.0, 550.0, 780.0, 2000.0, 1e6):
        nu = 1.0 / lam  # nm^-1 in the fluid frame
        a, j = O.radiative(sc, x, nu)
        # Planck in the reference's units (2hc 1e27 nu^3 / expm1(hc 1e9 nu / k T), nu in
        # nm^-1); expm1 because h nu / k T_e ~ 1e-7 (the reference's exp(x) - 1 in
        # calc_planck would lose 9 digits here)
        planck = (2 * const.h * const.c * 1e27 * nu ** 3 /
                  math.expm1(const.h * const.c * 1e9 * nu / (const.k * d["T_e"])))
        assert j / a == pytest.approx(planck, rel=1e-12)
        f = const.c / (lam * 1e-9)  # Hz
        X = f / nus
        jnu = (ne * math.sqrt(2) * PI * e_cgs ** 2 * nus / (3 * sp.kv(2, 1 / sc.theta_e) * c)
               * (X ** 0.5 + 2 ** (11 / 12) * X ** (1 / 6)) ** 2 * math.exp(-X ** (1 / 3)))
        Bnu = 2 * h * f ** 3 / c ** 2 / math.expm1(h * f / (k * Te))
        a_cgs = jnu / Bnu  # per cm
        assert a == pytest.approx(a_cgs * 9.62e14, rel=1e-7)
