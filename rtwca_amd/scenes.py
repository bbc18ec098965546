"""Scene presets: the reference's initialize_world cases and the benchmark configs.

`initialize_world(name)` mirrors main/initialization.jl:78-172: it returns the
scene (instead of `include`-ing a *.source directory) and the default camera
arguments (position, direction, pointing, beta) for that scene.

`config(n)` returns the five benchmark workloads (scene, camera, integrator,
wavelengths).  Values that the reference does not fix for these new scenes are
decided here, once:
  * cfg1 sphere: absorption a_nu = 100 per unit length, so the unit sphere is
    optically thick (tau >= -log(atol) = 23 within a chord of 0.23);
  * cfg2 disk: the reference's theta wedge [15pi/32, 17pi/32] with a_nu = 1;
  * cfg3/4 thick disk: r from the reference-formula ISCO to 12, H = 0.1 r,
    truncated at |z| <= 4H, kappa = 0.2, tolerance rtol = atol = 1e-6 (cfg3),
    1e-8 (cfg4); cfg4, the differentiable workload, softens the disk edges over
    DIFF_EDGE_WIDTH so that its image is C^1 in the parameters;
  * cfg5 torus + jet: see `kerr_synch`; 64 static cameras at r = 50 on a circle
    tilted 60 degrees from the equator (`inclined_orbit_cameras`), FOV 60/60,
    rtol = atol = 1e-6, max_steps = 2e4.
"""
import cmath
import math
from fractions import Fraction

from . import abi

PI = math.pi


def wavelengths(lo=380, hi=780, n=64):
    """range(lo, hi, length=n) as Julia computes it (correctly rounded nodes)."""
    if n == 1:
        return [float(lo)]
    step = Fraction(hi - lo, n - 1)
    return [float(Fraction(lo) + step * k) for k in range(n)]


def f32(x):
    import struct
    return struct.unpack("f", struct.pack("f", x))[0]


def kerr_isco_ref(a, M=1.0, f32_mode=False):
    """calc_orb_limits prograde ISCO (main/Kerr.source/source_geometry.jl:20-26),
    including its (1 +- chi)^(3/2) exponent."""
    if f32_mode:
        import numpy as np
        af, Mf = np.float32(a), np.float32(M)
        rs = np.float32(2) * Mf
        chi = np.float32(2) * af / rs
        Z1 = 1.0 + float(np.float32(1) - chi * chi) ** (1 / 3) * (float(np.float32(1) + chi) ** 1.5 + float(np.float32(1) - chi) ** 1.5)
        Z2 = math.sqrt(float(np.float32(3) * (chi * chi)) + Z1 * Z1)
        return float(np.float32(float(Mf) * (3 + Z2 - math.sqrt((3 - Z1) * (3 + Z1 + 2 * Z2)))))
    rs = 2 * M
    chi = 2 * a / rs
    Z1 = 1 + (1 - chi * chi) ** (1 / 3) * ((1 + chi) ** 1.5 + (1 - chi) ** 1.5)
    Z2 = math.sqrt(3 * (chi * chi) + Z1 * Z1)
    return M * (3 + Z2 - math.sqrt((3 - Z1) * (3 + Z1 + 2 * Z2)))


# radial edge width (M) of the softened thick disk of the differentiable workload
DIFF_EDGE_WIDTH = 1.0


def kerr_isco_ref_spin_slope(a, M=1.0):
    """d/da of kerr_isco_ref(a, M) (FP64 form), by complex step (the formula is
    analytic for 0 < a < M)."""
    h = 1e-20
    chi = 2 * complex(a, h) / (2 * M)
    Z1 = 1 + (1 - chi * chi) ** (1 / 3) * ((1 + chi) ** 1.5 + (1 - chi) ** 1.5)
    Z2 = cmath.sqrt(3 * (chi * chi) + Z1 * Z1)
    return (M * (3 + Z2 - cmath.sqrt((3 - Z1) * (3 + Z1 + 2 * Z2)))).imag / h


def kerr_circ_beta(r, a, M=1.0):
    """calc_circ_geod_beta (main/Kerr.source/source_geometry.jl:44-49)."""
    rs = 2 * M
    D = r * r - r * rs + a * a
    return math.sqrt(M) * (r * r - 2 * a * math.sqrt(M * r) + a * a) / (math.sqrt(D) * (r ** 1.5 + a * math.sqrt(M)))


def _base(kind):
    s = abi.Scene()
    s.kind = kind
    s.flags = 0
    s.mass = 1.0
    s.spin = 0.0
    s.temperature = 5778.0
    s.absorb = 1.0
    s.density = 1.0
    s.r_in = 0.0
    s.r_out = 1.0
    s.th_lo = 0.0
    s.th_hi = PI
    s.h_over_r = 0.1
    s.z_cut = 4.0
    s.r_escape = 100.0
    s.bfield = 10.0
    s.theta_e = 10.0
    s.jet_half_angle = 0.0
    s.r_jet = 0.0
    s.torus_r_max = 12.0
    return s


def minkowski_sphere(temperature=5778.0, absorb=1.0, radius=1.0):
    """main/Minkowski_Sphere_Spherical.source: emitting ball r <= 1, a_nu = 1, T = 5778 K
    (source_model.jl:224,240,255)."""
    s = _base(abi.SCENE_MINKOWSKI_SPHERE)
    s.temperature = temperature
    s.absorb = absorb
    s.r_out = radius
    return s


def schwarzschild_disk(temperature=5778.0, absorb=1.0, r_in=6.0, r_out=12.0):
    """main/Schwarzschild_Simple_Disk_Spherical.source: wedge 6 <= r <= 12,
    15pi/32 <= theta <= 17pi/32 (source_model.jl:273-280), a_nu = 1, T = 5778 K."""
    s = _base(abi.SCENE_SCHWARZSCHILD_DISK)
    s.temperature = temperature
    s.absorb = absorb
    s.r_in, s.r_out = r_in, r_out
    s.th_lo, s.th_hi = 15 * PI / 32, 17 * PI / 32
    return s


def kerr_disk(spin=0.998, ref_f32=True):
    """main/Kerr.source: alpha = 0.998f0 (source_geometry.jl:13), wedge from the
    reference-formula ISCO to r = 12, 31pi/64 <= theta <= 33pi/64 (source_model.jl:73-80),
    a_nu = 1f-1, T = 5778 K.  ref_f32 keeps the Float32 evaluation of the reference."""
    s = _base(abi.SCENE_KERR_DISK)
    s.flags = abi.SCENE_REF_F32 if ref_f32 else 0
    s.spin = f32(spin) if ref_f32 else spin
    s.absorb = f32(0.1) if ref_f32 else 0.1
    s.r_in = kerr_isco_ref(spin, 1.0, ref_f32)
    s.r_out = 12.0
    s.th_lo, s.th_hi = 31 * PI / 64, 33 * PI / 64
    return s


def kerr_thick_ff(spin=0.9, temperature=1e4, density=1.0, kappa=0.2, r_in=None, r_out=12.0,
                  h_over_r=0.1, z_cut=4.0, gaunt_fp64=False, edge_width=0.0):
    """Kerr geometry (main/Kerr.source/source_geometry.jl in FP64) with a Gaussian thick
    disk rho = rho_s exp(-z^2 / 2H^2), H = 0.1 r, |z| <= 4H, radiating free-free with the
    van Hoof Gaunt factor (main/emission_functions.jl:54-78) and Kirchhoff's law.
    edge_width > 0 softens the edges (quintic smootherstep windows of that width at
    the radial edges and of one scale height below the vertical cut); with the
    default inner edge (the reference-formula ISCO of the spin) the scene also
    carries d r_in / d spin."""
    s = _base(abi.SCENE_KERR_THICK_FF)
    s.spin = spin
    s.temperature = temperature
    s.density = density
    s.absorb = kappa
    s.r_in = kerr_isco_ref(spin) if r_in is None else r_in
    s.r_out = r_out
    s.h_over_r = h_over_r
    s.z_cut = z_cut
    s.edge_width = edge_width
    if edge_width > 0 and r_in is None:
        s.r_in_spin_slope = kerr_isco_ref_spin_slope(spin)
    if gaunt_fp64:
        s.flags |= abi.SCENE_GAUNT_FP64
    return s


def powehi_m87(ref_f32=True):
    """main/Powehi.source, the reference's M87* model: Kerr geometry with alpha = 0.9f0
    (source_geometry.jl:12); a wedge disk from the reference-formula ISCO to r = 12,
    31pi/64 <= theta <= 33pi/64, blackbody at 5778 K with a_nu = 1f-2
    (source_model.jl:172,190-194,231); a conical jet r >= 10 outside 5..175 degrees
    with power-law synchrotron in B^phi = 10 G (:51-57, :233-238)."""
    s = _base(abi.SCENE_POWEHI_M87)
    s.flags = abi.SCENE_REF_F32 if ref_f32 else 0
    s.spin = f32(0.9) if ref_f32 else 0.9
    s.temperature = 5778.0
    s.absorb = f32(1e-2)
    s.r_in = kerr_isco_ref(0.9, 1.0, ref_f32)
    s.r_out = 12.0
    s.th_lo, s.th_hi = 31 * PI / 64, 33 * PI / 64
    s.bfield = 10.0
    s.r_jet = 10.0
    return s


def wormhole():
    """main/Wormhole.source: the Dneg wormhole of Interstellar (rho = 1, half-length
    0.005, lensing width 0.05; source_geometry.jl:7-13), no matter; rays that leave
    through r(l) > 100 take the skybox of their side (needs both skyboxes loaded)."""
    s = _base(abi.SCENE_WORMHOLE)
    s.r_escape = 100.0
    return s


def minkowski_cart_sphere(temperature=5778.0, absorb=1.0, radius=1.0):
    """main/Minkowski_Sphere_Cartesian.source: flat space in (t, x, y, z), the unit
    sphere about the origin, a_nu = 1, T = 5778 K (source_model.jl:51-89); rays stop
    100 from the origin (:158-172) and every ray adds skybox 1 in its direction
    (:406-431; needs skybox 1 loaded)."""
    s = _base(abi.SCENE_MINKOWSKI_CART_SPHERE)
    s.temperature = temperature
    s.absorb = absorb
    s.r_out = radius
    s.r_escape = 100.0
    return s


def synchro_test(bfield=10.0, radius=1.0):
    """main/Synchro_Test.source: flat space in (t, x, y, z), a static cylinder
    sqrt(x^2 + y^2) <= 1 (source_model.jl:216-223) in the uniform field B^z = 10 G
    (:50-56) radiating power-law synchrotron (p = 2.5, gamma 30 .. 1e9, with the
    anisotropy of the M87* jet; :152-213); rays stop 100 from the origin."""
    s = _base(abi.SCENE_SYNCHRO_TEST)
    s.bfield = bfield
    s.r_out = radius
    s.r_escape = 100.0
    return s


def kerr_synch(spin=0.94, r_in=6.0, r_max=12.0, density=1e45, jet_ratio=1e-4, jet_half_angle_deg=10.0,
               r_jet=10.0, bfield=10.0, theta_e=10.0):
    """Kerr geometry (main/Kerr.source/source_geometry.jl in FP64) with a Fishbone-Moncrief
    torus (constant l = u_phi u^t of the circular orbit at the pressure maximum r_max, inner
    edge r_in) and a conical jet (|cos theta| >= cos(half-angle), r >= r_jet, n_e ~ r^-2 with
    the density going to zero continuously at the wall and the base, radial outflow at
    Lorentz factor 2), thermal electrons at theta_e in B = bfield Gauss radiating thermal
    synchrotron (Leung, Gammie & Noble 2011) with Kirchhoff absorption.  The jet base
    radius r_jet = 10 is the reference's M87* jet base (main/Powehi.source/source_model.jl:
    233-237); lengths are in units of M87*'s M.

    density is n_e (cm^-3) at the torus pressure maximum, jet_ratio the jet's density scale
    relative to it.  Optical frequencies lie deep in the exponential tail of this electron
    population ((nu / nu_s)^(1/3) ~ 100 at 550 nm, nu_s = 0.54 GHz): the opacity falls by
    ~11 decades from 780 to 380 nm and Doppler factors of the torus and jet flows move it
    by up to e^(+-30).  The defaults give the torus a median optical depth of ~3 at 780 nm
    (up to ~1e5 for edge-on, Doppler-boosted chords), ~1e-1..1e2 at 550 nm and < 1e-3 at
    380 nm, and a jet that carries a tenth to a half of the flux in views near its axis.
    The density vanishes continuously at every matter boundary: a jump in an opacity this
    steep makes an RK45 stage sum overflow exp(-tau)."""
    s = _base(abi.SCENE_KERR_SYNCH)
    s.spin = spin
    s.r_in = r_in
    s.r_out = r_in
    s.torus_r_max = r_max
    s.density = density
    s.absorb = jet_ratio
    s.jet_half_angle = jet_half_angle_deg * PI / 180
    s.r_jet = r_jet
    s.bfield = bfield
    s.theta_e = theta_e
    return s


def inclined_orbit_cameras(n_poses=64, r=50.0, incl_deg=60.0, x_pix=512, fov=60.0):
    """Static cameras at radius r on a circle whose plane is tilted by incl_deg from the
    equatorial plane, at orbital phases 2 pi k / n_poses, each looking at the hole."""
    i = incl_deg * PI / 180
    cams = []
    for k in range(n_poses):
        psi = 2 * PI * k / n_poses
        z = math.sin(psi) * math.sin(i)
        th = math.acos(max(-1.0, min(1.0, z)))
        ph = math.atan2(math.sin(psi) * math.cos(i), math.cos(psi))
        cams.append(abi.make_camera([0, r, th, ph], [0, 0], 0.0, [-PI / 2, 0, 0], x_pix, fov, fov))
    return cams


def initialize_world(name):
    """main/initialization.jl:78-172 (scenes available natively)."""
    n = name.lower()
    if n == "schwarzschild_t,r,θ,ϕ_simpledisk":
        return schwarzschild_disk(), [0, 25, PI / 3, 0], [PI / 2, 0], [0.0, 0.0, -PI / 2], 0.0
    if n == "kerr_boyer_lindquist":
        return kerr_disk(), [0, 25, 0.52 * PI, 0], [PI / 2, 0], [0.0, 0.0, -PI / 2], 0.0
    if n == "m87":  # initialization.jl:100-108
        return powehi_m87(), [0, 35, 0.3, 0], [PI / 2, PI / 6], [0.0, 0.0, PI], 0.0
    if n == "wormhole":  # initialization.jl:140-149
        return wormhole(), [0, 5, PI / 2, 0], [PI / 2, 0], [0.0, 0.0, -PI / 2], 0.0
    if n == "minkowski_t,r,θ,ϕ_sphere":
        return minkowski_sphere(), [0, 4, PI / 2, 0], [PI / 2, 0], [0.0, 0.0, -PI / 2], 0.0
    if n == "synch":  # initialization.jl:120-128
        return synchro_test(), [0.0, 0.0, 0.0, 0.0], [0.0, 0.0], [PI / 8, 0.0, PI], 0.0
    if n == "minkowski_t,x,y,z_sphere":  # initialization.jl:130-138
        return minkowski_cart_sphere(), [0.0, 0.0, 0.0, -4.0], [0.0, 0.0], [0.0, 0.0, 0.0], 0.0
    raise ValueError("scene %r is not provided by rtwca_amd" % name)


class Workload:
    def __init__(self, name, scene, cameras, integ, lam):
        self.name = name
        self.scene = scene
        self.cameras = cameras
        self.integ = integ
        self.lam = lam

    @property
    def x_pix(self):
        return self.cameras[0].x_pix


def config(idx, x_pix=None, precision=abi.PREC_FP64, schedule=abi.SCHED_PERSISTENT, n_lambda=64):
    """Benchmark workloads 1..5 (see the module docstring for the choices made)."""
    lam = wavelengths(380, 780, n_lambda)
    if idx == 1:
        sc = minkowski_sphere(temperature=5772.0, absorb=100.0)
        cam = abi.make_camera([0, 10, PI / 2, 0], [PI / 2, 0], 0.0, [0, 0, -PI / 2],
                              x_pix or 256, 60, 60)
        it = abi.make_integrator(1e-8, 1e-10, 0.1, 1e4, precision, schedule)
        return Workload("minkowski_sphere_256", sc, [cam], it, lam)
    if idx == 2:
        sc = schwarzschild_disk(temperature=1e4, absorb=1.0, r_in=6.0, r_out=20.0)
        cam = abi.make_camera([0, 30, 80 * PI / 180, 0], [PI / 2, 0], 0.0, [0, 0, -PI / 2],
                              x_pix or 512, 40, 40)
        it = abi.make_integrator(1e-8, 1e-10, 0.1, 1e4, precision, schedule)
        return Workload("schwarzschild_disk_512", sc, [cam], it, lam)
    if idx == 3:
        sc = kerr_thick_ff(spin=0.9)
        beta = kerr_circ_beta(5.0, 0.9)
        cam = abi.make_camera([0, 5, PI / 2, 0], [0, 0], beta, [-PI / 2, 0, 0],
                              x_pix or 1024, 90, 90)
        it = abi.make_integrator(1e-6, 1e-6, 0.1, 1e4, precision, schedule, 3.0)
        return Workload("kerr_a0.9_thick_ff_1024", sc, [cam], it, lam)
    if idx == 4:
        sc = kerr_thick_ff(spin=0.9, gaunt_fp64=True, edge_width=DIFF_EDGE_WIDTH)
        cam = abi.make_camera([0, 10, 60 * PI / 180, 0], [0, 0], 0.0, [-PI / 2, 0, 0],
                              x_pix or 256, 60, 60)
        it = abi.make_integrator(1e-8, 1e-8, 0.1, 1e4, precision, schedule)
        return Workload("kerr_thick_ff_adjoint_256", sc, [cam], it, lam)
    if idx == 5:
        sc = kerr_synch()
        cams = inclined_orbit_cameras(64, 50.0, 60.0, x_pix or 512, 60.0)
        it = abi.make_integrator(1e-6, 1e-6, 0.1, 2e4, precision, schedule, 3.0)
        return Workload("kerr_a0.94_synch_torus_jet_64x512", sc, cams, it, lam)
    raise ValueError("unknown config %r" % idx)
