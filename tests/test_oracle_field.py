"""Pins of the oracle's magnetic fields: the trilinear field map
(InterpolatedBFieldMap + fieldMapXYZ) and the solenoid used to sample it
(SolenoidBField), and the library's map sampler against the oracle.

- the complete elliptic integrals against scipy, including the reference's
  use of k^2 as the modulus (SolenoidBField.cpp:111,148-150: Boost's
  ellint_1(k) takes the modulus, scipy's ellipk(m) the parameter m = k^2);
- the solenoid's normalisation B(0) = bMagCenter and its symmetries
  (Tests/UnitTests/Core/MagneticField/SolenoidBFieldTests.cpp:27-110);
- the map against closed forms: exact at grid points, exact (to rounding) for
  a field linear in x, y, z, the reference's validity domain (upper edge
  exclusive at the last grid point), and scipy's RegularGridInterpolator.
"""
import math

import numpy as np
import pytest
import scipy.interpolate
import scipy.special

from acts_amd.detector import TESLA, FieldMap


def test_elliptic_integrals_match_scipy(oracle_mod):
    for k in np.linspace(0.0, 0.999, 40):
        K, E = oracle_mod.elliptic(k)
        assert K == pytest.approx(scipy.special.ellipk(k * k), rel=1e-14, abs=1e-15)
        assert E == pytest.approx(scipy.special.ellipe(k * k), rel=1e-14, abs=1e-15)


def test_solenoid_uses_k2_as_modulus(oracle_mod):
    # B_z of one coil off axis, restated with scipy: K and E of modulus k^2,
    # i.e. scipy parameter m = k^4
    R, L, nc, b0 = 1200.0, 6000.0, 1, 2.0 * TESLA
    r, z = 700.0, 250.0
    B = oracle_mod.solenoid_eval(R, L, nc, b0, np.array([[r, 0.0, z]]))[0]
    # the scale normalises B(0) to b0: recompute it from the on-axis formula
    # of the single coil (coil at z = 0): B_z(0, 0) = scale / (2 R)
    scale = b0 * 2 * R
    k2 = 4 * R * r / ((R + r) ** 2 + z ** 2)
    k = math.sqrt(k2)
    Kq, Eq = scipy.special.ellipk(k2 * k2), scipy.special.ellipe(k2 * k2)
    bz = scale * k / (4 * math.pi * math.sqrt(R * r)) * (
        ((R + r) * k2 - 2 * r) / (2 * r * (1 - k2)) * Eq + Kq)
    br = scale * k * z / (4 * math.pi * math.sqrt(R * r ** 3)) * (
        (2 - k2) / (2 - 2 * k2) * Eq - Kq)
    assert B[2] == pytest.approx(bz, rel=1e-12)
    assert B[0] == pytest.approx(br, rel=1e-12)
    # and it is not the textbook field (modulus k)
    Kt, Et = scipy.special.ellipk(k2), scipy.special.ellipe(k2)
    bz_textbook = scale * k / (4 * math.pi * math.sqrt(R * r)) * (
        ((R + r) * k2 - 2 * r) / (2 * r * (1 - k2)) * Et + Kt)
    assert abs(B[2] - bz_textbook) > 1e-3 * abs(bz_textbook)


def test_solenoid_normalisation_and_symmetry(oracle_mod):
    R, L, nc, b0 = 1200.0, 6000.0, 100, 2.0 * TESLA
    B = oracle_mod.solenoid_eval(R, L, nc, b0, np.zeros((1, 3)))[0]
    assert np.linalg.norm(B) == pytest.approx(b0, abs=1e-6 * TESLA)
    rng = np.random.default_rng(1)
    pts = rng.uniform([-1900, -1900, -3900], [1900, 1900, 3900], size=(50, 3))
    Bp = oracle_mod.solenoid_eval(R, L, nc, b0, pts)
    Bm = oracle_mod.solenoid_eval(R, L, nc, b0, pts * np.array([1, 1, -1]))
    # B_z even and B_r odd in z
    assert Bm[:, 2] == pytest.approx(Bp[:, 2], rel=1e-9, abs=1e-14)
    assert Bm[:, :2] == pytest.approx(-Bp[:, :2], rel=1e-9, abs=1e-14)
    # the transverse field is radial
    cross = Bp[:, 0] * pts[:, 1] - Bp[:, 1] * pts[:, 0]
    assert np.abs(cross).max() < 1e-12 * np.abs(Bp).max() * 2000
    # field inside the solenoid of the order of the centre value (the k^2
    # modulus makes it differ from a physical solenoid)
    Bc = oracle_mod.solenoid_eval(R, L, nc, b0, np.array([[300.0, 0, 500.0]]))[0]
    assert Bc[2] == pytest.approx(b0, rel=0.1)


def test_library_sampler_matches_oracle_solenoid(oracle_mod):
    fm = FieldMap.solenoid(n=(9, 11, 13), point_min=(-1500, -1500, -3500),
                           point_max=(1500, 1500, 3500), n_coils=20)
    X, Y, Z = np.meshgrid(fm.points(0), fm.points(1), fm.points(2), indexing="ij")
    pts = np.stack([X, Y, Z], axis=-1).reshape(-1, 3)
    ref = oracle_mod.solenoid_eval(1200.0, 6000.0, 20, 2.0 * TESLA, pts)
    assert np.array_equal(fm.values, ref)


def linear_map(n=(5, 6, 7), lo=(-100.0, -50.0, -200.0), hi=(100.0, 70.0, 400.0)):
    pts = [lo[a] + np.arange(n[a]) * ((hi[a] - lo[a]) / (n[a] - 1)) for a in range(3)]
    X, Y, Z = np.meshgrid(*pts, indexing="ij")
    A = np.array([[0.3, -0.2, 0.1], [0.05, 0.4, -0.3], [-0.1, 0.2, 0.25]])
    c = np.array([1.0, -2.0, 0.5])
    P = np.stack([X, Y, Z], axis=-1).reshape(-1, 3)
    return FieldMap(n, lo, hi, P @ A.T + c), A, c, pts


def test_map_exact_at_grid_points_and_for_linear_fields(oracle_mod):
    fm, A, c, pts = linear_map()
    X, Y, Z = np.meshgrid(pts[0][:-1], pts[1][:-1], pts[2][:-1], indexing="ij")
    P = np.stack([X, Y, Z], axis=-1).reshape(-1, 3)
    B, st = oracle_mod.field_eval(fm, P)
    assert not st.any()
    sel = np.stack(np.meshgrid(range(4), range(5), range(6), indexing="ij"), -1).reshape(-1, 3)
    idx = (sel[:, 0] * 6 + sel[:, 1]) * 7 + sel[:, 2]
    assert np.array_equal(B, fm.values[idx])
    rng = np.random.default_rng(2)
    Q = rng.uniform([-100, -50, -200], [99.9, 69.9, 399.9], size=(500, 3))
    B, st = oracle_mod.field_eval(fm, Q)
    assert not st.any()
    assert B == pytest.approx(Q @ A.T + c, rel=1e-12, abs=1e-12)


def test_map_validity_domain(oracle_mod):
    fm, *_ = linear_map()
    last = np.array([100.0, 70.0, 400.0])
    inside = np.array([[0.0, 0.0, 0.0], [-100.0, -50.0, -200.0],
                       [np.nextafter(100.0, 0), 0, 0]])
    outside = np.array([[np.nextafter(-100.0, -200), 0, 0], [100.0, 0.0, 0.0],
                        [0.0, 70.0, 0.0], [0.0, 0.0, 400.0], [0.0, 0.0, 1e4],
                        [np.nan, 0.0, 0.0]])
    _, st = oracle_mod.field_eval(fm, inside)
    assert not st.any()
    _, st = oracle_mod.field_eval(fm, outside)
    assert (st == 0x501).all()
    assert last is not None


def test_map_against_scipy_interpolator(oracle_mod):
    fm = FieldMap.solenoid(n=(11, 11, 21), point_min=(-2000, -2000, -4000),
                           point_max=(2000, 2000, 4000), n_coils=20)
    grid = tuple(fm.points(a) for a in range(3))
    vals = fm.values.reshape(11, 11, 21, 3)
    rng = np.random.default_rng(4)
    Q = rng.uniform([-1999, -1999, -3999], [1599, 1599, 3599], size=(300, 3))
    B, st = oracle_mod.field_eval(fm, Q)
    assert not st.any()
    for comp in range(3):
        interp = scipy.interpolate.RegularGridInterpolator(grid, vals[..., comp])
        assert B[:, comp] == pytest.approx(interp(Q), rel=1e-10, abs=1e-16)
