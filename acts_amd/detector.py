"""Detector geometry, material and field descriptions for the fitter.

The geometry is a flat table of surfaces (acts_kf_surface): the direct
navigator only needs each track's ordered surface sequence, so there is no
tracking-geometry hierarchy here.  Builders cover the benchmark detectors:
a silicon barrel of cylinders and, optionally, endcap discs.
"""
from __future__ import annotations

from dataclasses import dataclass, field
import math

import ctypes as C

import numpy as np

from . import capi

# Units (Core/include/Acts/Definitions/Units.hpp)
MM = 1.0
CM = 10.0
M = 1000.0
GEV = 1.0
MEV = 1e-3
TESLA = 0.000299792458
NS = 299792458000.0 * 1e-9  # 1 ns in mm (c = 1)
UM = 1e-3

# PDG muon mass (float, ParticleHypothesis.hpp:113)
MUON_MASS = 0.1056583755


@dataclass
class Material:
    """Material::fromMolarDensity arguments (floats in the reference)."""
    x0: float
    l0: float
    ar: float
    z: float
    molar_rho: float


def silicon() -> Material:
    """Silicon as Tests/CommonHelpers/.../PredefinedMaterials.hpp:23-27."""
    return Material(x0=9.370 * CM, l0=46.52 * CM, ar=28.0855, z=14.0,
                    molar_rho=(2.329 / 28.0855) * 1.0 / (CM ** 3))


def beryllium() -> Material:
    """Beryllium as Tests/CommonHelpers/.../PredefinedMaterials.hpp:17-21."""
    return Material(x0=35.28 * CM, l0=42.10 * CM, ar=9.012, z=4.0,
                    molar_rho=(1.848 / 9.012) * 1.0 / (CM ** 3))


@dataclass
class SurfaceDesc:
    type: int
    rotation: np.ndarray  # 3x3, columns = local axes
    translation: np.ndarray
    bounds: tuple[float, float]
    material: int = -1
    sensitive: bool = True


@dataclass
class Detector:
    surfaces: list[SurfaceDesc] = field(default_factory=list)
    materials: list[tuple[Material, float]] = field(default_factory=list)
    name: str = ""
    # per surface (dimension, bound indices, sigmas) of its measurements;
    # None: 2D (loc0, loc1) with the generator's resolutions
    kinds: tuple | None = None

    def add_material(self, mat: Material, thickness: float) -> int:
        self.materials.append((mat, thickness))
        return len(self.materials) - 1

    def surface_table(self):
        arr = (capi.Surface * len(self.surfaces))()
        for i, s in enumerate(self.surfaces):
            arr[i].type = s.type
            arr[i].material = s.material
            arr[i].flags = capi.SURFACE_SENSITIVE if s.sensitive else 0
            rot = np.asarray(s.rotation, dtype=np.float64).reshape(9)
            for k in range(9):
                arr[i].rotation[k] = rot[k]
            for k in range(3):
                arr[i].translation[k] = float(s.translation[k])
            arr[i].bounds[0] = float(s.bounds[0])
            arr[i].bounds[1] = float(s.bounds[1])
        return arr

    def material_table(self):
        n = max(1, len(self.materials))
        arr = (capi.MaterialSlab * n)()
        for i, (m, t) in enumerate(self.materials):
            arr[i].x0 = m.x0
            arr[i].l0 = m.l0
            arr[i].ar = m.ar
            arr[i].z = m.z
            arr[i].molar_rho = m.molar_rho
            arr[i].thickness = t
        return arr


BARREL_RADII = (32.0, 72.0, 116.0, 172.0, 260.0, 360.0, 500.0, 660.0, 820.0, 1000.0)


def barrel_detector(radii=BARREL_RADII, half_length=3000.0, thickness=0.3,
                    material: Material | None = None) -> Detector:
    """Cylindrical silicon barrel: one cylinder per radius, centred on z."""
    det = Detector(name="barrel")
    mi = det.add_material(material or silicon(), thickness)
    for r in radii:
        det.surfaces.append(SurfaceDesc(capi.SURFACE_CYLINDER, np.eye(3),
                                        np.zeros(3), (r, half_length), mi, True))
    return det


def endcap_z_positions(z_min=600.0, z_max=2900.0, n=7) -> list[float]:
    return [z_min + (z_max - z_min) * i / (n - 1) for i in range(n)]


def barrel_endcap_detector(radii=BARREL_RADII, half_length=3000.0,
                           disc_z=None, disc_rmin=30.0, disc_rmax=1000.0,
                           thickness=0.3) -> Detector:
    """Barrel plus 2 x 7 endcap discs at z = +-(600 .. 2900) mm."""
    det = barrel_detector(radii, half_length, thickness)
    det.name = "barrel+endcaps"
    mi = 0
    zs = endcap_z_positions() if disc_z is None else list(disc_z)
    for side in (-1.0, 1.0):
        for z in zs:
            det.surfaces.append(SurfaceDesc(capi.SURFACE_DISC, np.eye(3),
                                            np.array([0.0, 0.0, side * z]),
                                            (disc_rmin, disc_rmax), mi, True))
    return det


def double_layer_detector(radii=BARREL_RADII, half_length=3000.0, disc_z=None,
                          disc_rmin=30.0, disc_rmax=1000.0, gap=2.0,
                          thickness=0.3) -> Detector:
    """Barrel of double layers (a second sensitive cylinder `gap` mm outside
    each layer) plus the 2x7 single endcap discs: up to 27 crossings."""
    det = Detector(name="double-layer barrel+endcaps")
    mi = det.add_material(silicon(), thickness)
    for r in radii:
        for dr in (0.0, gap):
            det.surfaces.append(SurfaceDesc(capi.SURFACE_CYLINDER, np.eye(3), np.zeros(3),
                                            (r + dr, half_length), mi, True))
    zs = endcap_z_positions() if disc_z is None else list(disc_z)
    for side in (-1.0, 1.0):
        for z in zs:
            det.surfaces.append(SurfaceDesc(capi.SURFACE_DISC, np.eye(3),
                                            np.array([0.0, 0.0, side * z]),
                                            (disc_rmin, disc_rmax), mi, True))
    return det


RAGGED_RADII = (32.0, 52.0, 72.0, 94.0, 116.0, 144.0, 172.0, 216.0, 260.0, 310.0, 360.0,
                430.0, 500.0, 660.0, 820.0)


def ragged_detector(radii=RAGGED_RADII, half_length=1000.0, gap=2.0,
                    disc_z=(1200.0, 1800.0, 2400.0), disc_rmin=30.0, disc_rmax=900.0,
                    thickness=0.3) -> Detector:
    """The ragged-batching detector: 15 barrel double layers (a second
    sensitive cylinder `gap` mm outside each, r = 32..822 mm) of half-length
    1 m, and 2 x 3 endcap discs at |z| = 1.2, 1.8, 2.4 m (r = 30..900 mm).
    A central track crosses all 30 cylinders; the further forward, the
    fewer barrel layers and the more discs it crosses, down to a few
    surfaces at |eta| = 3.5: with 5 % holes and the 6-measurement minimum,
    tracks carry 6 to 30 measurements (mean ~20)."""
    det = Detector(name="ragged: 15 double-layer barrel + 2x3 discs")
    mi = det.add_material(silicon(), thickness)
    for r in radii:
        for dr in (0.0, gap):
            det.surfaces.append(SurfaceDesc(capi.SURFACE_CYLINDER, np.eye(3), np.zeros(3),
                                            (r + dr, half_length), mi, True))
    for side in (-1.0, 1.0):
        for z in disc_z:
            det.surfaces.append(SurfaceDesc(capi.SURFACE_DISC, np.eye(3),
                                            np.array([0.0, 0.0, side * z]),
                                            (disc_rmin, disc_rmax), mi, True))
    return det


# measurement kinds of a telescope plane: (dimension, bound indices, sigmas)
PIXEL = (2, (0, 1), (25 * 1e-3, 50 * 1e-3))
STRIP0 = (1, (0, 0), (100 * 1e-3,))
STRIP1 = (1, (1, 1), (150 * 1e-3,))


def telescope_detector(shift=3000.0) -> Detector:
    """The reference's cubic test telescope (Tests/CommonHelpers/src/
    CubicTrackingGeometry.cpp:32-112): six planes normal to x at x = -2,
    -1, 1 -+ 0.001, 2 -+ 0.001 m, each 0.5 mm of beryllium, the last four
    rotated by +-0.026 rad about their normal (stereo strips); local x is
    global z and local y global y (rotation by 90 degrees about y).  The
    planes are shifted by `shift` along x, so that tracks from the origin
    start where the reference's start at x = -3 m does.  `kinds` gives the
    measurement of each plane as in FitterTestsCommon.hpp:145-155: 2D pixels
    (25 / 50 um) on the first two, 1D strips on loc0 (100 um) and loc1
    (150 um) alternating on the others."""
    det = Detector(name="telescope")
    mi = det.add_material(beryllium(), 0.5)
    xs = (-2000.0, -1000.0, 999.0, 1001.0, 1999.0, 2001.0)
    angles = (0.0, 0.0, 0.026, -0.026, 0.026, -0.026)
    ry = np.array([[0.0, 0.0, -1.0], [0.0, 1.0, 0.0], [1.0, 0.0, 0.0]])
    for x, a in zip(xs, angles):
        rz = np.array([[math.cos(a), -math.sin(a), 0.0], [math.sin(a), math.cos(a), 0.0],
                       [0.0, 0.0, 1.0]])
        det.surfaces.append(SurfaceDesc(capi.SURFACE_PLANE, ry @ rz,
                                        np.array([x + shift, 0.0, 0.0]), (0.0, 0.0), mi, True))
    det.kinds = (PIXEL, PIXEL, STRIP0, STRIP1, STRIP0, STRIP1)
    return det


@dataclass
class ConstantField:
    bx: float = 0.0
    by: float = 0.0
    bz: float = 2.0 * TESLA

    def struct(self) -> capi.Field:
        f = capi.Field()
        f.kind = capi.FIELD_CONSTANT
        f.constant[0], f.constant[1], f.constant[2] = self.bx, self.by, self.bz
        return f


class FieldMap:
    """Field values on the points of a regular xyz grid (the input of the
    reference's fieldMapXYZ, BFieldMapUtils.cpp:109-186): n[a] points per
    axis from point_min[a] to point_max[a], values[(ix*ny + iy)*nz + iz]."""

    def __init__(self, n, point_min, point_max, values: np.ndarray):
        self.n = tuple(int(v) for v in n)
        self.point_min = tuple(float(v) for v in point_min)
        self.point_max = tuple(float(v) for v in point_max)
        self.values = np.ascontiguousarray(values, dtype=np.float64).reshape(-1, 3)
        assert self.values.shape[0] == self.n[0] * self.n[1] * self.n[2]

    def struct(self) -> capi.Field:
        f = capi.Field()
        f.kind = capi.FIELD_GRID_XYZ
        for a in range(3):
            f.n[a] = self.n[a]
            f.point_min[a] = self.point_min[a]
            f.point_max[a] = self.point_max[a]
        f.values = self.values.ctypes.data_as(C.POINTER(C.c_double))
        return f

    def points(self, axis: int) -> np.ndarray:
        n, lo, hi = self.n[axis], self.point_min[axis], self.point_max[axis]
        return lo + np.arange(n) * ((hi - lo) / (n - 1))

    @classmethod
    def solenoid(cls, n=(201, 201, 401), point_min=(-2000.0, -2000.0, -4000.0),
                 point_max=(2000.0, 2000.0, 4000.0), radius=1200.0, length=6000.0,
                 n_coils=100, b_center=2.0 * TESLA, n_threads=0) -> "FieldMap":
        """SolenoidBField sampled on the grid points (library sampler)."""
        lib = capi.load_library()
        cfg = capi.Solenoid(radius=radius, length=length, n_coils=n_coils,
                            b_mag_center=b_center)
        values = np.zeros((n[0] * n[1] * n[2], 3), np.float64)
        nn = (C.c_int32 * 3)(*n)
        lo = (C.c_double * 3)(*point_min)
        hi = (C.c_double * 3)(*point_max)
        capi.check(lib.acts_kf_solenoid_map(C.byref(cfg), nn, lo, hi,
                                            values.ctypes.data_as(C.POINTER(C.c_double)),
                                            n_threads), "acts_kf_solenoid_map")
        fm = cls(n, point_min, point_max, values)
        fm.solenoid_config = dict(radius=radius, length=length, n_coils=n_coils,
                                  b_center=b_center)
        return fm


def curvilinear_rotation(direction) -> np.ndarray:
    """CurvilinearSurface::referenceFrame (CurvilinearSurface.cpp:25-39)."""
    t = np.asarray(direction, dtype=np.float64)
    t = t / np.linalg.norm(t)
    ref = np.array([0.0, 0.0, 1.0]) if abs(t[2]) < 0.999995 else np.array([1.0, 0.0, 0.0])
    u = np.cross(ref, t)
    u /= np.linalg.norm(u)
    v = np.cross(t, u)
    return np.stack([u, v, t], axis=1)
