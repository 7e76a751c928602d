"""Synthetic muon events (inputs of the batched fit) from the native generator.

Workloads follow the benchmark configurations:
  * ``barrel``  : 10 barrel layers, pT 1-10 GeV, |eta| < 1, Bz = 2 T
  * ``endcap``  : barrel + 2x7 discs, |eta| < 2.5 (6-17 measurements)
Start parameters sit on a curvilinear plane through the true vertex and are
the truth smeared by ten times the measurement scale:
sigma(loc0) = 0.15 mm, sigma(loc1) = 0.5 mm, sigma(phi) = sigma(theta) =
1 mrad, sigma(q/p)/|q/p| = 10 %, sigma(t) = 1 ns; the start covariance is
diag(sigma^2).
"""
from __future__ import annotations

import ctypes as C
import dataclasses
from dataclasses import dataclass

import numpy as np

from . import capi
from .detector import (ConstantField, Detector, FieldMap, MUON_MASS, NS, UM,
                       barrel_detector, barrel_endcap_detector, double_layer_detector, ragged_detector,
                       telescope_detector)


@dataclass
class Events:
    n_tracks: int
    start_params: np.ndarray   # [n,6]
    start_cov: np.ndarray      # [n,36]
    start_plane: np.ndarray    # [n,12]
    seq_offsets: np.ndarray    # [n+1] int64
    seq_surfaces: np.ndarray   # int32
    meas_offsets: np.ndarray   # [n+1] int64
    meas_surfaces: np.ndarray  # int32
    meas_values: np.ndarray    # [m,2]
    meas_cov: np.ndarray       # [m,4]
    truth_start: np.ndarray    # [n,6]
    truth_meas: np.ndarray     # [m,6]
    meas_is_outlier: np.ndarray  # [m] int32
    # measurement subspaces (None: every measurement 2D on loc0, loc1)
    meas_dims: np.ndarray | None = None      # [m] uint8
    meas_indices: np.ndarray | None = None   # [m,2] uint8

    @property
    def max_sequence_length(self) -> int:
        if self.n_tracks == 0:
            return 0
        return int(np.max(np.diff(self.seq_offsets)))

    def batch_in(self) -> capi.BatchIn:
        """acts_kf_batch_in over these (host) arrays; keep `self` alive."""
        b = capi.BatchIn()
        b.n_tracks = self.n_tracks
        for name in ("start_params", "start_cov", "start_plane", "seq_offsets",
                     "seq_surfaces", "meas_offsets", "meas_surfaces", "meas_values",
                     "meas_cov"):
            arr = getattr(self, name)
            assert arr.flags["C_CONTIGUOUS"]
            setattr(b, name, arr.ctypes.data)
        b.max_sequence_length = self.max_sequence_length
        for name in ("meas_dims", "meas_indices"):
            arr = getattr(self, name)
            if arr is not None:
                assert arr.flags["C_CONTIGUOUS"] and arr.dtype == np.uint8
                setattr(b, name, arr.ctypes.data)
        return b

    def subset(self, idx) -> "Events":
        """Events for the track indices `idx` (in that order)."""
        idx = np.asarray(idx, dtype=np.int64)
        so, mo = self.seq_offsets, self.meas_offsets
        seq = [self.seq_surfaces[so[i]:so[i + 1]] for i in idx]
        msl = [np.arange(mo[i], mo[i + 1]) for i in idx]
        mid = np.concatenate(msl) if len(msl) else np.zeros(0, np.int64)
        new_so = np.zeros(len(idx) + 1, np.int64)
        new_so[1:] = np.cumsum([len(s) for s in seq])
        new_mo = np.zeros(len(idx) + 1, np.int64)
        new_mo[1:] = np.cumsum([len(m) for m in msl])
        cat = lambda xs, dt: (np.ascontiguousarray(np.concatenate(xs)).astype(dt)
                              if len(xs) else np.zeros(0, dt))
        return Events(
            n_tracks=len(idx),
            start_params=np.ascontiguousarray(self.start_params[idx]),
            start_cov=np.ascontiguousarray(self.start_cov[idx]),
            start_plane=np.ascontiguousarray(self.start_plane[idx]),
            seq_offsets=new_so, seq_surfaces=cat(seq, np.int32),
            meas_offsets=new_mo,
            meas_surfaces=np.ascontiguousarray(self.meas_surfaces[mid]),
            meas_values=np.ascontiguousarray(self.meas_values[mid]),
            meas_cov=np.ascontiguousarray(self.meas_cov[mid]),
            truth_start=np.ascontiguousarray(self.truth_start[idx]),
            truth_meas=np.ascontiguousarray(self.truth_meas[mid]),
            meas_is_outlier=np.ascontiguousarray(self.meas_is_outlier[mid]),
            meas_dims=None if self.meas_dims is None
            else np.ascontiguousarray(self.meas_dims[mid]),
            meas_indices=None if self.meas_indices is None
            else np.ascontiguousarray(self.meas_indices[mid]))


def with_state_slots(events: Events, slots) -> Events:
    """The same tracks for the layer navigator: `slots` state slots per
    track (an int for every track, or one per track) instead of the surface
    sequence, whose entries the layer navigator does not read (-1 here)."""
    n = events.n_tracks
    lens = np.broadcast_to(np.asarray(slots, dtype=np.int64), (n,))
    if np.any(lens < 0):
        raise ValueError("negative slot count")
    so = np.zeros(n + 1, np.int64)
    so[1:] = np.cumsum(lens)
    return dataclasses.replace(events, seq_offsets=so,
                               seq_surfaces=np.full(int(so[-1]), -1, np.int32))


def particle_options(mass: float = MUON_MASS) -> capi.Options:
    o = capi.default_options()
    o.abs_pdg = 13
    o.mass = mass
    o.abs_charge = 1.0
    return o


def gen_config(n_tracks: int, seed: int, *, pt=(1.0, 10.0), pt_log=False,
               eta=(-1.0, 1.0), phi=(-np.pi, np.pi), sigma_loc=(15 * UM, 50 * UM),
               hole_probability=0.0, outlier_fraction=0.0,
               outlier_count=(1, 3), outlier_distance=20.0,
               min_measurements=1, n_threads=0) -> capi.GenConfig:
    c = capi.GenConfig()
    c.n_tracks = n_tracks
    c.seed = seed
    c.pt_min, c.pt_max = pt
    c.pt_log_uniform = 1 if pt_log else 0
    c.n_threads = n_threads
    c.eta_min, c.eta_max = eta
    c.phi_min, c.phi_max = phi
    c.vertex_sigma_z = 50.0
    c.sigma_loc0, c.sigma_loc1 = sigma_loc
    start = (10 * sigma_loc[0], 10 * sigma_loc[1], 1e-3, 1e-3, 0.1, 1.0 * NS)
    for i, v in enumerate(start):
        c.start_sigma[i] = v
    c.hole_probability = hole_probability
    c.outlier_fraction = outlier_fraction
    c.outlier_min, c.outlier_max = outlier_count
    c.outlier_distance = outlier_distance
    c.min_measurements = min_measurements
    c.max_attempts = 16
    return c


def generate(detector: Detector, field: ConstantField, cfg: capi.GenConfig,
             mass: float = MUON_MASS) -> Events:
    lib = capi.load_library()
    surf = detector.surface_table()
    mats = detector.material_table()
    fld = field.struct()
    opt = particle_options(mass)
    handle = C.c_void_p()
    capi.check(lib.acts_kf_gen_run(C.byref(cfg), surf, len(detector.surfaces), mats,
                                   len(detector.materials), C.byref(fld), C.byref(opt),
                                   C.byref(handle)), "acts_kf_gen_run")
    try:
        n, ns, nm = C.c_int64(), C.c_int64(), C.c_int64()
        lib.acts_kf_gen_sizes(handle, C.byref(n), C.byref(ns), C.byref(nm))
        n, ns, nm = n.value, ns.value, nm.value
        ev = Events(
            n_tracks=n,
            start_params=np.zeros((n, 6)), start_cov=np.zeros((n, 36)),
            start_plane=np.zeros((n, 12)),
            seq_offsets=np.zeros(n + 1, np.int64), seq_surfaces=np.zeros(ns, np.int32),
            meas_offsets=np.zeros(n + 1, np.int64), meas_surfaces=np.zeros(nm, np.int32),
            meas_values=np.zeros((nm, 2)), meas_cov=np.zeros((nm, 4)),
            truth_start=np.zeros((n, 6)), truth_meas=np.zeros((nm, 6)),
            meas_is_outlier=np.zeros(nm, np.int32))
        ptr = lambda a: a.ctypes.data
        lib.acts_kf_gen_fetch(handle, ptr(ev.start_params), ptr(ev.start_cov),
                              ptr(ev.start_plane), ptr(ev.seq_offsets), ptr(ev.seq_surfaces),
                              ptr(ev.meas_offsets), ptr(ev.meas_surfaces),
                              ptr(ev.meas_values), ptr(ev.meas_cov), ptr(ev.truth_start),
                              ptr(ev.truth_meas), ptr(ev.meas_is_outlier))
    finally:
        lib.acts_kf_gen_free(handle)
    return ev


def apply_measurement_kinds(events: Events, detector: Detector, seed: int) -> Events:
    """Measurements of the kinds the detector's surfaces give
    (Detector.kinds: dimension, bound indices, sigmas per surface), smeared
    from the generator's truth with those sigmas, as the reference's
    MeasurementsCreator does for a resolution map (Tests/CommonHelpers/
    .../MeasurementsCreator.hpp:51-174, FitterTestsCommon.hpp:145-155).  A
    1D measurement keeps its value and variance in the first entries of
    meas_values and meas_cov."""
    rng = np.random.default_rng(seed)
    nm = len(events.meas_surfaces)
    dims = np.zeros(nm, np.uint8)
    idx = np.zeros((nm, 2), np.uint8)
    vals = np.zeros((nm, 2))
    cov = np.zeros((nm, 4))
    noise = rng.normal(size=(nm, 2))
    for m in range(nm):
        d, ix, sig = detector.kinds[events.meas_surfaces[m]]
        dims[m] = d
        idx[m] = ix
        for i in range(d):
            vals[m, i] = events.truth_meas[m, ix[i]] + sig[i] * noise[m, i]
            cov[m, 3 * i] = sig[i] ** 2
    import dataclasses
    return dataclasses.replace(events, meas_values=vals, meas_cov=cov, meas_dims=dims,
                               meas_indices=idx)


# ---- benchmark workloads ---------------------------------------------------
WORKLOADS = {
    # 1,000 muons, barrel only, |eta| < 1, uniform Bz = 2 T
    "barrel": dict(detector=barrel_detector, n_tracks=1000, eta=(-1.0, 1.0),
                   pt=(1.0, 10.0)),
    # 100,000 muons, barrel + 2x7 endcap discs, |eta| < 2.5
    "endcap": dict(detector=barrel_endcap_detector, n_tracks=100000, eta=(-2.5, 2.5),
                   pt=(1.0, 10.0)),
    # 500,000 muons, pT log-uniform 0.5-100 GeV, in a 201x201x401 field map
    # (x, y in +-2 m, z in +-4 m) sampled from a 2 T solenoid of radius 1.2 m
    # and length 6 m (100 coils)
    "solenoid": dict(detector=barrel_endcap_detector, n_tracks=500000, eta=(-2.5, 2.5),
                     pt=(0.5, 100.0), pt_log=True, field="solenoid"),
    # 2,000,000 muons with 6-30 measurements (ragged_detector: 15 barrel
    # double layers of half-length 1 m and 2x3 discs, |eta| < 3.5, so the
    # number of crossed layers depends on eta), 5 % hole probability per
    # crossed layer, at least 6 measurements, 10 % of the tracks with 1-3
    # outliers at 20 sigma (rejected by a chi2 cut of 25), pT log-uniform
    # 0.5-100 GeV, uniform Bz = 2 T.  A track with a rejected outlier never
    # counts all its measurements, so with the direct navigator it only ends
    # at the loop-protection path limit (half a helix turn); in a finite field
    # map it would end with FIELD_OUT_OF_BOUNDS instead, as in the reference.
    # the reference's own fitter-test geometry: 10,000 muons along x through
    # six beryllium planes (two 2D pixel planes, four stereo 1D strip
    # planes), pT 0.5-10 GeV, |eta| < 0.1, |phi| < 0.1, no field
    "telescope": dict(detector=telescope_detector, n_tracks=10000, eta=(-0.1, 0.1),
                      phi=(-0.1, 0.1), pt=(0.5, 10.0), field="zero"),
    "ragged": dict(detector=ragged_detector, n_tracks=2_000_000, eta=(-3.5, 3.5),
                   pt=(0.5, 100.0), pt_log=True, hole_probability=0.05,
                   outlier_fraction=0.1, min_measurements=6,
                   options=dict(outlier_chi2_cut=25.0)),
}

FIELD_MAP_POINTS = (201, 201, 401)


def workload(name: str, n_tracks: int | None = None, seed: int = 42,
             n_threads: int = 0, field_map_points=None):
    """(detector, field, events) for a named benchmark workload."""
    w = WORKLOADS[name]
    det = w["detector"]()
    if w.get("field") == "solenoid":
        field = FieldMap.solenoid(n=field_map_points or FIELD_MAP_POINTS, n_threads=n_threads)
    elif w.get("field") == "zero":
        field = ConstantField(0.0, 0.0, 0.0)
    else:
        field = ConstantField()
    n = w["n_tracks"] if n_tracks is None else n_tracks
    cfg = gen_config(n, seed, pt=w["pt"], pt_log=w.get("pt_log", False), eta=w["eta"],
                     phi=w.get("phi", (-np.pi, np.pi)),
                     hole_probability=w.get("hole_probability", 0.0),
                     outlier_fraction=w.get("outlier_fraction", 0.0),
                     min_measurements=w.get("min_measurements", 1), n_threads=n_threads)
    ev = generate(det, field, cfg)
    if det.kinds is not None:
        ev = apply_measurement_kinds(ev, det, seed)
    return det, field, ev


def workload_options(name: str, **overrides):
    """KalmanFitterOptions of a named workload (e.g. its outlier chi2 cut)."""
    from .fitter import KalmanFitterOptions
    kw = dict(WORKLOADS[name].get("options", {}))
    kw.update(overrides)
    return KalmanFitterOptions(**kw)
