"""Algorithmic floating-point work of the fit kernel.

Counts are the number of floating-point additions, subtractions,
multiplications, divisions and square roots written in the kernel source
(acts_amd/csrc/kf_device.hpp, kf_fit.hpp, kf_kernel.hpp) per code block;
transcendental calls (sin, cos, atan2, log) count as one operation each.
The per-track totals use the kernel's own per-track counters (accepted
steps S, attempted steps T, track states N, measurement states M) with

  FLOP_track = T*F_TRIAL + S*F_STEP + N*(F_SURF + F_MAT + F_NAV)
               + M*(F_UPD + F_UPDATE_STATE) + (N-1)*F_SMOOTH

Only the states up to the track tip (the last measurement) are counted,
which slightly under-counts tracks that keep propagating past their last
measurement.
"""
from __future__ import annotations

import numpy as np

# per attempted RK step (EigenStepper::step trial: 2 field points, k2..k4,
# error estimate)
F_TRIAL = 96
# per accepted RK step: k1, time, transport matrix D (dk/dT, dk/dL chains),
# 8x8 block update, parameter update and renormalisation, accuracy update,
# the post-step target intersection and both path-limit checks
F_STEP = 616 + 50 + 6
# per track state: covariance transport to the surface (reference frame,
# freeToPath, F2B (1+ds) Jt B2F chain, J C J^T), global->local, bound angles,
# jacobian re-initialisation
F_SURF = 1255 + 30
# per track state: next-target search (one intersection and one status update)
F_NAV = 100
# per state with material: path correction, slab scaling, Bethe + Highland +
# Landau sigma (float) and the double-precision state/covariance update
F_MAT = 200
# per measurement: 2D gain-matrix update with chi2
F_UPD = 267
# per measurement: EigenStepper::update (bound->free, jacToGlobal)
F_UPDATE_STATE = 100
# per smoothed state: G = Cf J^T Cp^-1 (with a 6x6 LU inverse), smoothed
# parameters and covariance
F_SMOOTH = 2276


def flops_per_track(n_steps, n_trials, n_states, n_meas) -> np.ndarray:
    s = np.asarray(n_steps, np.float64)
    t = np.asarray(n_trials, np.float64)
    n = np.asarray(n_states, np.float64)
    m = np.asarray(n_meas, np.float64)
    return (t * F_TRIAL + s * F_STEP + n * (F_SURF + F_MAT + F_NAV)
            + m * (F_UPD + F_UPDATE_STATE) + np.maximum(n - 1, 0) * F_SMOOTH)


def batch_flops(result) -> float:
    """Total algorithmic FLOP of one fit of the batch held in `result`."""
    return float(np.sum(flops_per_track(result.n_steps, result.n_step_trials,
                                        result.n_states, result.n_measurements)))


def forward_flops(result) -> float:
    """Algorithmic FLOP of the forward-filter kernel alone (no smoother)."""
    t = np.asarray(result.n_step_trials, np.float64)
    s = np.asarray(result.n_steps, np.float64)
    n = np.asarray(result.n_states, np.float64)
    m = np.asarray(result.n_measurements, np.float64)
    return float(np.sum(t * F_TRIAL + s * F_STEP + n * (F_SURF + F_MAT + F_NAV)
                        + m * (F_UPD + F_UPDATE_STATE)))
