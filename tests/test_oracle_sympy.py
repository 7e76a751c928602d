"""Pins of the oracle's SympyStepper (SympyStepper.cpp:141-305; the vacuum
kernel generate_sympy_stepper.py:157-444 writes).

The ATLAS-form step and the EigenStepper's RKN4 step are the same Runge-
Kutta method written differently, with the same step-size control: on the
same inputs they take the same steps, so positions, directions and times
agree to rounding.  The SympyStepper carries the bound-to-free jacobian
through tangents of its stages instead of an 8x8 transport matrix; the
transported jacobian is pinned against central finite differences and
against the EigenStepper's (a dropped field term or a wrong tangent in one
column fails both), and whole fits with either stepper agree on every
output.  The analytic helix pins the trajectory at the integration-test
tolerances.
"""
import math

import numpy as np
import pytest

import acts_amd as A
from acts_amd import capi
from acts_amd.detector import curvilinear_rotation
from acts_amd.fitter import ALL_STATE_OUTPUTS

from parity import cov_rel_err, param_rel_err
from test_oracle_propagation import (BZ, EPS_DIR, EPS_POS, helix, no_material_options,
                                     start_at_origin, surface)


def _prop(stepper, plane, params, target, cov=None):
    import oracle
    o = no_material_options()
    o.stepper = stepper
    return oracle.propagate(A.ConstantField(bz=BZ), o, plane, params, cov, target, 1.0)


CASES = [(0.3, 1.2, 1.0, 1), (2.9, 0.8, 3.0, -1), (-1.7, 2.0, 0.7, 1), (3.1, 1.5708, 10.0, -1)]


def _target(kind, d0, theta):
    if kind == "cylinder":
        return surface(capi.SURFACE_CYLINDER, bounds=(120.0, 1e5))
    if kind == "disc":
        return surface(capi.SURFACE_DISC, translation=(0, 0, math.copysign(500.0, math.cos(theta))),
                       bounds=(0.0, 1e5))
    return surface(capi.SURFACE_PLANE, rotation=curvilinear_rotation(d0), translation=300.0 * d0)


@pytest.mark.parametrize("kind", ["cylinder", "disc", "plane"])
@pytest.mark.parametrize("phi,theta,p,q", CASES)
def test_sympy_equals_eigen_propagation(oracle_mod, kind, phi, theta, p, q):
    plane0, d0 = start_at_origin(phi, theta)
    start = np.array([0.0, 0.0, phi, theta, q / p, 0.0])
    tgt = _target(kind, d0, theta)
    re_ = _prop(capi.STEPPER_EIGEN, plane0, start, tgt, cov=np.eye(6))
    rs = _prop(capi.STEPPER_SYMPY, plane0, start, tgt, cov=np.eye(6))
    assert re_[0] == rs[0]  # same outcome (a disc parallel to the track is unreachable)
    if re_[0]:
        return
    # parameters, path length and the transported jacobian
    assert np.all(param_rel_err(rs[1], re_[1]) < 1e-10), (rs[1], re_[1])
    assert rs[4] == pytest.approx(re_[4], rel=1e-12)
    assert np.allclose(rs[3], re_[3], rtol=1e-9, atol=1e-11), (rs[3], re_[3])


@pytest.mark.parametrize("kind", ["cylinder", "disc", "plane"])
@pytest.mark.parametrize("phi,theta,p,q", CASES[:3])
def test_sympy_jacobian_against_finite_differences(oracle_mod, kind, phi, theta, p, q):
    plane0, d0 = start_at_origin(phi, theta)
    start = np.array([0.0, 0.0, phi, theta, q / p, 0.0])
    tgt = _target(kind, d0, theta)
    rc, ep, _, jac, _, _ = _prop(capi.STEPPER_SYMPY, plane0, start, tgt, cov=np.eye(6))
    assert rc == 0
    steps = np.array([1e-4, 1e-4, 1e-6, 1e-6, 1e-6 * abs(q / p), 1e-4])
    fd = np.zeros((6, 6))
    for j in range(6):
        cols = []
        for sgn in (1, -1):
            x = start.copy()
            x[j] += sgn * steps[j]
            r, e, *_ = _prop(capi.STEPPER_SYMPY, plane0, x, tgt)
            assert r == 0
            cols.append(e)
        diff = cols[0] - cols[1]
        diff[2] = math.remainder(diff[2], 2 * math.pi)
        fd[:, j] = diff / (2 * steps[j])
    scale = np.maximum(np.abs(fd), np.abs(jac)).max(axis=1, keepdims=True) + 1e-12
    assert (np.abs(jac - fd) / scale).max() < 0.025


def test_sympy_helix(oracle_mod):
    phi, theta, p, q = 0.3, 1.2, 1.0, 1
    plane0, _ = start_at_origin(phi, theta)
    start = np.array([0.0, 0.0, phi, theta, q / p, 0.0])
    cyl = surface(capi.SURFACE_CYLINDER, bounds=(300.0, 1e5))
    rc, ep, _, _, path, _ = _prop(capi.STEPPER_SYMPY, plane0, start, cyl)
    assert rc == 0
    pos, d = helix(phi, theta, p, q, path)
    assert abs(math.hypot(pos[0], pos[1]) - 300.0) < EPS_POS
    assert abs(ep[1] - pos[2]) < EPS_POS
    assert abs(math.remainder(ep[2] - math.atan2(d[1], d[0]), 2 * math.pi)) < EPS_DIR


@pytest.mark.parametrize("workload,n,seed", [("barrel", 1000, 101), ("endcap", 3000, 102),
                                             ("solenoid", 1500, 103)])
def test_sympy_fit_equals_eigen_fit(oracle_mod, workload, n, seed):
    det, field, ev = A.workload(workload, n_tracks=n, seed=seed)
    out = {}
    for stepper in (capi.STEPPER_EIGEN, capi.STEPPER_SYMPY):
        o = A.workload_options(workload, stepper=stepper)
        out[stepper], _ = oracle_mod.fit(det, field, o, ev, n_threads=0, outputs=ALL_STATE_OUTPUTS)
    e, s = out[capi.STEPPER_EIGEN], out[capi.STEPPER_SYMPY]
    same = (e.status == s.status) & (e.n_states == s.n_states)
    print(workload, "differing tracks", int(np.count_nonzero(~same)))
    assert np.count_nonzero(~same) <= 0.002 * n
    ok = same & (e.status == 0) & (e.chi2 <= 100 * np.maximum(e.ndf, 1))
    so = ev.seq_offsets
    idx = np.concatenate([np.arange(so[t], so[t] + e.n_states[t]) for t in np.nonzero(ok)[0]])
    assert np.array_equal(e.n_steps[ok], s.n_steps[ok])
    pe = param_rel_err(s.smoothed[idx], e.smoothed[idx]).max()
    pj = np.abs(s.jacobian[idx] - e.jacobian[idx]).max()
    pc = cov_rel_err(s.predicted_cov[idx], e.predicted_cov[idx]).max()
    print(workload, "smoothed", pe, "jacobian", pj, "predicted cov", pc)
    assert pe < 1e-9
    assert pc < 1e-9
    assert pj < 1e-9 * max(1.0, np.abs(e.jacobian[idx]).max())
