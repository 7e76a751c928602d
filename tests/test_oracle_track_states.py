"""Pins of the oracle's track-state table (RootTrackStatesWriter's columns,
RootTrackStatesWriter.cpp:364-735; unbiased parameters,
TrackHelpersUnbiased.cpp:21-54) and of the host-side tree builder.

The checks use what the mathematics fixes rather than the writer's
formulas: the unbiased state at a track's last measurement is the
prediction (the smoothed state there is the filtered one, and removing the
measurement undoes the update); in the measurement plane the unbiased
covariance and estimate are the information-form difference of the
smoothed state and the measurement; the columns of one parameter set obey
the kinematic identities (p = |px,py,pz| = 1/|q/p|, pT = |px,py|,
eta = asinh(pz/pT), phi and theta of the momentum); hit positions lie on
their surfaces; the state types count the fit's measurements, outliers and
holes; pulls against the truth have zero mean and unit width where the fit
is consistent.
"""
import math

import numpy as np
import pytest

import acts_amd as A
from acts_amd import capi
from acts_amd import track_states as TS
from acts_amd.fitter import ALL_STATE_OUTPUTS

FLOAT_REL = 4e-7  # a few float ulps


def _fit(oracle_mod, workload, n, seed):
    det, field, ev = A.workload(workload, n_tracks=n, seed=seed)
    opts = A.workload_options(workload)
    r, _ = oracle_mod.fit(det, field, opts, ev, n_threads=0, outputs=ALL_STATE_OUTPUTS)
    tab = oracle_mod.track_states(det, field, opts, ev, r, ev.truth_meas)
    return det, ev, r, tab


@pytest.fixture(scope="module")
def barrel(oracle_mod):
    return _fit(oracle_mod, "barrel", 1500, 71)


@pytest.fixture(scope="module")
def ragged(oracle_mod):
    return _fit(oracle_mod, "ragged", 600, 72)


@pytest.fixture(scope="module")
def telescope(oracle_mod):
    return _fit(oracle_mod, "telescope", 800, 73)


def _close(a, b, rel=FLOAT_REL, abs_=0.0):
    return np.abs(a - b) <= rel * np.maximum(np.abs(a), np.abs(b)) + abs_


def test_state_types_count_the_fit(ragged):
    _, ev, r, tab = ragged
    st, trk = tab.column("stateType"), tab.column("track")
    for t in np.nonzero(r.status == 0)[0][:300]:
        rows = np.arange(ev.seq_offsets[t], ev.seq_offsets[t] + r.n_states[t])
        assert np.all(trk[rows] == t)
        assert np.count_nonzero(st[rows] == TS.MEASUREMENT) == r.n_measurements[t]
        assert np.count_nonzero(st[rows] == TS.OUTLIER) == r.n_outliers[t]
        assert np.count_nonzero(st[rows] == TS.HOLE) == r.n_holes[t]
        # slots past the tip hold no state
        tail = np.arange(ev.seq_offsets[t] + r.n_states[t], ev.seq_offsets[t + 1])
        assert np.all(st[tail] == -1)
    for t in np.nonzero(r.status != 0)[0]:
        assert np.all(st[ev.seq_offsets[t]:ev.seq_offsets[t + 1]] == -1)
    # unbiased parameters and hit columns exist exactly for states with a
    # measurement (measurements and outliers)
    valid = st >= 0
    link = (st == TS.MEASUREMENT) | (st == TS.OUTLIER)
    has = tab.column("has_params")
    assert np.all(has[valid & link] == 0xF) and np.all(has[valid & ~link] == 0x7)
    assert np.all(np.isfinite(tab.column("pull_x_hit")[link]))
    assert np.all(np.isnan(tab.column("pull_x_hit")[valid & ~link]))
    assert np.all(np.isnan(tab.column("eLOC0_ubs")[valid & ~link]))


@pytest.mark.parametrize("which", ["barrel", "telescope"])
def test_unbiased_at_the_tip_is_the_prediction(request, which):
    """On the telescope the tip is a 1D strip measurement of loc1."""
    _, ev, r, tab = request.getfixturevalue(which)
    ok = np.nonzero(r.status == 0)[0]
    tip = ev.seq_offsets[ok] + r.n_states[ok] - 1
    assert np.all(tab.column("stateType")[tip] == TS.MEASUREMENT)
    for q in TS.PARAMS:
        a, b = tab.column(f"{q}_ubs")[tip], tab.column(f"{q}_prt")[tip]
        # the update and its removal cancel to ~1e-12 relative; scale by the
        # prediction's error for parameters near zero
        err = tab.column(f"err_{q}_prt")[tip]
        assert np.all(np.abs(a - b) <= 1e-6 * np.maximum(np.abs(b), err)), q
        assert np.all(_close(tab.column(f"err_{q}_ubs")[tip], err, 1e-5)), q


def test_unbiased_in_the_measurement_plane_is_the_information_difference(barrel):
    _, ev, r, tab = barrel
    rows = np.nonzero(tab.column("dim_hit") == 2)[0]
    mi = r.state_measurement[rows]
    Cs = r.smoothed_cov[rows].reshape(-1, 6, 6)[:, :2, :2]
    xs = r.smoothed[rows][:, :2]
    V = ev.meas_cov[mi].reshape(-1, 2, 2)
    m = ev.meas_values[mi]
    # (Cs^-1 - V^-1)^-1 and Cu (Cs^-1 xs - V^-1 m)
    Ci, Vi = np.linalg.inv(Cs), np.linalg.inv(V)
    Cu = np.linalg.inv(Ci - Vi)
    xu = np.einsum("nij,nj->ni", Cu, np.einsum("nij,nj->ni", Ci, xs) - np.einsum("nij,nj->ni", Vi, m))
    for i, q in enumerate(("eLOC0", "eLOC1")):
        got_err = tab.column(f"err_{q}_ubs")[rows].astype(np.float64)
        want_err = np.sqrt(Cu[:, i, i])
        assert np.all(np.abs(got_err - want_err) <= 1e-5 * want_err), q
        got = tab.column(f"{q}_ubs")[rows].astype(np.float64)
        assert np.all(np.abs(got - xu[:, i]) <= 1e-5 * want_err + 1e-6 * np.abs(xu[:, i])), q


@pytest.mark.parametrize("s", TS.PARAM_SETS)
def test_kinematic_identities(ragged, s):
    _, _, _, tab = ragged
    rows = np.nonzero(np.isfinite(tab.column(f"eLOC0_{s}")))[0]
    f = {q: tab.column(f"{q}_{s}")[rows].astype(np.float64)
         for q in ("px", "py", "pz", "pT", "eta", "ePHI", "eTHETA", "eQOP")}
    p = np.sqrt(f["px"] ** 2 + f["py"] ** 2 + f["pz"] ** 2)
    assert np.all(_close(p, 1 / np.abs(f["eQOP"]), 2e-6))
    assert np.all(_close(f["pT"], np.hypot(f["px"], f["py"]), 2e-6))
    assert np.all(_close(f["eta"], np.arcsinh(f["pz"] / f["pT"]), 2e-6, 1e-6))
    dphi = np.remainder(np.arctan2(f["py"], f["px"]) - f["ePHI"] + np.pi, 2 * np.pi) - np.pi
    assert np.all(np.abs(dphi) < 1e-5)
    assert np.all(np.abs(np.arctan2(f["pT"], f["pz"]) - f["eTHETA"]) < 1e-5)


def test_positions_lie_on_their_surfaces(ragged, oracle_mod):
    import oracle
    det, ev, r, tab = ragged
    rows = np.nonzero(tab.column("dim_hit") == 2)[0][:400]
    surf = r.state_surface[rows]
    table = det.surface_table()
    for sets in (("g_x_hit", "g_y_hit", "g_z_hit", "l_x_hit", "l_y_hit"),
                 ("g_x_smt", "g_y_smt", "g_z_smt", "eLOC0_smt", "eLOC1_smt"),
                 ("t_x", "t_y", "t_z", "t_eLOC0", "t_eLOC1")):
        g = np.stack([tab.column(c)[rows] for c in sets[:3]], 1).astype(np.float64)
        loc = np.stack([tab.column(c)[rows] for c in sets[3:]], 1).astype(np.float64)
        for k in range(len(rows)):
            st = table[surf[k]]
            rc, back = oracle.surface_local(st, False, g[k])
            assert rc == 0, sets
            # float rounding of the global position (~6e-8 of its norm)
            # moves the local position by as much, in mm or, for the disc's
            # angle, in mm over the radius
            tol = 4e-7 * np.linalg.norm(g[k]) + 4e-7 * np.abs(loc[k])
            if st.type == capi.SURFACE_DISC:
                tol[1] = 4e-7 * np.linalg.norm(g[k]) / abs(loc[k][0]) + 4e-7 * abs(loc[k][1])
            assert np.all(np.abs(back[:2] - loc[k]) <= tol), (sets, k, back, loc[k])
    t_r = tab.column("t_r")[rows].astype(np.float64)
    assert np.all(_close(t_r, np.hypot(tab.column("t_x")[rows], tab.column("t_y")[rows]), 2e-6))
    tdir = np.stack([tab.column(c)[rows] for c in ("t_dx", "t_dy", "t_dz")], 1)
    assert np.allclose(np.linalg.norm(tdir, axis=1), 1, atol=1e-6)


def test_pulls_against_truth(barrel):
    _, _, _, tab = barrel
    link = tab.column("dim_hit") == 2
    # positions: unit width; angles and q/p: narrower, because the truth has
    # no random scattering and the covariance has (the fit's pulls at the
    # first state measure 0.56 / 0.47 / 0.33 for phi / theta / q/p)
    for name, lo in (("pull_eLOC0_smt", 0.7), ("pull_eLOC1_smt", 0.7), ("pull_eLOC0_ubs", 0.7),
                     ("pull_eLOC1_ubs", 0.7), ("pull_x_hit", 0.7), ("pull_y_hit", 0.7),
                     ("pull_ePHI_smt", 0.2), ("pull_eTHETA_smt", 0.2), ("pull_eQOP_smt", 0.2)):
        v = tab.column(name)[link]
        v = v[np.abs(v) < 10]
        assert abs(v.mean()) < 0.05, (name, v.mean())
        assert lo < v.std() < 1.15, (name, v.std())
    # residuals are differences against the truth (phi wrapped), so a pull
    # times its error is the residual
    for q in TS.PARAMS[:5]:
        res = tab.column(f"res_{q}_smt")[link].astype(np.float64)
        pe = (tab.column(f"pull_{q}_smt")[link].astype(np.float64)
              * tab.column(f"err_{q}_smt")[link])
        assert np.all(np.abs(res - pe) <= 1e-5 * np.abs(res) + 1e-12), q
        if q != "ePHI":
            d = tab.column(f"{q}_smt")[link].astype(np.float64) - tab.column(f"t_{q}")[link]
            assert np.all(np.abs(res - d) <= 2e-6 * np.maximum(np.abs(tab.column(f"{q}_smt")[link]), 1e-3)), q


def test_table_without_truth(oracle_mod):
    det, field, ev = A.workload("barrel", n_tracks=50, seed=73)
    opts = A.KalmanFitterOptions()
    r, _ = oracle_mod.fit(det, field, opts, ev, n_threads=0, outputs=ALL_STATE_OUTPUTS)
    tab = oracle_mod.track_states(det, field, opts, ev, r, None)
    with_truth = oracle_mod.track_states(det, field, opts, ev, r, ev.truth_meas)
    for name in ("t_x", "t_eQOP", "res_eLOC0_smt", "pull_eQOP_ubs"):
        assert np.all(np.isnan(tab.column(name)))
    for name in ("eLOC0_smt", "err_eQOP_ubs", "pull_x_hit", "g_z_hit", "pT_prt"):
        assert np.array_equal(tab.column(name), with_truth.column(name), equal_nan=True)


def test_tree_push_rules():
    """Two tracks on a hand-made table: the vectors come tip first; a hole
    pushes no residuals of present parameter sets and no hit residual, but
    NaN residuals of the absent unbiased set; a failed track is skipped."""
    tab = TS.TrackStatesTable.empty(6)
    tab.floats[:] = np.arange(6, dtype=np.float32)[None, :] + 100 * np.arange(capi.TS_NFLOAT)[:, None]
    st = tab.column("stateType")
    # track 0: slots 0-2 (measurement, hole, measurement); track 1 (failed):
    # slot 3; track 2: slot 4 (outlier) and an unused slot 5
    st[:] = [TS.MEASUREMENT, TS.HOLE, TS.MEASUREMENT, -1, TS.OUTLIER, -1]
    tab.column("dim_hit")[:] = [2, 0, 2, 0, 2, 0]
    tab.column("has_params")[:] = [0xF, 0x7, 0xF, 0, 0xF, 0]
    for c in ("res_eLOC0_ubs", "eLOC0_ubs"):
        tab.column(c)[1] = np.nan
    seq = np.array([0, 3, 4, 6])
    n_states = np.array([3, 1, 1])
    status = np.array([0, 3, 0])
    tr = TS.tree(tab, n_states, status, np.array([2, 1, 0]), np.array([7, 8, 9, 10, 11, 12]), seq)
    assert list(tr["track_nr"]) == [0, 2] and list(tr["nStates"]) == [3, 1]
    v, o = tr["stateType"]
    assert list(o) == [0, 3, 4] and list(v) == [TS.MEASUREMENT, TS.HOLE, TS.MEASUREMENT, TS.OUTLIER]
    v, o = tr["module_id"]
    assert list(v) == [9, 8, 7, 11]
    v, o = tr["eLOC0_prt"]
    assert list(o) == [0, 3, 4]
    v, o = tr["res_eLOC0_prt"]  # the hole (row 1) pushes nothing
    assert list(o) == [0, 2, 3] and list(v) == list(tab.column("res_eLOC0_prt")[[2, 0, 4]])
    v, o = tr["res_x_hit"]
    assert list(o) == [0, 2, 3]
    v, o = tr["res_eLOC0_ubs"]  # absent set: NaN pushed for the hole
    assert list(o) == [0, 3, 4] and np.isnan(v[1])
    assert list(tr["nUnbiased"]) == [2, 1] and list(tr["nPredicted"]) == [3, 1]
    v, o = tr["unbiased"]
    assert list(v) == [True, False, True, True]


def test_parquet_round_trip(barrel, tmp_path):
    pq = pytest.importorskip("pyarrow.parquet")
    _, ev, r, tab = barrel
    tr = TS.tree(tab, r.n_states, r.status, r.n_measurements, r.state_surface, ev.seq_offsets)
    path = str(tmp_path / "trackstates.parquet")
    TS.write_parquet(path, tr)
    t = pq.read_table(path)
    assert t.num_rows == np.count_nonzero(r.status == 0)
    for name in ("eLOC0_smt", "pull_x_hit", "stateType", "nMeasurements", "t_eQOP"):
        assert name in t.column_names
    first = t.column("eLOC0_smt")[0].as_py()
    assert len(first) == r.n_states[np.nonzero(r.status == 0)[0][0]]
    assert math.isclose(first[0], float(tr["eLOC0_smt"][0][0]))


def test_1d_strips_in_the_table(telescope):
    """Strip rows of the telescope: the hit dimension is 1; the hit's local
    position is the measured component with 0 for the other; the x hit
    columns describe the measured component (the writer's res[0]), the y
    columns are NaN; the unbiased variance and estimate of the measured
    parameter are the 1D information difference of the smoothed state and
    the measurement."""
    _, ev, r, tab = telescope
    rows = np.nonzero(tab.column("dim_hit") == 1)[0]
    assert len(rows) == 4 * int(np.count_nonzero(r.status == 0))
    mi = r.state_measurement[rows]
    assert np.all(ev.meas_dims[mi] == 1)
    a = ev.meas_indices[mi, 0].astype(int)
    m = ev.meas_values[mi, 0]
    V = ev.meas_cov[mi, 0]
    lx, ly = tab.column("l_x_hit")[rows], tab.column("l_y_hit")[rows]
    assert np.all(np.where(a == 0, lx, ly) == m.astype(np.float32))
    assert np.all(np.where(a == 0, ly, lx) == 0)
    for c in ("res_y_hit", "err_y_hit", "pull_y_hit"):
        assert np.all(np.isnan(tab.column(c)[rows])), c
    xp = r.predicted[rows, a]
    assert np.all(_close(tab.column("res_x_hit")[rows], m - xp, 1e-6, 1e-9))
    assert np.all(_close(tab.column("err_x_hit")[rows], np.sqrt(V), 1e-6))
    Cs = r.smoothed_cov[rows].reshape(-1, 6, 6)[np.arange(len(rows)), a, a]
    xs = r.smoothed[rows, a]
    Cu = 1.0 / (1.0 / Cs - 1.0 / V)
    xu = Cu * (xs / Cs - m / V)
    q = np.array(["eLOC0", "eLOC1"])[a]
    for i in range(2):
        sel = a == i
        name = ("eLOC0", "eLOC1")[i]
        got_err = tab.column(f"err_{name}_ubs")[rows[sel]].astype(np.float64)
        assert np.all(np.abs(got_err - np.sqrt(Cu[sel])) <= 1e-5 * np.sqrt(Cu[sel])), name
        got = tab.column(f"{name}_ubs")[rows[sel]].astype(np.float64)
        assert np.all(np.abs(got - xu[sel]) <= 1e-5 * np.sqrt(Cu[sel]) + 1e-6 * np.abs(xu[sel]))
    assert len(q) == len(rows)
