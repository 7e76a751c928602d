"""Track-state table: the output of RootTrackStatesWriter for a fitted batch.

The reference writes one ROOT tree entry per fitted track whose vector
branches hold one value per track state, in trackStatesReversed() order
(the tip first) (Examples/Io/Root/src/RootTrackStatesWriter.cpp:310-740).
Here the per-state values are computed on the GPU by
acts_kf_track_states_device into a dense column-major table (one row per
state slot of the batch's sequence layout, float columns as the writer's
clampValue<float> stores them); `tree()` turns that table into the writer's
per-track vectors, applying its push rules, and `write_parquet()` stores
them under the writer's branch names (ROOT itself is not a dependency).

Branches the table cannot provide: particle_ids_* (the generator has no
particle barcodes) and the geometry identifiers, which `tree()` takes from
an optional per-surface (volume, layer, sensitive) table (default: volume
0, layer 0, sensitive = the surface index).
"""
from __future__ import annotations

import numpy as np

from . import capi

PARAM_SETS = ("prt", "flt", "smt", "ubs")  # predicted, filtered, smoothed, unbiased
PARAM_SET_FLAGS = ("predicted", "filtered", "smoothed", "unbiased")
PARAM_COUNTS = ("nPredicted", "nFiltered", "nSmoothed", "nUnbiased")
PARAMS = ("eLOC0", "eLOC1", "ePHI", "eTHETA", "eQOP", "eT")


def _float_columns() -> list[str]:
    names = ["chi2", "pathLength", "t_x", "t_y", "t_z", "t_r", "t_dx", "t_dy", "t_dz"]
    names += [f"t_{p}" for p in PARAMS]
    names += ["l_x_hit", "l_y_hit", "g_x_hit", "g_y_hit", "g_z_hit",
              "res_x_hit", "res_y_hit", "err_x_hit", "err_y_hit", "pull_x_hit", "pull_y_hit"]
    for s in PARAM_SETS:
        names += [f"{p}_{s}" for p in PARAMS]
        names += [f"res_{p}_{s}" for p in PARAMS]
        names += [f"err_{p}_{s}" for p in PARAMS]
        names += [f"pull_{p}_{s}" for p in PARAMS]
        names += [f"g_x_{s}", f"g_y_{s}", f"g_z_{s}", f"px_{s}", f"py_{s}", f"pz_{s}",
                  f"eta_{s}", f"pT_{s}"]
    return names


FLOAT_COLUMNS = tuple(_float_columns())
INT_COLUMNS = ("stateType", "dim_hit", "has_params", "track")
assert len(FLOAT_COLUMNS) == capi.TS_NFLOAT and len(INT_COLUMNS) == capi.TS_NINT
_FIDX = {n: i for i, n in enumerate(FLOAT_COLUMNS)}
_IIDX = {n: i for i, n in enumerate(INT_COLUMNS)}

# state types (RootTrackStatesWriter.hpp:89-95)
MEASUREMENT, OUTLIER, HOLE, MATERIAL, UNKNOWN = 0, 1, 2, 3, 4


class TrackStatesTable:
    """The dense table on the host: floats [TS_NFLOAT][n_slots], ints
    [TS_NINT][n_slots]."""

    def __init__(self, floats: np.ndarray, ints: np.ndarray):
        assert floats.shape[0] == capi.TS_NFLOAT and ints.shape[0] == capi.TS_NINT
        self.floats = floats
        self.ints = ints
        self.n_slots = floats.shape[1]

    @classmethod
    def empty(cls, n_slots: int) -> "TrackStatesTable":
        return cls(np.zeros((capi.TS_NFLOAT, n_slots), np.float32),
                   np.zeros((capi.TS_NINT, n_slots), np.int32))

    def column(self, name: str) -> np.ndarray:
        if name in _FIDX:
            return self.floats[_FIDX[name]]
        return self.ints[_IIDX[name]]

    def out_struct(self) -> capi.TrackStatesOut:
        o = capi.TrackStatesOut()
        o.n_slots = self.n_slots
        o.columns = self.floats.ctypes.data
        o.ints = self.ints.ctypes.data
        return o


def device_table(fitter, dev_events, dev_result, truth=None, stream=None):
    """Run the table kernel on a batch fitted on the device.

    `dev_events` is a DeviceEvents, `dev_result` a DeviceResult of the fit
    holding the per-state outputs the kernel reads (see
    acts_kf_track_states_device), `truth` an optional device tensor [m][6].
    Returns (floats, ints) device tensors; `.cpu()` them into a
    TrackStatesTable."""
    import torch
    n_slots = dev_events.n_slots
    dev = dev_result.t["status"].device
    floats = torch.empty((capi.TS_NFLOAT, n_slots), dtype=torch.float32, device=dev)
    ints = torch.empty((capi.TS_NINT, n_slots), dtype=torch.int32, device=dev)
    o = capi.TrackStatesOut()
    o.n_slots = n_slots
    o.columns = floats.data_ptr()
    o.ints = ints.data_ptr()
    fitter.track_states_device(dev_events.batch_in(), dev_result.batch_out(),
                               truth.data_ptr() if truth is not None else None, o, stream)
    return floats, ints


def host_table(floats, ints) -> TrackStatesTable:
    return TrackStatesTable(floats.cpu().numpy(), ints.cpu().numpy())


def _ragged(values: np.ndarray, keep: np.ndarray, track_of_row: np.ndarray, n_out: int):
    """Values of the kept rows and per-track offsets."""
    counts = np.bincount(track_of_row[keep], minlength=n_out)
    offsets = np.zeros(n_out + 1, np.int64)
    np.cumsum(counts, out=offsets[1:])
    return values[keep], offsets


def tree(table: TrackStatesTable, n_states: np.ndarray, status: np.ndarray,
         n_measurements: np.ndarray, state_surface: np.ndarray, seq_offsets: np.ndarray,
         geometry_ids: np.ndarray | None = None, event_nr: int = 0) -> dict:
    """The writer's tree for one event: {branch: array} for per-track
    branches and {branch: (values, offsets)} for vector branches, one entry
    per fitted track (status 0) in batch order.

    Push rules of the writer (RootTrackStatesWriter.cpp:364-735): every
    state pushes the state type, chi2, path length, geometry ids, truth and
    hit position (NaN without a measurement), the presence flag and values,
    errors, global position and momentum of each parameter set; residuals
    and pulls of a present parameter set only for states with a
    measurement, NaN for an absent set; the hit residual, error, pull and
    dim_hit only with the predicted set: for states with a measurement, or
    NaN and 0 when the state has no prediction."""
    fitted = np.nonzero(status == 0)[0]
    n_out = len(fitted)
    ns = n_states[fitted].astype(np.int64)
    # rows in trackStatesReversed() order: slot seq_offsets[t] + n_states[t] - 1 - j
    track_of_row = np.repeat(np.arange(n_out), ns)
    starts = np.zeros(n_out + 1, np.int64)
    np.cumsum(ns, out=starts[1:])
    j = np.arange(starts[-1]) - starts[track_of_row]
    rows = seq_offsets[fitted][track_of_row] + ns[track_of_row] - 1 - j
    stype = table.ints[capi.TS_STATE_TYPE, rows]
    assert np.all(stype >= 0), "a fitted track's state is missing from the table"
    has = table.ints[capi.TS_HAS_PARAMS, rows]
    link = table.ints[capi.TS_DIM_HIT, rows] > 0
    every = np.ones(len(rows), bool)

    out: dict = {"event_nr": np.full(n_out, event_nr, np.uint32),
                 "track_nr": fitted.astype(np.uint32),
                 "nStates": ns.astype(np.uint32),
                 "nMeasurements": n_measurements[fitted].astype(np.uint32)}

    def vec(name, values, keep=every):
        out[name] = _ragged(values, keep, track_of_row, n_out)

    surf = state_surface[rows]
    if geometry_ids is None:
        ids = np.zeros((len(rows), 3), np.int32)
        ids[:, 2] = surf
    else:
        ids = np.asarray(geometry_ids, np.int32)[surf]
    vec("volume_id", ids[:, 0])
    vec("layer_id", ids[:, 1])
    vec("module_id", ids[:, 2])
    vec("stateType", stype)
    for name in ("chi2", "pathLength", "t_x", "t_y", "t_z", "t_r", "t_dx", "t_dy", "t_dz",
                 *[f"t_{p}" for p in PARAMS], "l_x_hit", "l_y_hit", "g_x_hit", "g_y_hit",
                 "g_z_hit"):
        vec(name, table.column(name)[rows])
    prt = (has & 1) != 0
    hit_keep = link & prt | ~prt
    for name in ("res_x_hit", "res_y_hit", "err_x_hit", "err_y_hit", "pull_x_hit",
                 "pull_y_hit"):
        vec(name, table.column(name)[rows], hit_keep)
    vec("dim_hit", np.where(prt, table.ints[capi.TS_DIM_HIT, rows], 0), hit_keep)
    for p, s in enumerate(PARAM_SETS):
        present = (has >> p & 1) != 0
        out[PARAM_COUNTS[p]] = np.bincount(track_of_row[present], minlength=n_out).astype(
            np.int32)
        vec(PARAM_SET_FLAGS[p], present)
        for q in PARAMS:
            vec(f"{q}_{s}", table.column(f"{q}_{s}")[rows])
            vec(f"err_{q}_{s}", table.column(f"err_{q}_{s}")[rows])
        for q in ("g_x", "g_y", "g_z", "px", "py", "pz", "eta", "pT"):
            vec(f"{q}_{s}", table.column(f"{q}_{s}")[rows])
        keep = present & link | ~present
        for q in PARAMS:
            vec(f"res_{q}_{s}", table.column(f"res_{q}_{s}")[rows], keep)
            vec(f"pull_{q}_{s}", table.column(f"pull_{q}_{s}")[rows], keep)
    return out


def write_parquet(path: str, tr: dict) -> None:
    """Store a tree() as a Parquet file with the writer's branch names (one
    row per track, list columns for the vector branches)."""
    import pyarrow as pa
    import pyarrow.parquet as pq
    cols = {}
    for name, v in tr.items():
        if isinstance(v, tuple):
            values, offsets = v
            cols[name] = pa.ListArray.from_arrays(pa.array(offsets.astype(np.int32)),
                                                  pa.array(values))
        else:
            cols[name] = pa.array(v)
    pq.write_table(pa.table(cols), path)
