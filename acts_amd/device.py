"""Device-resident batches (torch tensors on the HIP device as plain memory).

torch is used only to own device memory and to expose its current stream;
the fit itself runs in libacts_kf.so through acts_kf_fit_batch_device.
"""
from __future__ import annotations

import numpy as np
import torch

from . import capi
from .events import Events
from .fitter import STATE_ARRAYS, TRACK_ARRAYS, TRACK_VECTOR_ARRAYS

_TORCH_DT = {np.float64: torch.float64, np.float32: torch.float32,
             np.int32: torch.int32, np.int64: torch.int64, np.uint32: torch.int32}


def to_device(arr: np.ndarray, device) -> torch.Tensor:
    t = torch.from_numpy(np.ascontiguousarray(arr).view(
        np.int32 if arr.dtype == np.uint32 else arr.dtype))
    return t.to(device)


class DeviceEvents:
    """The input arrays of a batch, resident on `device`."""

    NAMES = ("start_params", "start_cov", "start_plane", "seq_offsets", "seq_surfaces",
             "meas_offsets", "meas_surfaces", "meas_values", "meas_cov")

    def __init__(self, events: Events, device="cuda:0"):
        self.n_tracks = events.n_tracks
        self.max_sequence_length = events.max_sequence_length
        self.n_slots = int(events.seq_offsets[-1])
        self.t = {name: to_device(getattr(events, name), device) for name in self.NAMES}
        for name in ("meas_dims", "meas_indices"):
            arr = getattr(events, name, None)
            if arr is not None:
                self.t[name] = torch.from_numpy(np.ascontiguousarray(arr)).to(device)

    def batch_in(self) -> capi.BatchIn:
        b = capi.BatchIn()
        b.n_tracks = self.n_tracks
        for name in self.NAMES + ("meas_dims", "meas_indices"):
            if name in self.t:
                setattr(b, name, self.t[name].data_ptr())
        b.max_sequence_length = self.max_sequence_length
        return b


class DeviceResult:
    """Output arrays of a batch on `device` (per-state arrays on request)."""

    def __init__(self, n_tracks: int, n_slots: int, outputs, device="cuda:0"):
        self.n_tracks = n_tracks
        self.t = {}
        for name, dt in TRACK_ARRAYS.items():
            self.t[name] = torch.zeros(n_tracks, dtype=_TORCH_DT[dt], device=device)
        for name in outputs:
            if name in TRACK_VECTOR_ARRAYS:
                self.t[name] = torch.zeros((n_tracks, TRACK_VECTOR_ARRAYS[name]),
                                           dtype=torch.float64, device=device)
                continue
            dt, w = STATE_ARRAYS[name]
            shape = (n_slots,) if w == 1 else (n_slots, w)
            self.t[name] = torch.zeros(shape, dtype=_TORCH_DT[dt], device=device)

    def batch_out(self) -> capi.BatchOut:
        b = capi.BatchOut()
        for name in capi.BATCH_OUT_FIELDS:
            t = self.t.get(name)
            setattr(b, name, t.data_ptr() if t is not None else None)
        return b

    def host(self, name: str) -> np.ndarray:
        return self.t[name].cpu().numpy()

    def host_rows(self, name: str, rows: np.ndarray) -> np.ndarray:
        """Rows `rows` of an output array, gathered on the device first."""
        t = self.t[name]
        idx = torch.from_numpy(np.ascontiguousarray(rows, dtype=np.int64)).to(t.device)
        return t.index_select(0, idx).cpu().numpy()
