"""The C-ABI library loads and exports every function include/*.h declares.

CPU only: no device is touched (only pure host entry points are called).
"""
import ctypes as C
import os
import re

import pytest

from acts_amd import capi

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(REPO, "include", h) for h in ("acts_kf.h", "acts_kf_gen.h")]
DECL = re.compile(r"^[A-Za-z_][\w\s\*]*?\b(acts_kf_\w+)\s*\(", re.M)


def declared_functions():
    names = []
    for h in HEADERS:
        with open(h) as f:
            text = re.sub(r"/\*.*?\*/", "", f.read(), flags=re.S)
        names += DECL.findall(text)
    return sorted(set(names))


def test_headers_declare_entry_points():
    names = declared_functions()
    for must in ("acts_kf_create", "acts_kf_fit_batch", "acts_kf_fit_batch_device",
                 "acts_kf_sync", "acts_kf_destroy", "acts_kf_status_string"):
        assert must in names


def test_library_exports_every_declared_symbol(lib):
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_ctypes_mirror_covers_every_declared_symbol():
    assert sorted(capi.SIGNATURES) == declared_functions()


def test_default_options_are_the_reference_defaults():
    o = capi.default_options()
    assert o.multiple_scattering == 1 and o.energy_loss == 1 and o.use_joseph == 0
    assert o.step_tolerance == pytest.approx(1e-4)
    assert o.max_steps == 1000 and o.max_target_skipping == 100
    assert o.max_runge_kutta_step_trials == 10000
    assert o.initial_step_size == 10000.0
    assert o.loop_protection == 1 and o.loop_fraction == 0.5
    assert o.near_limit == pytest.approx(-0.1)
    assert o.abs_pdg == 13
    assert o.mass == pytest.approx(0.1056583755, rel=1e-7)


def test_status_strings():
    assert "ok" in capi.status_string(capi.STATUS_OK).lower() or capi.status_string(0)
    for code in (capi.STATUS_UPDATE_FAILED, capi.STATUS_SMOOTH_FAILED,
                 capi.STATUS_STEP_COUNT_LIMIT, capi.STATUS_NOT_ON_SURFACE):
        assert capi.status_string(code)
    assert capi.status_string(0x7777)


def test_structs_match_header_sizes():
    # the ctypes layouts follow the header field by field; spot-check sizes
    assert C.sizeof(capi.Surface) == 4 * 4 + 8 * 14
    assert C.sizeof(capi.MaterialSlab) == 6 * 4
    assert C.sizeof(capi.BatchOut) == 8 * len(capi.BATCH_OUT_FIELDS)
    assert C.sizeof(capi.BatchIn) == 8 * 10 + 8 + 2 * 8


@pytest.mark.parametrize("field,value", [("smoother", 7), ("stepper", 2), ("precision", 5),
                                         ("reference_surface", 9), ("reference_strategy", 3)])
def test_create_rejects_unknown_options(lib, field, value):
    """Options are validated before any device call: an unknown smoother,
    precision mode, reference surface or extrapolation strategy fails
    acts_kf_create with an error code and a message naming the option."""
    import acts_amd as A
    det = A.barrel_detector()
    o = A.KalmanFitterOptions(reference_surface=capi.REFERENCE_PERIGEE).struct()
    setattr(o, field, value)
    surf, mats = det.surface_table(), det.material_table()
    fld = A.ConstantField().struct()
    ctx = C.c_void_p()
    rc = lib.acts_kf_create(surf, len(det.surfaces), mats, len(det.materials), C.byref(fld),
                            C.byref(o), 0, C.byref(ctx))
    assert rc in (capi.ERR_UNSUPPORTED, capi.ERR_INVALID_ARGUMENT)
    assert not ctx.value
    msg = capi.last_error().lower()
    assert field.split("_")[-1] in msg or field.split("_")[0] in msg, msg

