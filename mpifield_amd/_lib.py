"""ctypes binding of libmpfhip.so (include/mpf_hip.h).

The library is the only compute path: importing a kernel wrapper without the built
library, or calling one without a ROCm GPU, raises -- there is no CPU fallback.
"""

from __future__ import annotations

import ctypes
import os
from pathlib import Path

_HERE = Path(__file__).resolve().parent
LIB_PATH = Path(os.environ.get("MPF_HIP_LIBRARY", _HERE / "libmpfhip.so"))

_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_INT = ctypes.c_int
_DBL = ctypes.c_double

MPF_OK = 0
F32, F64 = 0, 1
MODE_ATOMIC, MODE_FIXED = 0, 1
SCALE_AUTO = -100000
KIND_LINEAR, KIND_QUADRATIC, KIND_SPLINE = 0, 1, 2


class MpfError(RuntimeError):
    """A failed call into libmpfhip.so, with the library's message."""


def _geom(ndim: int) -> list:
    return [_DBL] * (2 * ndim) + [_I64] * (2 * ndim)


_SCATTER_IN = [_P, _I64, _P, _I64, _INT, _I64, _I64, _P]
_GATHER_IN = [_P, _I64, _INT, _I64, _P, _I64, _P, _I64]
_PATHS_IN = [_P, _I64, _P, _I64, _P, _INT, _I64, _P]
_EDGES_GATHER_IN = [_P, _I64, _INT, _I64, _P, _P, _I64]
_TAIL = [_P, _P]

#: every exported entry point of include/mpf_hip.h with its argument types
SIGNATURES: dict[str, list] = {
    "mpf_abi_version": [],
    "mpf_last_error": [],
    "mpf_context_create": [_INT],
    "mpf_context_destroy": [_P],
    "mpf_profile_enable": [_INT],
    "mpf_profile_reset": [],
    "mpf_profile_query": [ctypes.c_char_p, _P, _P],
    "mpf_scatter": [_INT, _INT, _INT, _P, _I64, _P, _I64, _INT, _I64, _I64, _P, _P, _P, _P, _P, _INT, _INT, _P, _P],
    "mpf_scatter_noalloc": [_INT, _INT, _INT, _P, _I64, _P, _I64, _INT, _I64, _I64, _P, _P, _P, _P, _P, _INT, _INT,
                            _P],
    "mpf_gather": [_INT, _INT, _INT, _P, _I64, _INT, _I64, _P, _I64, _P, _I64, _P, _P, _P, _P, _P, _P],
    "mpf_set_binned_min": [_I64],
    "mpf_tiled_variant": [_INT, _INT, _INT, _INT, _INT],
    "mpf_tiled_variant_shape": [_INT, _P, _P],
    "mpf_gather_derivative": [_INT, _INT, _INT, _P, _I64, _INT, _I64, _P, _I64, _P, _I64, _P, _P, _P, _P, _P, _P],
    "mpf_deposit_paths": [_INT, _INT, _INT, _P, _I64, _P, _I64, _P, _INT, _I64, _P, _P, _P, _P, _P, _INT, _INT, _P,
                          _P],
    "mpf_deposit_paths_binned": [_INT, _INT, _INT, _P, _P, _I64, _P, _INT, _I64, _P, _P, _P, _P, _P, _P, _P, _P],
    "mpf_halo_fill": [_INT, _P, _I64, _I64, _P, _P, _P],
    "mpf_halo_fold": [_INT, _P, _I64, _I64, _P, _P, _P],
    "mpf_divergence": [_INT, _P, _P, _P, _P, _P],
    "mpf_gradient": [_INT, _P, _P, _P, _P, _P],
    "mpf_bin_sizes": [_INT, _P, _P],
    "mpf_patch_plan_create": [_INT, _I64, _P, _I64, _I64],
    "mpf_patch_plan_destroy": [_P],
    "mpf_patch_plan_items": [_P],
    "mpf_patches_scatter": [_P, _P, _I64, _P, _I64, _INT, _I64, _P, _INT, _INT, _P, _P],
    "mpf_patches_gather": [_P, _P, _I64, _INT, _P, _I64, _P, _I64, _INT, _P, _P],
    "mpf_bin_markers": [_INT, _INT, _INT, _P, _I64, _P, _I64, _INT, _I64, _I64, _P, _P, _P, _P, _P, _P, _P, _P, _P,
                        _INT, _P, _P],
    "mpf_scatter_binned": [_INT, _INT, _INT, _P, _P, _INT, _I64, _I64, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P,
                           _P],
    "mpf_gather_binned": [_INT, _INT, _INT, _P, _INT, _I64, _P, _P, _P, _P, _I64, _P, _P, _P, _P, _P, _P, _P],
    "mpf_rebin_markers": [_INT, _INT, _INT, _P, _I64, _P, _I64, _INT, _I64, _I64, _P, _P, _P, _P, _P, _P, _P, _P,
                          _P, _P, _P, _INT, _P, _P],
    "mpf_bin_items": [_INT, _P, _P, _P, _P, _P, _P],
    "mpf_gather_derivative_binned": [_INT, _INT, _INT, _P, _INT, _I64, _P, _P, _P, _I64, _P, _I64, _P, _I64, _P, _P,
                                     _P, _P, _P, _P],
    "mpf_fixed_scale": [_P, _I64, _INT, _I64, _I64, _INT, _P, _P, _P],
    "mpf_unsort_rows": [_P, _P, _P, _I64, _I64, _INT, _P],
    "mpf_perm_plan_bytes": [_I64, _P],
    "mpf_perm_plan_build": [_P, _I64, _P, _P, _P],
    "mpf_perm_rows": [_P, _I64, _INT, _P, _I64, _P, _I64, _I64, _INT, _P, _P],
    "mpf_gather_edges": [_INT, _INT, _INT, _P, _I64, _INT, _I64, _P, _P, _I64, _P, _P, _P, _P, _P, _P],
    "mpf_scatter_forms": [_INT, _INT, _INT, _P, _I64, _P, _I64, _INT, _I64, _I64, _P, _P, _P, _P, _P, _INT, _INT, _P,
                          _P],
    "mpf_gather_forms": [_INT, _INT, _INT, _P, _I64, _INT, _I64, _P, _I64, _P, _I64, _P, _P, _P, _P, _P, _P],
    "mpf_scatter_forms_binned": [_INT, _INT, _INT, _P, _P, _INT, _I64, _I64, _P, _P, _P, _P, _P, _P, _P, _P, _P],
    "mpf_gather_forms_binned": [_INT, _INT, _INT, _P, _INT, _I64, _P, _P, _I64, _P, _P, _P, _P, _P, _P, _P],
}
for _d in (1, 2, 3):
    for _kind in ("linear", "quadratic"):
        SIGNATURES[f"mpf_scatter_{_kind}_{_d}d"] = _SCATTER_IN + _geom(_d) + [_INT] + _TAIL
        SIGNATURES[f"mpf_gather_{_kind}_{_d}d"] = _GATHER_IN + _geom(_d) + _TAIL
    SIGNATURES[f"mpf_scatter_spline_{_d}d"] = _SCATTER_IN + _geom(_d) + [_INT, _INT] + _TAIL
    SIGNATURES[f"mpf_gather_spline_{_d}d"] = _GATHER_IN + _geom(_d) + [_INT] + _TAIL
    SIGNATURES[f"mpf_deposit_paths_{_d}d"] = _PATHS_IN + _geom(_d) + [_INT] + _TAIL
    SIGNATURES[f"mpf_deposit_paths_spline_{_d}d"] = _PATHS_IN + _geom(_d) + [_INT, _INT] + _TAIL
    SIGNATURES[f"mpf_gather_edges_{_d}d"] = _EDGES_GATHER_IN + _geom(_d) + _TAIL
    SIGNATURES[f"mpf_gather_edges_spline_{_d}d"] = _EDGES_GATHER_IN + _geom(_d) + [_INT] + _TAIL
    SIGNATURES[f"mpf_scatter_forms_spline_{_d}d"] = _SCATTER_IN + _geom(_d) + [_INT, _INT, _INT] + _TAIL
    SIGNATURES[f"mpf_gather_forms_spline_{_d}d"] = _GATHER_IN + _geom(_d) + [_INT, _INT] + _TAIL

_RESTYPES = {"mpf_last_error": ctypes.c_char_p, "mpf_context_create": _P, "mpf_context_destroy": None,
             "mpf_set_binned_min": _I64,
             "mpf_patch_plan_create": _P, "mpf_patch_plan_destroy": None, "mpf_patch_plan_items": _I64}


class Patch(ctypes.Structure):
    """The C struct mpf_patch of include/mpf_hip.h."""

    _fields_ = [
        ("lower", ctypes.c_double * 3),
        ("cell_size", ctypes.c_double * 3),
        ("num_nodes", ctypes.c_int64 * 3),
        ("code", ctypes.c_int64 * 3),
        ("kind", ctypes.c_int32),
        ("degree", ctypes.c_int32),
        ("first_marker", ctypes.c_int64),
        ("num_markers", ctypes.c_int64),
        ("field_offset", ctypes.c_int64),
    ]

_LIB: ctypes.CDLL | None = None


def load(path: Path | str | None = None) -> ctypes.CDLL:
    """Load libmpfhip.so and declare its signatures; raises if it is not built."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    where = Path(path) if path is not None else LIB_PATH
    if not where.exists():
        raise MpfError(
            f"{where} is not built: run `python -m mpifield_amd.csrc.build` (or __graft_entry__.build()); "
            "mpifield_amd has no CPU fallback"
        )
    lib = ctypes.CDLL(str(where))
    for name, argtypes in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = argtypes
        fn.restype = _RESTYPES.get(name, _INT)
    if path is None:
        _LIB = lib
        # the launch variants and binning threshold tuned on the MI355X (autotune.py)
        from mpifield_amd import autotune

        table = autotune.TABLE
        if table.exists():
            import json

            data = json.loads(table.read_text())
            autotune.apply_table(lib, data)
            if "binned" in data:
                lib.mpf_set_binned_min(int(data["binned"]["binned_min"]))
    return lib


def check(rc: int) -> None:
    """Raise MpfError with the library's message for a non-zero status."""
    if rc != MPF_OK:
        message = load().mpf_last_error().decode(errors="replace")
        kind = {1: ValueError, 3: NotImplementedError}.get(rc)
        if kind is not None:
            raise kind(message)
        raise MpfError(f"libmpfhip.so: {message} (status {rc})")


def call(name: str, *args) -> None:
    check(getattr(load(), name)(*args))
