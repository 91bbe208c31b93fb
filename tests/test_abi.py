"""CPU checks of the drop-in boundary: the library builds, loads and exports every entry
point include/mpf_hip.h declares, with the argument lists the ctypes binding uses; and the
host-side plugin API keeps the reference's names, checks and messages."""

from __future__ import annotations

import ctypes
import re
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "mpf_hip.h"


def declared() -> set[str]:
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"^\s*(?:const\s+)?[\w\s\*]+?\b(mpf_\w+)\s*\(", text, flags=re.M))


def test_library_exports_every_declared_entry_point():
    from mpifield_amd import _lib

    names = declared()
    assert len(names) >= 35
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # the binding declares exactly the header's entry points
    assert set(_lib.SIGNATURES) == names


def test_abi_version_and_error_slot_work_without_a_gpu():
    from mpifield_amd import _lib

    lib = _lib.load()
    assert lib.mpf_abi_version() == 1
    # the slot keeps a thread's last error, so it is empty only in a fresh process
    fresh = subprocess.run([sys.executable, "-c", "import ctypes, sys; lib = ctypes.CDLL(sys.argv[1]); "
                            "lib.mpf_last_error.restype = ctypes.c_char_p; print(repr(lib.mpf_last_error()))",
                            str(_lib.LIB_PATH)], capture_output=True, text=True, check=True)
    assert fresh.stdout.strip() == "b''"


def test_argument_errors_need_no_gpu():
    """Argument checks run on the host before any device work."""
    from mpifield_amd import _lib

    lib = _lib.load()
    lower = np.zeros(1)
    cell = np.ones(1)
    num = np.array([5], dtype=np.int64)
    code = np.array([0], dtype=np.int64)
    rc = lib.mpf_scatter(2, 9, 1, None, 0, None, 0, 1, 0, 1, None, lower.ctypes.data, cell.ctypes.data,
                         num.ctypes.data, code.ctypes.data, 0, _lib.SCALE_AUTO, None, None)
    assert rc == 1 and b"spline degree" in lib.mpf_last_error()
    code2 = np.array([2], dtype=np.int64)
    rc = lib.mpf_gather(1, 0, 1, None, 0, 1, 0, None, 1, None, 0, lower.ctypes.data, cell.ctypes.data,
                        num.ctypes.data, code2.ctypes.data, None, None)
    assert rc == 1 and b"clamped" in lib.mpf_last_error()
    rc = lib.mpf_deposit_paths(1, 5, 1, None, 0, None, 0, None, 1, 0, None, lower.ctypes.data, cell.ctypes.data,
                               num.ctypes.data, code.ctypes.data, 0, _lib.SCALE_AUTO, None, None)
    assert rc == 1 and b"spline edge degrees" in lib.mpf_last_error()
    # the bins' limits: 32-bit offsets and at most 64 value rows, refused before any device work
    rc = lib.mpf_bin_markers(1, 2, 1, None, 0, None, 0, 1, 1 << 31, 0, lower.ctypes.data, cell.ctypes.data,
                             num.ctypes.data, code.ctypes.data, None, None, None, None, None, _lib.SCALE_AUTO, None,
                             None)
    assert rc == 1 and b"n < 2^31" in lib.mpf_last_error()
    rc = lib.mpf_bin_markers(1, 2, 1, None, 0, None, 0, 1, 10, 65, lower.ctypes.data, cell.ctypes.data,
                             num.ctypes.data, code.ctypes.data, None, None, None, None, None, _lib.SCALE_AUTO, None,
                             None)
    assert rc == 1 and b"ncomp <= 64" in lib.mpf_last_error()


def test_plugin_api_names_and_messages():
    import mpifield_amd as m

    assert repr(m.QUADRATIC) == "Weighting('quadratic', ndim=(1, 2, 3), halo=2, degree=2)"
    assert repr(m.CUBIC) == "Weighting('cubic', ndim=(1, 2, 3), halo=2, degree=3)"
    assert repr(m.QUADRATIC_EDGES) == "EdgeWeighting('quadratic', ndim=(1, 2, 3), halo=2, degree=2)"
    assert (m.WHITNEY.degree, m.QUADRATIC_EDGES.degree, m.CUBIC_EDGES.degree) == (1, 2, 3)
    with pytest.raises(ValueError, match="1, 2 or 3 dimensions"):
        m.Weighting("bad", scatter={4: print}, gather={})
    with pytest.raises(ValueError, match="has no 2D scatter kernel"):
        m.Weighting("mine", scatter={1: print}, gather={}).scatter_kernel(2)
    with pytest.raises(TypeError):
        m.LINEAR.scatter[2] = print  # read-only mapping
    with pytest.raises(ValueError, match="does not take clamped axes"):
        m.QUADRATIC.scatter_block(np.zeros((1, 1)), np.ones((1, 1)), np.zeros((4, 1)),
                                  m.BlockGrid.from_geometry([0.0], [1.0], [4], clamped=[True]))
    assert m.splines(2).edges is m.QUADRATIC_EDGES and m.splines(3, clamped=True).nodes is m.CUBIC
    with pytest.raises(ValueError, match="no built-in splines of degree 5"):
        m.splines(5)


def test_block_grid_from_geometry():
    """test_weighting.py:190-197."""
    from mpifield_amd import BlockGrid

    grid = BlockGrid.from_geometry((0.0, 0.0), (0.5, 0.5), (3, 3), (True, False))
    assert grid.geometry == (0.0, 0.0, 0.5, 0.5, 3, 3, 1, 0)
    with pytest.raises(ValueError, match="differ in length"):
        BlockGrid.from_geometry((0.0,), (0.5, 0.5), (3, 3))
    with pytest.raises(ValueError, match="either periodic or clamped"):
        BlockGrid.from_geometry((0.0,), (0.5,), (3,), (True,), (True,))


def test_product_path_fails_loudly_without_a_gpu(monkeypatch):
    import torch

    from mpifield_amd import _lib
    from mpifield_amd import kernels as K

    monkeypatch.setattr(torch.cuda, "is_available", lambda: False)
    with pytest.raises(_lib.MpfError, match="needs a ROCm GPU"):
        K.device()


def test_missing_library_is_an_error(tmp_path):
    from mpifield_amd import _lib

    with pytest.raises(_lib.MpfError, match="no CPU fallback"):
        _lib.load(tmp_path / "absent.so")


def test_binned_route_stops_below_two_to_the_31_markers():
    """Bins hold 32-bit offsets: a batch of 2^31 markers or more takes the direct kernels (the
    rule of mpf_binned.hip's binned_route, which the autograd Functions share)."""
    from mpifield_amd import kernels as K

    num, code = (256, 256, 256), (1, 1, 1)
    assert K.binned_route(1, 2, 3, num, code, (1 << 31) - 1)
    assert not K.binned_route(1, 2, 3, num, code, 1 << 31)
    assert not K.binned_route(1, 2, 3, num, code, K.BINNED_MIN_MARKERS - 1)
    assert K.binned_route(1, 2, 3, num, code, K.BINNED_MIN_MARKERS)
    # clamped axes bin for the 3D splines only; any degree and component count does on tiles
    assert K.binned_route(2, 3, 3, num, (2, 1, 1), 1 << 20)
    assert not K.binned_route(1, 2, 3, num, (2, 1, 1), 1 << 20)
    assert not K.binned_route(2, 2, 1, (256, 256), (2, 1), 1 << 20)
    assert not K.binned_route(2, 2, 1, (7, 64, 64), (2, 1, 1), 1 << 20)  # a clamped axis shorter than a tile
    assert not K.derivative_tiled_ok(2, 2, 1, num, (2, 1, 1))
    assert K.binned_route(2, 3, 3, num, code, 1 << 20)
    assert K.binned_route(1, 2, 5, (100, 36, 52), (1, 1, 1), 1 << 20)  # partial periodic tiles
    # a periodic axis needs two tiles and a footprint: 8 x 8 x 16 tiles, quadratic stencils
    assert not K.binned_route(1, 2, 3, (9, 64, 64), (1, 1, 1), 1 << 20)
    assert K.binned_route(1, 2, 3, (10, 64, 64), (1, 1, 1), 1 << 20)
    assert K.binned_route(1, 2, 3, (9, 64, 64), (0, 1, 1), 1 << 20)
    # the tiled gradient keeps its narrower rule
    assert not K.derivative_tiled_ok(2, 3, 3, num, code)
    assert not K.derivative_tiled_ok(1, 2, 3, (100, 36, 52), (1, 1, 1))
    assert K.derivative_tiled_ok(1, 2, 3, num, code)


def test_permutation_plan_sizes_and_argument_errors_need_no_gpu():
    """mpf_perm_plan_bytes: q (2 B), dst (4 B per slot of whole 2^15 chunks) and lo (2 B) per
    marker, each part 256-byte aligned; no plan outside 1 .. 2^28 markers; mpf_perm_rows
    refuses a bad direction, dtype or stride before any device work."""
    from mpifield_amd import _lib

    lib = _lib.load()

    def size(n):
        out = ctypes.c_int64(-1)
        assert lib.mpf_perm_plan_bytes(n, ctypes.addressof(out)) == 0
        return out.value

    def a256(b):
        return (b + 255) // 256 * 256

    for n in (1, 1000, 1 << 15, (1 << 15) + 1, 3_000_000, 1 << 28):
        chunks = -(-n // (1 << 15))
        assert size(n) == 2 * a256(2 * n) + a256(4 * chunks * (1 << 15)), n
    assert size(0) == 0 and size((1 << 28) + 1) == 0 and size(-5) == 0
    assert lib.mpf_perm_plan_bytes(10, None) == 1
    rc = lib.mpf_perm_rows(None, 10, 2, None, 10, None, 10, 1, _lib.F32, None, None)
    assert rc == 1 and b"direction" in lib.mpf_last_error()
    rc = lib.mpf_perm_rows(None, 10, 0, None, 10, None, 10, 1, 7, None, None)
    assert rc == 1 and b"dtype" in lib.mpf_last_error()
    rc = lib.mpf_perm_rows(None, 10, 0, None, 9, None, 10, 1, _lib.F32, None, None)
    assert rc == 1 and b"strides" in lib.mpf_last_error()
    rc = lib.mpf_perm_rows(None, (1 << 28) + 1, 0, None, 1 << 29, None, 1 << 29, 1, _lib.F32, None, None)
    assert rc == 1 and b"2^28" in lib.mpf_last_error()
    # nothing to move is not an error
    assert lib.mpf_perm_rows(None, 0, 0, None, 0, None, 0, 3, _lib.F32, None, None) == 0
