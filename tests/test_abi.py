"""The C-ABI library loads and exports every entry point include/orgpu.h declares (CPU only)."""
import ctypes
import os
import re

import pytest

from orthoroute_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "orgpu.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(orgpu_[a-z_0-9]+)\s*\(", txt)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "orgpu_route" in syms and "orgpu_lattice_build" in syms
    assert set(_abi.EXPORTED) == set(syms)


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_abi.lib_path()):
        pytest.skip("liborgpu.so not built (run python -m orthoroute_amd.build)")
    lib = ctypes.CDLL(_abi.lib_path())
    for s in declared_symbols():
        assert hasattr(lib, s), s


def test_lattice_count_host_only():
    """orgpu_lattice_count needs no GPU: closed-form sizes from the header's descriptor."""
    if not os.path.exists(_abi.lib_path()):
        pytest.skip("liborgpu.so not built")
    from orthoroute_amd.device import lattice_counts
    from orthoroute_amd.lattice import edge_count
    assert lattice_counts(96, 96, 6, "reference") == (55_296, 441_600)
    assert lattice_counts(96, 96, 6, "adjacent") == (55_296, 201_600)
    assert lattice_counts(600, 400, 18, "adjacent") == (4_320_000, 16_782_000)
    assert lattice_counts(600, 400, 18, "reference") == (4_320_000, 138_224_000)
    assert lattice_counts(300, 300, 12, "reference")[1] == 21_594_000
    assert lattice_counts(1000, 800, 24, "adjacent")[1] == 75_156_800
    assert edge_count(12, 10, 6, "reference") == lattice_counts(12, 10, 6, "reference")[1]


def test_error_reporting():
    if not os.path.exists(_abi.lib_path()):
        pytest.skip("liborgpu.so not built")
    lib = _abi.load()
    d = _abi.LatticeDesc()
    d.X, d.Y, d.L = 0, 5, 4
    n, e = ctypes.c_int64(), ctypes.c_int64()
    rc = lib.orgpu_lattice_count(ctypes.byref(d), ctypes.byref(n), ctypes.byref(e))
    assert rc == -1
    assert b"bad shape" in lib.orgpu_last_error()


def test_solver_seam_fails_loudly_without_a_gpu():
    """No silent CPU fallback: the solver seam raises where no GPU is visible
    (the reference's caller then takes its own CPU path, UPF:1819-1821)."""
    import numpy as np
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    from orthoroute_amd.solver import GpuSolver
    with pytest.raises(RuntimeError):
        GpuSolver(np.array([0, 1, 1], np.int32), np.array([1], np.int32))
