"""CPU checks of the C ABI: the HIP library loads and exports every entry point that
include/wl.h declares (no compute calls: this container has no GPU)."""
import ctypes
import os
import re

import pytest

from __graft_entry__ import ROOT, load_package

HEADER = os.path.join(ROOT, "include", "wl.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\*?\s+\*?(wl_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    for s in ("wl_flow_create", "wl_mom_step", "wl_pois_solve", "wl_pressure_force", "wl_last_error"):
        assert s in syms


def test_library_exports_every_declared_symbol():
    wl = load_package()
    if not os.path.exists(wl.LIB_PATH):
        pytest.fail(f"{wl.LIB_PATH} is not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(wl.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_python_binding_covers_header():
    wl = load_package()
    from waterlily_amd import _abi
    assert set(declared_symbols()) == set(_abi.SIGNATURES)


def test_host_helpers():
    wl = load_package()
    # coarsen_mask KATs of test/test_poisson.jl:44-46
    assert wl.coarsen_mask((18, 18, 6)) == (True, True, True)
    assert wl.coarsen_mask((18, 18, 4)) == (True, True, False)
    assert wl.coarsen_mask((18, 17, 6)) == (True, False, True)
    import numpy as np
    p = np.zeros((4, 5))
    p[...] = (np.arange(4)[:, None] + 1 - 1.5) + (np.arange(5)[None, :] + 1 - 1.5) + 3
    assert wl.L2(p) == 187  # test/test_flow.jl:46-49
