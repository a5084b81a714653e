"""The C-ABI library loads and exports every function include/c3d_hip.h declares."""
import ctypes
import os
import re

from conftest import ROOT, product


def declared_functions():
    text = open(os.path.join(ROOT, "include", "c3d_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:void|int)\s+(\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ("solveFluxKinetic_hip", "solveH_hip", "solveHInit_hip", "c3d_closure",
                 "c3d_update", "c3d_flux_kinetic", "c3d_flux_moment", "c3d_ghost_fill"):
        assert must in names


def test_library_exports_every_declared_symbol():
    P = product()
    lib = P.load()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_front_end_covers_header():
    P = product()
    bound = {n for n, _, _ in P.SIGNATURES}
    assert set(declared_functions()) == bound


def test_library_is_gfx950_and_not_oracle():
    # the product must not link the CPU oracle
    path = os.path.join(ROOT, "c3d-2015_amd", "libc3d_hip.so")
    data = open(path, "rb").read()
    assert b"orc_closure" not in data and b"liboracle" not in data
    assert b"gfx950" in data
