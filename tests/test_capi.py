"""The C ABI library: it loads without a GPU and exports every entry point
include/msim_hip.h declares (no compute calls here)."""
from __future__ import annotations

import ctypes
import os
import re

import microsimulator_amd as msa

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_entry_points():
    names = set()
    include = os.path.join(ROOT, "include")
    for header in os.listdir(include):
        if header.endswith(".h"):
            text = open(os.path.join(include, header)).read()
            names |= set(re.findall(r"\b(msim_hip_\w+)\s*\(", text))
    return names


def ensure_built():
    if not os.path.exists(msa.LIBRARY):
        from microsimulator_amd import build
        build.build()


def test_library_exports_every_declared_entry_point():
    ensure_built()
    declared = declared_entry_points()
    assert declared == set(msa.EXPORTS)
    lib = ctypes.CDLL(msa.LIBRARY)
    for name in sorted(declared):
        assert hasattr(lib, name), name


def test_python_mirror_binds_the_abi():
    ensure_built()
    lib = msa.library()
    for name in msa.EXPORTS:
        assert getattr(lib, name).restype is not None or name == "msim_hip_destroy"


def test_unsupported_features_are_not_advertised():
    assert msa.SUPPORTED == {"signals", "depth_averaged_flow"}
    assert set(msa.FEATURES) >= msa.SUPPORTED


def test_depth_averaged_entry_points_are_implemented():
    """Both depth-averaged ComputeBackend calls (backend.hpp:44-49) run the
    C ABI; only the resolved and cell-mechanics calls are refused."""
    stub = msa.HipBackend._unsupported
    assert msa.HipBackend.solve_depth_averaged_flow is not stub
    assert msa.HipBackend.solve_depth_averaged_port_flow is not stub
    assert msa.HipBackend.solve_resolved_flow is stub


def test_kernels_are_built_for_gfx950_only():
    ensure_built()
    blob = open(msa.LIBRARY, "rb").read()
    assert b"gfx950" in blob
    for other in (b"gfx942", b"gfx90a", b"gfx1100"):
        assert other not in blob
