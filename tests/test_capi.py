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


def test_only_implemented_features_are_advertised():
    assert msa.SUPPORTED == {"signals", "coupled_rates", "depth_averaged_flow", "resolved_flow"}
    assert set(msa.FEATURES) >= msa.SUPPORTED


def test_flow_entry_points_are_implemented():
    """Both depth-averaged ComputeBackend calls (backend.hpp:44-49), both
    resolved ones (50-55) and the coupled step (31-34) run the C ABI; the
    cell-mechanics calls are refused."""
    stub = msa.HipBackend._unsupported
    assert msa.HipBackend.solve_depth_averaged_flow is not stub
    assert msa.HipBackend.solve_depth_averaged_port_flow is not stub
    assert msa.HipBackend.advance_coupled is not stub
    assert msa.HipBackend.solve_resolved_flow is not stub
    assert msa.HipBackend.solve_resolved_port_flow is not stub
    assert msa.HipBackend.advance_species is stub


def test_rate_ops_follow_the_reference_order():
    """RateOp's codes (coupled_rates.hpp) as the C ABI's MSIM_RATE_* and the
    oracle number them."""
    from tests.helpers import OP_SELECT, OP_SIGNAL, OP_SURFACE, OP_VOLUME_CHANGE_RATE
    assert msa.RateOp.SELECT == OP_SELECT == 26
    assert msa.RateOp.SIGNAL == OP_SIGNAL == 27
    assert msa.RateOp.SURFACE_AREA == OP_SURFACE == 10
    assert msa.RateOp.VOLUME_CHANGE_RATE == OP_VOLUME_CHANGE_RATE == 28
    header = open(os.path.join(ROOT, "include", "msim_hip.h")).read()
    assert "MSIM_RATE_VOLUME_CHANGE_RATE" in header


def test_kernels_are_built_for_gfx950_only():
    """Every code object in the offload bundle targets gfx950 (the sort's
    header library names other targets in host strings, which is not code)."""
    ensure_built()
    blob = open(msa.LIBRARY, "rb").read()
    targets = set(re.findall(rb"amdgcn-amd-amdhsa-[-\w:+]*", blob))
    assert targets == {b"amdgcn-amd-amdhsa--gfx950"}
