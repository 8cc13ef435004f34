"""The reference-side `hip` backend (microsimulator_amd/binding/hip_backend.cpp)
compiled against the reference's own headers and linked with its CPU sources
and libmsim_hip.so (oracle/Makefile, target `binding`). Without a device the
reference's interface behaves as for an unavailable backend: no devices, and
construction throws runtime_error (simulation.cpp:26-36). No device is
needed."""
from __future__ import annotations

import pytest

from oracle import Binding, OracleError


@pytest.fixture(scope="module")
def binding():
    if not Binding.available():
        pytest.skip("oracle/_ref/libmsim_binding.so was not built (needs the reference tree)")
    return Binding()


def test_binding_loads_and_counts_devices(binding):
    import microsimulator_amd as msa
    assert binding.device_count() == msa.device_count()


def test_binding_without_a_device_throws_runtime_error(binding):
    import microsimulator_amd as msa
    if msa.device_count() > 0:
        pytest.skip("a device is present")
    with pytest.raises(OracleError, match="HIP backend device is unavailable") as e:
        binding.create()
    assert e.value.code == 2  # std::runtime_error
