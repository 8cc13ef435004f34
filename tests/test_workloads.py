"""The seeded workload generators."""
from __future__ import annotations

import numpy as np

from microsimulator_amd.workloads import brinkman_mobility, config, pillar_array


def test_channel_trap_layout():
    w = config("c1")
    assert (w.nx, w.ny) == (256, 128)
    assert w.fluid == 256 * 40 + 80 * 60
    assert np.all(w.mobility[w.obstacles != 0] == 0)
    assert np.all(w.mobility[w.obstacles == 0] > 0)
    assert w.phi.max() < 0.6 and np.all(w.phi[:, :40] == 0)


def test_generators_are_seeded():
    a, b = config("c1"), config("c1")
    assert np.array_equal(a.mobility, b.mobility)
    c = config("c1", seed=43)
    assert not np.array_equal(a.mobility, c.mobility)


def test_pillar_array_counts_and_clear_ends():
    w = pillar_array(512, 256, seed=3, pillars=60, radius=(8, 15), blobs=10)
    assert 50 <= w.meta["pillars"] <= 60
    assert np.all(w.obstacles[:10, :] == 0) and np.all(w.obstacles[-10:, :] == 0)
    assert 0.0 < w.phi.max() <= 0.7


def test_brinkman_mobility_closure():
    """m = k/(1 + k*C*phi^2/(1-phi)^3) with k = H^2/12 (flow.py:91-96, 361-384)."""
    m = brinkman_mobility(np.array([[10.0, 10.0]]), np.array([[0.0, 0.5]]),
                          np.zeros((1, 2), bool), 10.0)
    k = 100.0 / 12.0
    assert np.isclose(m[0, 0], k, rtol=1e-6)
    assert np.isclose(m[0, 1], k / (1 + k * 100.0 * 0.25 / 0.125), rtol=1e-6)
