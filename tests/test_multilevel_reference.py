"""Pins the CPU restatement of the multilevel preconditioner (tests/mg_reference.py)
by what the mathematics fixes: the aggregated operator is the Galerkin
product P^T A P, the cycle is a symmetric positive-definite operator, and on a
grid that needs no coarsening it is the exact inverse."""
from __future__ import annotations

import numpy as np

from oracle import Grid
from tests import mg_reference as mr
from tests.helpers import random_case


def system(oracle, nx, ny, seed=3):
    solid, mob = random_case(seed, nx, ny)
    a = oracle.assemble(Grid(nx, ny, 1, 1.0, 1.0, 1.0, solid), mob, 0)
    m = a["conductance"].reshape(nx, ny).astype(np.float32)

    def harm(p, q):
        s = p + q
        return np.where(s > 0, np.float32(2) * p * q / np.where(s > 0, s, 1), 0).astype(np.float32)
    cxl = np.zeros((nx, ny), np.float32)
    cxl[1:] = harm(m[1:], m[:-1])
    cyl = np.zeros((nx, ny), np.float32)
    cyl[:, 1:] = harm(m[:, 1:], m[:, :-1])
    return mr.level0(cxl, cyl, a["bc"], nx, ny, 1.0, 1.0), a


def test_aggregated_operator_is_galerkin(oracle):
    L, _ = system(oracle, 9, 7)
    C = mr.coarsen(L)
    A = mr.dense(L)
    cx, cy = C["bc"].shape
    P = np.zeros((9 * 7, cx * cy))
    for x in range(9):
        for y in range(7):
            P[x * 7 + y, (x // 2) * cy + y // 2] = 1.0
    assert np.allclose(mr.dense(C), P.T @ A @ P, rtol=1e-12, atol=1e-12)


def test_cycle_is_symmetric_positive_definite(oracle):
    L, _ = system(oracle, 64, 36)
    ml = mr.Multilevel(L)
    assert len(ml.levels) >= 3
    n = 64 * 36
    M = np.column_stack([ml.cycle(np.eye(n)[k].reshape(64, 36)).ravel() for k in range(n)])
    active = mr.diagonal(L).ravel() > 0
    Ma = M[np.ix_(active, active)]
    assert np.allclose(Ma, Ma.T, rtol=1e-10, atol=1e-12)
    assert np.linalg.eigvalsh(0.5 * (Ma + Ma.T)).min() > 0


def test_single_level_cycle_is_exact(oracle):
    L, a = system(oracle, 12, 10)
    ml = mr.Multilevel(L)
    assert len(ml.levels) == 1
    b = a["rhs"].reshape(12, 10).astype(np.float64)
    x = ml.cycle(b)
    assert np.allclose(mr.apply(L, x), b, rtol=1e-10, atol=1e-12)
