"""Pins the CPU restatement of the learned smoother: back-propagated gradients
against central finite differences, its homogeneity and oddness, its
reduction to damped Jacobi when the last layer is zero, tile locality and
masking."""
import numpy as np
import pytest

from tests import learned_reference as lr


def synthetic_level(nx, ny, seed):
    """A 5-point operator with a solid block, random couplings, Dirichlet
    boundary terms on the x ends."""
    rng = np.random.default_rng(seed)
    solid = np.zeros((nx, ny), bool)
    solid[nx // 3:nx // 3 + 5, ny // 4:ny // 2] = True
    m = np.where(solid, 0.0, rng.uniform(0.5, 2.0, (nx, ny)))
    wx = np.zeros((nx, ny))
    wx[1:, :] = 2.0 * m[1:, :] * m[:-1, :] / np.maximum(m[1:, :] + m[:-1, :], 1e-30)
    wy = np.zeros((nx, ny))
    wy[:, 1:] = 2.0 * m[:, 1:] * m[:, :-1] / np.maximum(m[:, 1:] + m[:, :-1], 1e-30)
    bc = np.zeros((nx, ny))
    bc[0, :] = 2.0 * m[0, :]
    bc[-1, :] = 2.0 * m[-1, :]
    L = dict(wx=wx, wy=wy, bc=bc)
    d = bc + wx + wy
    d[:-1, :] += wx[1:, :]
    d[:, :-1] += wy[:, 1:]
    L["dinv"] = np.where(d > 0, 1.0 / np.where(d > 0, d, 1.0), 0.0)
    return L


def solve(L, r):
    nx, ny = r.shape
    n = nx * ny
    A = np.column_stack([lr.apply_operator(L, np.eye(n)[k].reshape(nx, ny)).ravel()
                         for k in range(n)])
    act = np.diag(A) > 0
    e = np.zeros(n)
    e[act] = np.linalg.solve(A[np.ix_(act, act)], r.ravel()[act])
    return e.reshape(nx, ny)


@pytest.fixture(scope="module")
def case():
    L = synthetic_level(20, 70, 3)
    rng = np.random.default_rng(5)
    rp = rng.normal(size=(20, 70)) * (L["dinv"] > 0)
    ep = solve(L, rp)
    W1, W2, W3 = lr.initial_weights(1)
    # Weights three times the initial scale, so that the clip is active on a
    # part of both hidden layers and the gradient check covers its slope.
    W3 = rng.normal(0.0, 0.05, W3.shape)
    return L, rp, ep, (3.0 * W1, 3.0 * W2, W3)


def test_fixture_exercises_the_clip(case):
    L, rp, _, W = case
    clipped, total = np.zeros(2), np.zeros(2)
    for fx0, fy0 in lr.tiles(*rp.shape):
        _, _, c = lr.tile_forward(L, rp, W, fx0, fy0)
        for k, (h, m) in enumerate(((c["h1"], c["m1"]), (c["h2"], c["m2"]))):
            keep = np.broadcast_to(m, h.shape)
            clipped[k] += np.count_nonzero(np.abs(h)[keep] >= 1.0)
            total[k] += np.count_nonzero(keep)
    frac = clipped / total
    assert np.all(frac > 0.01) and np.all(frac < 0.9), frac


def test_gradients_match_finite_differences(case):
    L, rp, ep, W = case
    value, G1, G2, G3 = lr.gradients(L, rp, ep, W)
    assert value == pytest.approx(lr.loss(L, rp, ep, W), rel=1e-12)
    rng = np.random.default_rng(9)
    flat = lr.pack(*W)
    grad = lr.pack(G1, G2, G3)
    picks = np.concatenate([rng.choice(lr.SIZES[0], 8, replace=False),
                            lr.SIZES[0] + rng.choice(lr.SIZES[1], 12, replace=False),
                            lr.SIZES[0] + lr.SIZES[1] + np.arange(0, lr.SIZES[2], 16)])
    h = 1e-6
    for k in picks:
        up, dn = flat.copy(), flat.copy()
        up[k] += h
        dn[k] -= h
        fd = (lr.loss(L, rp, ep, lr.unpack(up)) - lr.loss(L, rp, ep, lr.unpack(dn))) / (2 * h)
        assert abs(fd - grad[k]) <= 1e-6 * max(1.0, abs(grad[k])) + 1e-8, (k, fd, grad[k])


def test_homogeneous_and_odd(case):
    L, rp, _, W = case
    z = lr.smoother(L, rp, W)
    np.testing.assert_allclose(lr.smoother(L, 3.5 * rp, W), 3.5 * z, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(lr.smoother(L, -rp, W), -z, rtol=1e-12, atol=1e-14)


def test_zero_last_layer_is_damped_jacobi(case):
    L, rp, ep, (W1, W2, W3) = case
    W = (W1, W2, np.zeros_like(W3))
    z = lr.smoother(L, rp, W)
    np.testing.assert_array_equal(z, lr.OMEGA * L["dinv"] * rp)
    d = ep - lr.OMEGA * L["dinv"] * rp
    expect = np.sum(d * lr.apply_operator(L, d)) / np.sum(ep * lr.apply_operator(L, ep))
    assert lr.loss(L, rp, ep, W) == pytest.approx(expect, rel=1e-13)


def test_tile_sees_only_its_region(case):
    L, rp, _, W = case
    n0, s0, _ = lr.tile_forward(L, rp, W, 0, 0)
    far = rp.copy()
    far[:, lr.TY + lr.HALO:] = 7.0  # outside the first tile's region
    n1, s1, _ = lr.tile_forward(L, far, W, 0, 0)
    assert s0 == s1
    np.testing.assert_array_equal(n0, n1)
    near = rp.copy()
    near[5, lr.TY + lr.HALO - 1] += 0.5 * s0 / max(L["dinv"][5, lr.TY + lr.HALO - 1], 1e-30)
    n2, _, _ = lr.tile_forward(L, near, W, 0, 0)
    assert not np.array_equal(n0, n2)


def test_correction_vanishes_on_solids(case):
    L, rp, _, W = case
    c = lr.correction(L, rp, W)
    assert np.all(c[L["dinv"] == 0] == 0.0)
    assert np.any(c[L["dinv"] > 0] != 0.0)


def test_bf16_rounding():
    x = np.array([1.0, 1.00390625, 1.005859375, -3.14159, 0.0, 1e-30])
    y = lr.bf16(x)
    assert y[0] == 1.0 and y[1] == 1.0  # a tie rounds to even
    assert y[2] == 1.0078125
    assert abs(y[3] + 3.140625) < 1e-12
    assert y[4] == 0.0
    np.testing.assert_allclose(lr.bf16(y), y)
