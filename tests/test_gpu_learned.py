"""The learned smoother on the GPU, through the C ABI, against the CPU
restatement in tests/learned_reference.py (itself pinned by finite
differences and invariants in tests/test_learned_reference.py).

Tolerances:
- With a zero last layer the learned cycle is the multilevel cycle with one
  damped-Jacobi post-smoothing sweep on the finest level (two elsewhere):
  against its float64 restatement to 2e-4 of the largest value, the
  multilevel cycle's own tolerance.
- The device network runs in bfloat16 with FP32 sums, so its forward output is
  compared with a restatement that rounds weights and stored activations to
  bfloat16 the same way. What is left is FP32 summation order and rare
  one-ulp bfloat16 flips: 5e-3 of the largest value.
- Gradients come from bfloat16 forward and backward passes and are compared
  with the exact float64 gradients: cosine >= 0.999 and relative norm error
  <= 5e-2, per layer.
"""
from __future__ import annotations

import numpy as np
import pytest

import microsimulator_amd as msa
from microsimulator_amd import learned
from microsimulator_amd.workloads import channel_trap
from tests import learned_reference as lr
from tests import mg_reference as mr
from oracle import Grid as OGrid
from tests.helpers import FIXED, NO_FLUX, c1_inputs, random_case

pytestmark = pytest.mark.gpu


def prepare(hip, nx, ny, seed=13, dy=1.0):
    solid, mob = random_case(seed, nx, ny, solid_fraction=0.1)
    g = msa.Grid(nx, ny, 1, 1.0, dy, 1.0, solid,
                 (FIXED, FIXED, NO_FLUX, NO_FLUX, NO_FLUX, NO_FLUX))
    hip.flow_prepare(g, mob, 0)
    return solid


def level0(hip):
    o = hip.learned_operator()
    return {k: v.astype(np.float64) for k, v in o.items()}


def random_weights(seed, last=0.1):
    W1, W2, _ = lr.initial_weights(seed)
    W3 = np.random.default_rng(seed + 1).normal(0.0, last, (9, lr.C1, 1))
    return lr.pack(W1, W2, W3).astype(np.float32)


def test_learned_requires_weights():
    with msa.HipBackend(0) as fresh:
        prepare(fresh, 64, 48)
        with pytest.raises(msa.InvalidArgument):
            fresh.flow_solve(msa.DepthAveragedFlowParameters(solver=msa.SOLVER_LEARNED))
        with pytest.raises(msa.InvalidArgument):
            fresh.learned_set_weights(np.zeros(10, np.float32))


def plain_part(L0, r):
    """The learned cycle without its network, restated in float64: two
    pre-smoothing sweeps, the V(2,2) cycle below, the coarse correction and
    one damped-Jacobi post-smoothing sweep on the finest level. Returns the
    iterate before post-smoothing, its residual and the finished cycle."""
    ml = mr.Multilevel({k: L0[k] for k in ("wx", "wy", "bc")})
    L = ml.levels[0]
    x = mr.smooth(L, None, r, 2)
    rc = mr.restrict(r - mr.apply(L, x), ml.levels[1]["bc"].shape)
    x = x + mr.CORRECTION * mr.prolong(ml.cycle(rc, 1), r.shape)
    rp = r - mr.apply(L, x)
    return L, x, rp, x + mr.OMEGA * L["dinv"] * rp


@pytest.mark.parametrize("shape", [(256, 128), (301, 77), (20, 13)])
def test_zero_last_layer_is_the_plain_cycle(hip, shape):
    nx, ny = shape
    solid = prepare(hip, nx, ny)
    w = random_weights(3)
    w[-lr.SIZES[2]:] = 0.0
    hip.learned_set_weights(w)
    r = np.random.default_rng(5).standard_normal(nx * ny).astype(np.float32)
    r[solid.reshape(-1) != 0] = 0.0
    z = hip.learned_precondition(r).astype(np.float64)
    _, _, _, ref = plain_part(level0(hip), r.astype(np.float64).reshape(nx, ny))
    err = np.abs(z - ref.ravel()).max() / np.abs(ref).max()
    print(shape, "relative error", err, flush=True)
    assert err <= 2e-4


@pytest.mark.parametrize("shape", [(256, 128), (301, 77)])
def test_learned_cycle_matches_cpu_restatement(hip, shape):
    nx, ny = shape
    solid = prepare(hip, nx, ny, seed=21, dy=1.5)
    w = random_weights(7)
    hip.learned_set_weights(w)
    L0 = level0(hip)
    r = np.random.default_rng(9).standard_normal(nx * ny).astype(np.float32)
    r[solid.reshape(-1) != 0] = 0.0
    z = hip.learned_precondition(r).astype(np.float64).reshape(nx, ny)
    rr = r.astype(np.float64).reshape(nx, ny)
    L, x, rp, _ = plain_part(L0, rr)
    ref = x + lr.smoother(L, rp, lr.unpack(w), bf=True)
    net = lr.correction(L, rp, lr.unpack(w), bf=True)
    err = np.abs(z - ref).max() / np.abs(ref).max()
    print(shape, "net part", np.abs(net).max() / np.abs(ref).max(), "relative error", err,
          flush=True)
    assert np.abs(net).max() > 0.01 * np.abs(ref).max()  # the network contributes
    assert err <= 5e-3


@pytest.fixture(scope="module")
def pairs(hip):
    """Harvested pairs of the 256 x 128 channel-and-trap device."""
    c = c1_inputs()
    g = msa.Grid(c["nx"], c["ny"], 1, 1.0, 1.0, c["dz"], c["obstacles"])
    hip.flow_prepare(g, c["mobility"], 0)
    hip.learned_set_weights(random_weights(11))
    r, e = hip.learned_harvest(8)
    L0 = level0(hip)
    print("harvested", len(r), "pairs", flush=True)
    ops = np.stack([L0["wx"], L0["wy"], L0["bc"], L0["dinv"]])[None].astype(np.float32)
    hip.learned_upload(ops, np.zeros(len(r), np.uint32), r, e)
    return L0, r.astype(np.float64), e.astype(np.float64)


def test_harvested_pairs_are_consistent(hip, pairs):
    """A e' = r' (the error solve is accurate) and the pairs are not trivial."""
    L0, r, e = pairs
    assert len(r) == 8
    for k in range(len(r)):
        res = mr.apply(L0, e[k]) - r[k]
        rel = np.linalg.norm(res) / np.linalg.norm(r[k])
        assert rel <= 1e-4, (k, rel)
        assert np.all(r[k][L0["dinv"] == 0] == 0.0)


def test_forward_matches_cpu_restatement(hip, pairs):
    L0, r, _ = pairs
    w = random_weights(11)
    hip.learned_set_weights(w)
    for k in (0, 5):
        z = hip.learned_apply(k).astype(np.float64).reshape(r[k].shape)
        ref = lr.smoother(L0, r[k], lr.unpack(w), bf=True)
        err = np.abs(z - ref).max() / np.abs(ref).max()
        print("pair", k, "forward relative error", err, flush=True)
        assert err <= 5e-3


def test_gradients_match_cpu_restatement(hip, pairs):
    L0, r, e = pairs
    w = random_weights(11, last=0.05)
    hip.learned_set_weights(w)
    for k in (1, 6):
        loss, grad = hip.learned_gradients(k)
        ref_loss, G1, G2, G3 = lr.gradients(L0, r[k], e[k], lr.unpack(w))
        print("pair", k, "loss", loss, "oracle", ref_loss, flush=True)
        assert loss == pytest.approx(ref_loss, rel=2e-2)
        ref = lr.pack(G1, G2, G3)
        a = 0
        for name, size in zip(("W1", "W2", "W3"), lr.SIZES):
            g, gr = grad[a:a + size], ref[a:a + size]
            a += size
            cos = g @ gr / (np.linalg.norm(g) * np.linalg.norm(gr))
            rel = np.linalg.norm(g - gr) / np.linalg.norm(gr)
            print(f"  {name}: cosine {cos:.6f} relative error {rel:.2e}", flush=True)
            assert cos >= 0.999
            assert rel <= 5e-2


def test_training_lowers_the_loss(hip, pairs):
    L0, r, e = pairs
    w0 = learned.initial_weights(0)
    hip.learned_set_weights(w0)

    def loss_over_all_pairs():
        return float(np.mean([hip.learned_gradients(k)[0] for k in range(len(r))]))
    before = loss_over_all_pairs()
    # With W3 = 0 the smoother is damped Jacobi, whose loss has a closed form.
    first = []
    for k in range(len(r)):
        d = e[k] - lr.OMEGA * L0["dinv"] * r[k]
        first.append(np.sum(d * mr.apply(L0, d)) / np.sum(e[k] * mr.apply(L0, e[k])))
    assert before == pytest.approx(float(np.mean(first)), rel=1e-3)
    losses, seconds = hip.learned_train(60, batch=4, learning_rate=3e-3, seed=1)
    after = loss_over_all_pairs()
    print("loss over the pairs", before, "->", after, "batch losses", losses[:3], "...",
          losses[-3:], "seconds", seconds, flush=True)
    assert after < 0.995 * before
    w1 = hip.learned_weights()
    assert np.any(w1[-lr.SIZES[2]:] != 0.0)


def test_fallback_to_jacobi_at_the_cap(hip, oracle):
    c = c1_inputs()
    g = msa.Grid(c["nx"], c["ny"], 1, 1.0, 1.0, c["dz"], c["obstacles"])
    hip.flow_prepare(g, c["mobility"], 0)
    hip.learned_set_weights(learned.initial_weights(0))
    p = msa.DepthAveragedFlowParameters(mean_inlet_speed=100.0, relative_tolerance=1e-8,
                                        solver=msa.SOLVER_LEARNED, fallback_iterations=3,
                                        check_interval=1)
    rep = hip.flow_solve(p)
    print("fallback", rep.fell_back, "iterations", rep.iterations, "true",
          rep.true_relative_residual, flush=True)
    assert rep.fell_back
    assert rep.iterations > 3
    # The fallback continues from the flexible-CG iterate and stops on the
    # FP64 true residual, the judge of every form.
    assert rep.true_relative_residual <= 1e-8
    assert rep.relative_residual == pytest.approx(rep.true_relative_residual, rel=1e-12)
    # The field is the solution's: against the oracle's FP64 solve.
    faces = hip.flow_download()[1].concatenated()
    ref_faces, _, ref_rep = oracle.solve_flow(
        OGrid(c["nx"], c["ny"], 1, 1.0, 1.0, c["dz"], c["obstacles"]), c["mobility"], 0, 100.0,
        1e-10, 100000)
    assert np.abs(faces - ref_faces).max() <= 1e-5 * ref_rep.max_speed
    p.fallback_iterations = 200
    rep = hip.flow_solve(p)
    assert not rep.fell_back
    assert rep.true_relative_residual <= 1e-8
