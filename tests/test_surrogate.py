"""The learned seed of the Langmuir-Knudsen interface solve (a 6-64-64-2 BF16
network on the MFMA units, FP32 accumulation).

Pins:
* the oracle's network (oracle/oracle_mlp.hpp) against an independent float64
  implementation (torch autograd) of the same network: forward and gradients
  to the accuracy BF16 rounding allows;
* the device's forward pass and gradient against the oracle's, which rounds
  at the same points (only the FP32 summation order differs);
* the converged interface solve: the network-seeded secant (falling back to
  the secant from the equilibrium rate, then to the reference's Picard solve)
  against the same solve without the network (secant from the equilibrium
  rate): states after N steps within 1e-10 relative L2, on the oracle and on
  the device.  Both differ from the reference's plain solve only on the drops
  where its damped Picard iteration stops at its 30-evaluation cap short of the
  fixed point (counted and bounded below);
* training on the solver's own samples: the loss falls, and the network's
  seed needs fewer map evaluations than the solve without it."""
import ctypes as C

import numpy as np
import pytest

from ice_amd import cases, native

from helpers import compare_states, rel_l2

P = native.MLP_PARAMS
OFF = {"W1": (0, (64, 16)), "b1": (1024, (64,)), "W2": (1088, (64, 64)), "b2": (5184, (64,)),
       "W3": (5248, (16, 64)), "b3": (6272, (16,))}


def random_params(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    p = np.zeros(P, np.float32)
    w1 = np.zeros((64, 16), np.float32)
    w1[:, :6] = rng.uniform(-1, 1, (64, 6)) * np.sqrt(6 / 6) * scale
    p[0:1024] = w1.ravel()
    p[1024:1088] = rng.uniform(-0.1, 0.1, 64)
    p[1088:5184] = (rng.uniform(-1, 1, (64, 64)) * np.sqrt(6 / 64) * scale).ravel()
    p[5184:5248] = rng.uniform(-0.1, 0.1, 64)
    w3 = np.zeros((16, 64), np.float32)
    w3[:2] = rng.uniform(-1, 1, (2, 64)) * np.sqrt(6 / 64) * scale
    p[5248:6272] = w3.ravel()
    p[6272:6274] = rng.uniform(-0.1, 0.1, 2)
    return p


def samples(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1.5, 1.5, (n, 6)).astype(np.float32)
    y = rng.uniform(0.2, 1.0, (n, 2)).astype(np.float32)
    return np.ascontiguousarray(x), np.ascontiguousarray(y)


def split(v):
    return {k: v[o:o + int(np.prod(s))].reshape(s) for k, (o, s) in OFF.items()}


def torch_reference(params, x, y):
    """float64 forward, loss and gradients of the network, no rounding"""
    import torch
    t = {k: torch.tensor(v, dtype=torch.float64, requires_grad=True) for k, v in split(params).items()}
    X = torch.zeros((x.shape[0], 16), dtype=torch.float64)
    X[:, :6] = torch.tensor(x, dtype=torch.float64)
    h1 = torch.relu(X @ t["W1"].T + t["b1"])
    h2 = torch.relu(h1 @ t["W2"].T + t["b2"])
    out = (h2 @ t["W3"].T + t["b3"])[:, :2]
    loss = ((out - torch.tensor(y, dtype=torch.float64)) ** 2).mean()
    loss.backward()
    g = np.zeros(P)
    for k, (o, s) in OFF.items():
        g[o:o + int(np.prod(s))] = t[k].grad.numpy().ravel()
    return out.detach().numpy(), float(loss), g


def lib_grad(lib, pre, params, x, y):
    g = np.zeros(P, np.float32)
    loss = np.zeros(1, np.float32)
    assert getattr(lib, pre + "mlp_grad")(None, x.shape[0], x, y, params, g, loss) == 0
    return g, float(loss[0])


def lib_forward(lib, pre, params, x):
    out = np.zeros((x.shape[0], 2), np.float32)
    assert getattr(lib, pre + "mlp_forward")(None, x.shape[0], x, params, out) == 0
    return out


def block_errors(g, ref):
    return {k: float(np.linalg.norm(split(g)[k] - split(ref)[k]) / max(np.linalg.norm(split(ref)[k]), 1e-30))
            for k in OFF}


def test_oracle_network_against_float64_autograd(oracle):
    params = random_params(1)
    x, y = samples(300, 2)
    out_t, loss_t, g_t = torch_reference(params, x, y)
    out_o = lib_forward(oracle, "orc_", params, x)
    g_o, loss_o = lib_grad(oracle, "orc_", params, x, y)
    assert rel_l2(out_o.T, out_t.T) < 2e-2
    assert abs(loss_o - loss_t) <= 3e-2 * loss_t
    errs = block_errors(g_o, g_t)
    print("oracle vs float64 autograd, relative L2 per tensor:", errs)
    assert max(errs.values()) < 5e-2
    # a transposed or dropped term is far outside that
    bad = g_o.copy()
    bad[1088:5184] = split(g_o)["W2"].T.ravel()
    assert block_errors(bad, g_t)["W2"] > 0.3


def test_seeded_solve_on_oracle(oracle):
    """An untrained network's seed (1, fallbacks included) lands within 1e-10 of
    the solve without the network (2) after 10 steps; both stay within 5e-2 of
    the reference's plain solve (0), whose Picard iteration stops short on a
    few drops."""
    runs = {}
    for mode in (0, 2, 1):
        case = cases.case_evaporation(n=16, steps=10)
        case.fams[0].surrogate = mode
        r = cases.Runner(oracle, case, "orc_")
        if mode == 1:
            assert oracle.orc_surrogate_set_params(r.ctx, 1, random_params(3, 0.3)) == 0
        r.step(10)
        runs[mode] = r
    case = runs[0].case
    err12 = compare_states(runs[1], runs[2], case)
    err02 = compare_states(runs[2], runs[0], case)
    print("network seed vs equilibrium seed", err12, "converged vs the reference's capped Picard", err02)
    assert err12 <= 1e-10
    assert 0 < err02 <= 5e-2


@pytest.mark.gpu
def test_device_network_against_oracle(gpu, oracle):
    params = random_params(4)
    x, y = samples(1000, 5)
    out_d = lib_forward(gpu, "ice_gpu_", params, x)
    out_o = lib_forward(oracle, "orc_", params, x)
    print("forward rel L2", rel_l2(out_d.T, out_o.T))
    assert rel_l2(out_d.T, out_o.T) < 1e-3
    g_d, loss_d = lib_grad(gpu, "ice_gpu_", params, x, y)
    g_o, loss_o = lib_grad(oracle, "orc_", params, x, y)
    errs = block_errors(g_d, g_o)
    print("device vs oracle gradient, relative L2 per tensor:", errs, "loss", loss_d, loss_o)
    assert abs(loss_d - loss_o) <= 1e-4 * loss_o
    assert max(errs.values()) < 2e-3


@pytest.mark.gpu
def test_trained_seed_on_evaporation_case(gpu):
    """Train on the plain solver's samples of the evaporation case; the seeded
    path stays within 1e-10 of the plain one and needs fewer map evaluations."""
    n, steps = 64, 6
    runs = {}
    stats = {}
    report = np.zeros(6, np.float32)
    for mode in (0, 2, 1):
        case = cases.case_evaporation(n=n, steps=steps)
        case.fams[0].surrogate = mode
        r = cases.Runner(gpu, case, "ice_gpu_")
        if mode == 1:
            assert gpu.ice_gpu_surrogate_train(r.ctx, 1, 200, 512, 3e-3, 7, report) == 0, r.error()
            print("training: samples %d, loss %.3e -> %.3e, rate ratio error mean %.3e max %.3e, %.1f ms" %
                  tuple(report))
            assert report[2] < 0.1 * report[1]
        assert gpu.ice_gpu_set_lk_stats(r.ctx, 1) == 0
        r.step(steps)
        s = np.zeros(4, np.int64)
        assert gpu.ice_gpu_lk_stats(r.ctx, 1, s) == 0
        stats[mode] = s
        runs[mode] = r
        print("mode", mode, "solves, evaluations, fallbacks, max evaluations:", s.tolist())
    err12 = compare_states(runs[1], runs[2], runs[0].case)
    err02 = compare_states(runs[2], runs[0], runs[0].case)
    print("network seed vs equilibrium seed", err12, "converged vs the reference's capped Picard", err02)
    assert err12 <= 1e-10 and 0 < err02 <= 5e-2
    # the same drops solve in every mode, up to the few whose state the converged solve moves
    assert stats[0][0] > 0 and abs(stats[1][0] - stats[0][0]) <= 0.01 * stats[0][0] and stats[1][0] == stats[2][0]
    per = {m: stats[m][1] / stats[m][0] for m in stats}
    print("map evaluations per solve: plain %.2f, secant from equilibrium %.2f, network seed %.2f" %
          (per[0], per[2], per[1]))
    assert per[1] < per[2] < per[0]


def retrain_run(gpu, every, explicit, steps=6, n=32, epochs=5):
    """The evaporation case with the network trained at step 0 and retrained every
    `every` steps, by ice_gpu_step itself (explicit False) or by explicit calls of
    ice_gpu_surrogate_retrain after the same steps: (state, network parameters)."""
    case = cases.case_evaporation(n=n, steps=steps)
    case.fams[0].surrogate = 1
    r = cases.Runner(gpu, case, "ice_gpu_")
    rep = np.zeros(6, np.float32)
    assert gpu.ice_gpu_surrogate_train(r.ctx, 1, 50, 256, 3e-3, 11, rep) == 0, r.error()
    p0 = np.zeros(P, np.float32)
    assert gpu.ice_gpu_surrogate_params(r.ctx, 1, p0) == 0
    if explicit:
        for _ in range(steps // every):
            r.step(every)
            assert gpu.ice_gpu_surrogate_retrain(r.ctx, 1, epochs) == 0, r.error()
    else:
        assert gpu.ice_gpu_surrogate_retrain_every(r.ctx, every, epochs) == 0
        r.step(steps)
    p1 = np.zeros(P, np.float32)
    assert gpu.ice_gpu_surrogate_params(r.ctx, 1, p1) == 0
    s = r.state(1, 1)
    r.close()
    return s, p0, p1


@pytest.mark.gpu
@pytest.mark.parametrize("every", [2, 6])
def test_online_retraining(gpu, every):
    """Online retraining every 2 and every 6 steps gives the network and the state
    that the same retraining done by explicit calls after the same steps gives, bit
    for bit; retraining moves the weights."""
    print("case evaporation 32^3 retrain every", every, flush=True)
    s_a, p0_a, p1_a = retrain_run(gpu, every, explicit=False)
    s_b, p0_b, p1_b = retrain_run(gpu, every, explicit=True)
    assert np.array_equal(p0_a, p0_b)
    assert not np.array_equal(p1_a, p0_a)
    assert np.array_equal(p1_a, p1_b)
    assert np.array_equal(s_a, s_b)
