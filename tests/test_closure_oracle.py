"""Closure network oracle (CPU): BF16 rounding, the loss definition, the
range fallback, and the FP64 backward pass pinned by central finite
differences over parameters of every layer."""
import numpy as np
import pytest

import oracle
from zsph_amd import closure


def random_problem(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.normal(0.0, 1.0, (n, closure.NF)).astype(np.float32)
    rec = rng.normal(0.0, 1.0, (n, closure.NT)).astype(np.float32)
    return x, rec


def random_net(seed):
    net = closure.init_network(seed, out_scale=0.8, zero_output=False)
    rng = np.random.default_rng(seed + 100)
    P = net.params()
    P[:] += rng.normal(0.0, 0.05, P.shape).astype(np.float32)  # non-zero biases too
    return net


def test_bf16_rounding_is_nearest_even():
    assert oracle.bf16_round(1.0 + 2.0 ** -9) == 1.0              # quarter ulp: down
    assert oracle.bf16_round(1.0 + 3 * 2.0 ** -9) == 1.0 + 2.0 ** -7  # three quarters: up
    assert oracle.bf16_round(1.0 + 2.0 ** -8) == 1.0              # tie: to the even 1.0
    assert oracle.bf16_round(1.0 + 3 * 2.0 ** -8) == 1.0 + 2 * 2.0 ** -7  # tie: to the even neighbour
    assert oracle.bf16_round(-3.0) == -3.0


def test_zero_output_layer_is_the_plain_residual():
    x, rec = random_problem(50, 1)
    net = closure.init_network(3)  # zero output layer: s = 1 exactly
    loss, _ = oracle.mlp_loss_grad(net, x, rec, w_rho=2.0, w_acc=0.5, grad=False)
    r = rec.astype(np.float64)
    res_rho = r[:, 0] + r[:, 1] - r[:, 2]
    res_a = r[:, 3:6] + r[:, 6:9] - r[:, 9:12]
    ref = np.mean(2.0 * res_rho ** 2 + 0.5 * (res_a ** 2).sum(1))
    assert loss == pytest.approx(ref, rel=1e-12)
    s = oracle.mlp_infer(net, x)
    assert np.array_equal(s, np.ones_like(s))


def test_range_fallback_keeps_plain_terms():
    x, _ = random_problem(20, 2)
    net = random_net(4)
    net.lo[:] = [-2.0] * closure.NF
    net.hi[:] = [2.0] * closure.NF
    x[:, 5] = 0.0
    x[3, 5] = 2.5   # outside the normalised range
    x[7, 5] = -2.01
    s = oracle.mlp_infer(net, x)
    assert np.array_equal(s[[3, 7]], np.ones((2, 2)))
    inside = np.all(np.abs(x) <= 2.0, axis=1)
    assert np.all(s[inside] != 1.0)


@pytest.mark.parametrize("seed", [0, 1])
def test_backward_matches_finite_differences(seed):
    x, rec = random_problem(7, 10 + seed)
    net = random_net(seed)
    loss, g = oracle.mlp_loss_grad(net, x, rec, w_rho=1.3, w_acc=0.7)
    P = net.params()
    rng = np.random.default_rng(seed)
    offs = np.cumsum([0, 24 * 64, 64, 64 * 64, 64, 64 * 64, 64, 64 * 2, 2])
    # every block of parameters (w0 b0 w1 b1 w2 b2 w3 b3), 6 entries each
    idx = np.concatenate([rng.integers(offs[k], offs[k + 1], 6) for k in range(8)])
    for i in idx:
        keep = P[i]
        # FP32 parameters: step on the FP32 grid, difference taken in FP64
        h = np.float32(max(1e-3 * abs(keep), 1e-3))
        P[i] = keep + h
        up = float(P[i])
        lp, _ = oracle.mlp_loss_grad(net, x, rec, w_rho=1.3, w_acc=0.7, grad=False)
        P[i] = keep - h
        down = float(P[i])
        lm, _ = oracle.mlp_loss_grad(net, x, rec, w_rho=1.3, w_acc=0.7, grad=False)
        P[i] = keep
        fd = (lp - lm) / (up - down)  # the FP32-rounded step actually taken
        assert fd == pytest.approx(g[i], rel=2e-5, abs=2e-6 * (1 + abs(loss))), f"parameter {i}"


def test_bf16_emulation_stays_close_to_fp64():
    x, rec = random_problem(200, 5)
    net = random_net(2)
    l64, g64 = oracle.mlp_loss_grad(net, x, rec)
    l16, g16 = oracle.mlp_loss_grad(net, x, rec, emulate_bf16=True)
    assert l16 == pytest.approx(l64, rel=2e-2)
    assert np.linalg.norm(g16 - g64) < 0.05 * np.linalg.norm(g64)
