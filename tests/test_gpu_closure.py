"""GPU: the MFMA closure network against its CPU oracle.

Forward (scale factors, range fallback), the training loss and its gradient
from the device's backward kernels against the FP64 oracle with the same
BF16 roundings (they differ only by FP32 accumulation order), determinism,
and training on the device: Adam steps fit a known scale-factor field.
"""
import sys

import numpy as np
import pytest

import oracle
from zsph_amd import closure
from tests.test_closure_oracle import random_net, random_problem

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("n", [1, 31, 1000])
def test_infer_matches_oracle(n):
    x, _ = random_problem(n, 7)
    net = random_net(1)
    net.lo[:] = [-2.5] * closure.NF
    net.hi[:] = [2.5] * closure.NF
    say(f"[mlp infer] n={n}")
    m = closure.Mlp(net)
    s = m.infer(x)
    ref = oracle.mlp_infer(net, x, emulate_bf16=True)
    outside = np.any(np.abs(x) > 2.5, axis=1)
    assert np.array_equal(s[outside], np.ones((int(outside.sum()), 2), np.float32))
    np.testing.assert_allclose(s, ref, rtol=0, atol=2e-3)
    m.close()


@pytest.mark.parametrize("n", [33, 4096, 20000])
def test_loss_and_gradient_match_oracle(n):
    x, rec = random_problem(n, 11)
    net = random_net(3)
    say(f"[mlp grad] n={n}")
    m = closure.Mlp(net)
    m.data(x, rec, w_rho=1.5, w_acc=0.5)
    loss, g = m.eval()
    lref, gref = oracle.mlp_loss_grad(net, x, rec, w_rho=1.5, w_acc=0.5, emulate_bf16=True)
    say(f"  loss gpu {loss:.6g} oracle {lref:.6g}; |g| {np.linalg.norm(gref):.4g}, "
        f"|dg| {np.linalg.norm(g - gref):.3g}")
    assert loss == pytest.approx(lref, rel=1e-4)
    scale = np.abs(gref).max()
    np.testing.assert_allclose(g, gref, rtol=0, atol=2e-3 * scale)
    assert np.linalg.norm(g - gref) < 2e-3 * np.linalg.norm(gref)
    # deterministic: a second evaluation is bitwise identical
    loss2, g2 = m.eval()
    assert loss2 == loss and np.array_equal(g, g2)
    # sub-ranges of the training set
    l_half, _ = m.eval(0, n // 2, grad=False)
    lh_ref, _ = oracle.mlp_loss_grad(net, x[:n // 2], rec[:n // 2], w_rho=1.5, w_acc=0.5, emulate_bf16=True,
                                      grad=False)
    assert l_half == pytest.approx(lh_ref, rel=1e-4)
    m.close()


def test_adam_fits_a_known_scale_field():
    # records whose optimal scale factors are s_d = 1 + 0.5 tanh(x0), s_v = 1 - 0.4 tanh(x1):
    # diff = 1, t_rho = cont + s_d*; visc = e_z, t_a = pa + s_v* e_z
    rng = np.random.default_rng(5)
    n = 16384
    x = rng.normal(0.0, 1.0, (n, closure.NF)).astype(np.float32)
    sd = 1.0 + 0.5 * np.tanh(x[:, 0])
    sv = 1.0 - 0.4 * np.tanh(x[:, 1])
    rec = np.zeros((n, closure.NT), np.float32)
    rec[:, 0] = rng.normal(size=n)
    rec[:, 1] = 1.0
    rec[:, 2] = rec[:, 0] + sd
    rec[:, 3:6] = rng.normal(size=(n, 3))
    rec[:, 8] = 1.0
    rec[:, 9:12] = rec[:, 3:6]
    rec[:, 11] += sv
    net = closure.set_normalisation(closure.init_network(0), x)
    m = closure.Mlp(net)
    m.data(x, rec)
    l0, _ = m.eval(grad=False)
    batch = 4096
    for it in range(600):
        m.adam((it % (n // batch)) * batch, batch, lr=2e-3)
    l1, _ = m.eval(grad=False)
    say(f"[mlp fit] loss {l0:.4g} -> {l1:.4g}")
    assert l1 < 0.02 * l0
    s = m.infer(x[:2000])
    assert np.abs(s[:, 0] - sd[:2000]).mean() < 0.03
    assert np.abs(s[:, 1] - sv[:2000]).mean() < 0.03
    # the trained parameters come back to the host and reproduce the device's output
    trained = m.get()
    ref = oracle.mlp_infer(trained, x[:2000], emulate_bf16=True)
    np.testing.assert_allclose(s, ref, rtol=0, atol=2e-3)
    m.close()
