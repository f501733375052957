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


def _interior_features(vel_fn, dim=3):
    """Oracle features of a lattice column with velocity field vel_fn, interior particles only."""
    from zsph_amd import FLUID, cases
    if dim == 3:
        c = cases.dambreak3d(dx=0.02, lx=0.5, wy=0.3, lz=0.5, cw=0.3, ch=0.3)
    else:
        c = cases.dambreak2d(dx=0.01, lx=0.5, lz=0.5, cw=0.3, ch=0.3)
    f = c.host.type == FLUID
    c.host.vel[f] = vel_fn(c.host.pos[f].astype(np.float64)).astype(np.float32)
    c.host.rhop[f, 0] = c.json.rho0  # uniform density, zero pressure: a clean kinematic field
    c.host.rhop[f, 1] = 0.0
    o = oracle.Oracle(c.params)
    o.upload(c.host)
    o.hash(0); o.sort_index(); o.sort_data(0, 1); o.boundary(1)
    F = o.features(1)
    pos = o.get(0, 1)[:, :3]  # posr of the sorted buffer
    typ = o.get(4, 1)
    h = c.params.h
    lo, hi = np.array([0.0, 0.0, 0.0]), np.array([0.3, 0.3 if dim == 3 else 0.0, 0.3])
    inner = (typ == FLUID) & np.all(pos[:, [0, 2]] > lo[[0, 2]] + 2.5 * h, axis=1) & \
        np.all(pos[:, [0, 2]] < hi[[0, 2]] - 2.5 * h, axis=1)
    if dim == 3:
        inner &= (pos[:, 1] > 2.5 * h) & (pos[:, 1] < 0.3 - 2.5 * h)
    assert inner.sum() > 20
    c0 = np.sqrt(c.params.cs2)
    return F[inner], h, c0, F[typ == FLUID]


def test_features_at_rest_in_the_bulk():
    F, h, c0, _ = _interior_features(lambda x: np.zeros_like(x))
    lam = F[:, 1]                                   # tr(L)/3: the lattice sum, ~2 % below 1 at h = 1.3 dx
    assert np.abs(lam - 1.0).max() < 0.03 and np.ptp(lam) < 1e-5
    np.testing.assert_allclose(F[:, 2], lam ** 3, rtol=1e-5)  # cubic lattice: L = lam I
    assert np.abs(F[:, 3]).max() < 1e-6 * lam.max()
    assert np.abs(F[:, 0]).max() < 0.03             # kernel sum ~ 1
    np.testing.assert_array_equal(F[:, 8:18], 0.0)  # no motion
    assert np.all(F[:, 23] == 0.0)                  # no dummy neighbours
    assert np.all(F[:, 20] == F[0, 20])             # same lattice neighbourhood everywhere


# For a linear velocity field v = B x the gradient estimate is exactly K = B L,
# and on a cubic lattice L = lam I (lam = feature 1): closed forms follow.
def test_features_of_a_rigid_rotation():
    om = 3.0  # v = om e_y x (x - x0): vorticity 2 om along y, no strain, no divergence
    x0 = np.array([0.15, 0.15, 0.15])
    F, h, c0, _ = _interior_features(lambda x: np.cross(np.array([0.0, om, 0.0]), x - x0))
    lam = F[:, 1]
    np.testing.assert_allclose(F[:, 10], 2 * om * lam * h / c0, rtol=1e-6)
    np.testing.assert_allclose(F[:, 12], 2 * om * lam * h / c0, rtol=1e-6)
    tiny = 1e-6 * om * h / c0
    assert np.abs(F[:, [8, 9, 11, 13]]).max() < tiny


def test_features_of_a_shear_and_an_expansion():
    g = 2.0  # v = (g z, 0, 0): curl_y = g lam, |dev strain| = g lam / sqrt(2)
    F, h, c0, _ = _interior_features(lambda x: np.stack([g * x[:, 2], 0 * x[:, 0], 0 * x[:, 0]], 1))
    lam = F[:, 1]
    np.testing.assert_allclose(F[:, 10], g * lam * h / c0, rtol=1e-6)
    np.testing.assert_allclose(F[:, 13], g * lam / np.sqrt(2) * h / c0, rtol=1e-6)
    e = 1.5  # v = e x: div = 3 e lam, no curl, no deviatoric strain
    F, h, c0, _ = _interior_features(lambda x: e * x)
    lam = F[:, 1]
    np.testing.assert_allclose(F[:, 8], 3 * e * lam * h / c0, rtol=1e-6)
    assert np.abs(F[:, [9, 10, 11, 12, 13]]).max() < 1e-6 * e * h / c0


def test_closure_with_unit_scale_equals_plain_governing():
    # a network whose output layer is zero gives s = 1: the closure step is the plain step
    from zsph_amd import cases
    from tests.helpers import jitter
    c = jitter(cases.dambreak2d(dx=0.02), seed=3)
    a = oracle.Oracle(c.params)
    b = oracle.Oracle(c.params)
    for o in (a, b):
        o.upload(c.host)
    b.set_closure(closure.init_network(1), emulate_bf16=True)
    for o in (a, b):
        o.hash(0); o.sort_index(); o.sort_data(0, 1); o.boundary(1); o.governing(1)
    np.testing.assert_array_equal(b.get(oracle.A_SCALE, 1), 1.0)
    np.testing.assert_allclose(b.get(9, 1), a.get(9, 1), rtol=1e-13, atol=1e-10)
    # and the parts add up to the plain result
    t = a.get(oracle.A_TERMS, 1)
    acc = a.get(9, 1)
    np.testing.assert_allclose(t[:, 0] + t[:, 1], acc[:, 3], rtol=1e-12, atol=1e-9)
    np.testing.assert_allclose(t[:, 2:5] + t[:, 5:8], acc[:, :3], rtol=1e-12, atol=1e-9)
