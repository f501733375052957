"""Shared test helpers: seeded states and tolerance checks.

Tolerances (the parity bar for floating-point results):
* state after one step (rho): |d| <= 1e-5 * |rho|;
* pressure: |d| <= 1e-5 * max(|p|, rho0 |g| H) (H = column height);
* sums over neighbours (acceleration, drho/dt, grad rho^L, div v):
  |d| <= 1e-5 * S_i, where S_i is the oracle's per-particle sum of the
  absolute values of the summed terms (relative error of a sum with
  cancellation is measured against the size of its terms).
Integer / index results (hash, sort order, cell table, neighbour sets,
gathers) must be bit-exact.
"""
import numpy as np

import oracle
from zsph_amd import FLUID

REL = 1e-5


def jitter(case, seed=12345, amp=0.1, vmax=1.0, drho=0.01):
    """Perturb fluid particles: +-amp*dx positions, U(-vmax,vmax) velocity,
    rho0 (1 + U(-drho, drho)) density with its EOS pressure."""
    rng = np.random.default_rng(seed)
    h = case.host
    f = h.type == FLUID
    n = int(f.sum())
    dx = case.json.dx
    h.pos[f] += rng.uniform(-amp, amp, (n, 3)).astype(np.float32) * np.float32(dx)
    if case.cpu.dim == 2:
        h.pos[:, 1] = 0.0
    h.vel[f] = rng.uniform(-vmax, vmax, (n, 3)).astype(np.float32)
    if case.cpu.dim == 2:
        h.vel[:, 1] = 0.0
    rho = case.json.rho0 * (1.0 + rng.uniform(-drho, drho, n))
    h.rhop[f, 0] = rho.astype(np.float32)
    if n <= 200000:
        h.rhop[f, 1] = np.array([oracle.eos_p(case.params, float(r)) for r in h.rhop[f, 0]], np.float32)
    else:  # the same EOS vectorised (any pressure is a valid initial state; both sides start from it)
        p = case.params
        e = h.rhop[f, 0].astype(np.float64) / p.rho0 - 1.0
        h.rhop[f, 1] = (p.cs2 * p.rho0 * e if p.gamma == 1.0
                        else p.rho0 * p.cs2 / p.gamma * np.expm1(p.gamma * np.log1p(e))).astype(np.float32)
    return case


def assert_close_scaled(got, ref, scale, what, rel=REL):
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    scale = np.asarray(scale, np.float64)
    err = np.abs(got - ref)
    bad = err > rel * scale + 1e-30
    if bad.any():
        k = int(np.argmax(err / (scale + 1e-30)))
        raise AssertionError(f"{what}: {int(bad.sum())} of {bad.size} values out of tolerance; worst at flat "
                             f"index {k}: got {got.ravel()[k]!r} ref {ref.ravel()[k]!r} scale {scale.ravel()[k]!r}")
    return float((err / (scale + 1e-30)).max()) if err.size else 0.0


def p_scale(case):
    H = case.json.ptc_num and (case.cpu.zmax - case.cpu.zmin)
    return case.json.rho0 * abs(case.json.g) * max(H, 1e-3)
