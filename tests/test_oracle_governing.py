"""Pins the oracle's continuity + momentum pass (oracle/zsph_oracle.hpp,
``governing``; the reference's delta_ns.cu:240-264) term by term, against
closed forms that do not restate the oracle's per-particle loop:

* two isolated fluid particles: every term of dRho/dt and of the acceleration
  (continuity, Antuono density diffusion, pressure, bulk viscosity from D,
  artificial viscosity, gravity), written out from the textbook delta-SPH
  formulas (Antuono et al. 2010; Monaghan 1992 for the alpha term) with the
  Wendland C2 gradient F = (1/r) dW/dr = -5 alpha_D (1 - q/2)^3 / h^2;
* a linear velocity field v = B x on every particle, random densities: the
  continuity sum equals -rho_i tr(B L_i) with L_i = -sum_j F V_j x_ij x_ij^T
  from a brute-force neighbour search, and D_i = tr(B L_i) over fluid
  neighbours (the SPH divergence of a linear field);
* random velocities: the alpha part of the viscous acceleration dissipates,
  sum_i m v_i . a_i = h c rho0 m alpha sum_pairs m F (v_ij . x_ij)^2 /
  ((r^2 + eta^2) rho_i rho_j) < 0 (F < 0), evaluated pair by pair.

A sign flip, a swapped density (V_i for V_j, rho_j for rho_i), a dropped
factor or a wrong velocity difference in either the continuity or the viscous
term fails these checks.  CPU only.
"""
import math

import numpy as np
import pytest
from scipy.spatial import cKDTree

import oracle
from oracle import A_ACC_SCALE, A_TERMS
from zsph_amd import A_ACC, A_GD, A_POSR, A_TYPE, A_VELP, FLUID, WALL, cases
from tests.helpers import jitter


def staged(case, gd_hook=None):
    o = oracle.Oracle(case.params)
    o.upload(case.host)
    o.hash(0)
    o.sort_index()
    o.sort_data(0, 1)
    o.wall(1)
    o.ldg(1)
    if gd_hook:
        o.set(A_GD, gd_hook(o.get(A_GD, 1)), 1)
    o.governing(1)
    return o


def wendland_F(cpu, r):
    """(1/r) dW/dr of W = alpha_D (1 - q/2)^4 (2q + 1).  alpha_D is the case's stored
    float32 constant (the reference keeps it in its float parameter block), which is
    21 / (16 pi h^3) in 3-D to float32 rounding."""
    h = float(cpu.h)
    ad = float(cpu.adh)
    assert ad == pytest.approx(21.0 / (16.0 * math.pi * h ** 3), rel=1e-7)
    q = r / h
    return np.where(q <= 2.0, -5.0 * ad * (1.0 - q / 2.0) ** 3 / h ** 2, 0.0)


# ---------------------------------------------------------------- two particles
def test_two_particle_closed_form():
    c = cases.dambreak3d(dx=0.05, lx=0.6, wy=0.3, lz=0.4, cw=0.25, ch=0.2)
    cpu = c.cpu
    h = c.host
    keep = np.zeros(c.n, bool)
    keep[:2] = True
    h.type[~keep] = WALL
    h.pos[~keep, 2] = np.float32(cpu.zmax)  # every other particle far above the pair
    h.vel[~keep] = 0.0
    # the pair: r = 0.9 h along a general direction, distinct densities, pressures, velocities
    hh = float(cpu.h)
    e = np.array([0.48, -0.6, 0.64])
    e /= np.linalg.norm(e)
    xi = np.array([0.3, 0.15, 0.12], np.float32)
    xj = (xi.astype(np.float64) - 0.9 * hh * e).astype(np.float32)
    h.type[:2] = FLUID
    h.pos[0], h.pos[1] = xi, xj
    h.vel[0] = (-0.5, 0.4, -0.3)   # v_ij . e < 0: the pair approaches
    h.vel[1] = (0.3, -0.2, 0.4)
    h.rhop[0] = (1003.0, 2100.0)
    h.rhop[1] = (998.5, 900.0)
    o = staged(c)
    typ = o.get(A_TYPE, 1)
    fl = np.where(typ == FLUID)[0]
    assert len(fl) == 2
    posr, velp, acc, gd = o.get(A_POSR, 1), o.get(A_VELP, 1), o.get(A_ACC, 1), o.get(A_GD, 1)
    i, j = (fl[0], fl[1]) if np.allclose(posr[fl[0], :3], xi) else (fl[1], fl[0])

    m, rho0, cs = float(cpu.m), float(cpu.rho0), float(cpu.cs)
    # the parameter block's float32 products (lib/init.cu:221-251), pinned to their definitions
    prm = c.params
    eta2, dhc, hcrm = float(prm.eta2), float(prm.delta_h_cs), float(prm.h_cs_rho_m)
    assert eta2 == pytest.approx(float(cpu.eta) ** 2, rel=3e-7)
    assert dhc == pytest.approx(float(cpu.delta) * hh * cs, rel=3e-7)        # delta h c0
    assert hcrm == pytest.approx(hh * cs * rho0 * m, rel=3e-7)              # h c0 rho0 m
    alpha, g = float(cpu.alpha), float(cpu.g)

    x = posr[i, :3] - posr[j, :3]               # x_ij
    r2 = float(x @ x)
    F = float(wendland_F(cpu, math.sqrt(r2)))
    assert F < 0.0
    ri, rj = posr[i, 3], posr[j, 3]
    pi_, pj = velp[i, 3], velp[j, 3]
    vij = velp[i, :3] - velp[j, :3]
    vx = float(vij @ x)                          # v_ij . x_ij (symmetric in i, j)
    Vi, Vj = m / ri, m / rj

    # gradient pass: D = div v over fluid neighbours; L has rank 1 (singular), so G is not renormalised
    Di, Dj = -F * vx * Vj, -F * vx * Vi
    Gi = F * (rj - ri) * Vj * x
    Gj = F * (ri - rj) * Vi * (-x)
    np.testing.assert_allclose(gd[i], [*Gi, Di], rtol=1e-12, atol=0)
    np.testing.assert_allclose(gd[j], [*Gj, Dj], rtol=1e-12, atol=0)

    def particle(a, b, xab, Va_, Vb_, ra, rb, pa, pb, Ga, Gb, Da, Db):
        cont = ra * vx * F * Vb_                                             # continuity
        psi = 2.0 * (ra - rb) * r2 / (r2 + eta2) - float((Ga + Gb) @ xab)
        diff = dhc * F * psi * Vb_                                           # Antuono diffusion
        a_p = -(pa + pb) * m * F / (ra * rb) * xab                           # pressure
        a_v = hcrm * F * (Da + Db + alpha * vx / (r2 + eta2)) / (ra * rb) * xab  # bulk + artificial viscosity
        a_tot = a_p + a_v + np.array([0.0, 0.0, g])
        return np.array([*a_tot, cont + diff]), cont, diff, a_v

    want_i, cont_i, diff_i, av_i = particle(i, j, x, Vi, Vj, ri, rj, pi_, pj, Gi, Gj, Di, Dj)
    want_j, cont_j, diff_j, av_j = particle(j, i, -x, Vj, Vi, rj, ri, pj, pi_, Gj, Gi, Dj, Di)
    np.testing.assert_allclose(acc[i], want_i, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(acc[j], want_j, rtol=1e-12, atol=1e-12)
    terms = o.get(A_TERMS, 1)
    np.testing.assert_allclose(terms[i, [0, 1]], [cont_i, diff_i], rtol=1e-12)
    np.testing.assert_allclose(terms[i, 5:8], av_i, rtol=1e-12)
    # physical signs: the pair approaches (v_ij . x_ij < 0), so both densities grow and the
    # viscous force opposes the approach (it does negative work on the relative motion)
    assert vx < 0.0 and cont_i > 0.0 and cont_j > 0.0
    assert float((av_i - av_j) @ vij) < 0.0


# ---------------------------------------------------------------- linear velocity field
def neighbour_pairs(posr, cpu, rc2):
    """Unordered neighbour pairs (a, b) with x_ab, r^2 and F(r) (FP64, kd-tree)."""
    x = posr[:, :3]
    pairs = cKDTree(x).query_pairs(math.sqrt(rc2) * (1 + 1e-9), output_type="ndarray")
    a, b = pairs[:, 0], pairs[:, 1]
    d = x[a] - x[b]
    r2 = np.einsum("ij,ij->i", d, d)
    ok = r2 <= rc2
    a, b, d, r2 = a[ok], b[ok], d[ok], r2[ok]
    return a, b, d, r2, wendland_F(cpu, np.sqrt(r2))


def pair_sum(n, a, b, wa, wb):
    """out_i = sum over pairs of wa (for i = a) and wb (for i = b); w of shape (npairs, ...)."""
    out = np.zeros((n,) + wa.shape[1:])
    np.add.at(out, a, wa)
    np.add.at(out, b, wb)
    return out


def test_continuity_of_linear_velocity_field():
    c = jitter(cases.dambreak3d(dx=0.04, lx=0.6, wy=0.32, lz=0.4, cw=0.3, ch=0.3), seed=21, amp=0.2, drho=0.02)
    h = c.host
    B = np.array([[0.8, -0.3, 0.5], [0.2, -1.1, 0.4], [-0.6, 0.1, 0.9]])
    h.vel[:] = (h.pos.astype(np.float64) @ B.T).astype(np.float32)  # fluid and dummies alike
    o = staged(c)
    posr, velp, typ = o.get(A_POSR, 1), o.get(A_VELP, 1), o.get(A_TYPE, 1)
    n = len(typ)
    f = typ == FLUID
    a, b, d, r2, F = neighbour_pairs(posr, c.cpu, float(c.params.rc2))
    V = float(c.cpu.m) / posr[:, 3]
    outer = d[:, :, None] * d[:, None, :]
    # L_i = -sum_j F V_j x_ij x_ij^T over all neighbours, and over fluid neighbours only
    L = pair_sum(n, a, b, (-F * V[b])[:, None, None] * outer, (-F * V[a])[:, None, None] * outer)
    fa, fb = (typ[a] == FLUID).astype(float), (typ[b] == FLUID).astype(float)
    Lf = pair_sum(n, a, b, (-F * V[b] * fb)[:, None, None] * outer, (-F * V[a] * fa)[:, None, None] * outer)
    # the stored velocities are the float32 roundings of B x: v_ij = B x_ij + e_ij with
    # |e_ij| <= |e_i| + |e_j|, which bounds each pair term's departure from the closed form
    e = np.linalg.norm(velp[:, :3] - posr[:, :3] @ B.T, axis=1)
    r = np.sqrt(r2)
    slack_w = np.abs(F) * r * (e[a] + e[b])
    slack = pair_sum(n, a, b, slack_w * V[b], slack_w * V[a])
    slack_f = pair_sum(n, a, b, slack_w * V[b] * fb, slack_w * V[a] * fa)
    terms = o.get(A_TERMS, 1)
    gd = o.get(A_GD, 1)
    scale = o.get(A_ACC_SCALE, 1)[:, 1]
    # continuity: rho_i sum_j (v_ij . x_ij) F V_j = rho_i tr(B sum F V x x^T) = -rho_i tr(B L_i)
    want = -posr[:, 3] * np.einsum("ab,iba->i", B, L)
    err = np.abs(terms[f, 0] - want[f])
    tol = posr[f, 3] * slack[f] + 1e-10 * scale[f]
    assert np.all(err <= tol), float((err / tol).max())
    assert f.sum() > 500 and np.abs(want[f]).max() > 100.0
    assert np.median(tol / np.abs(want[f])) < 1e-5  # the check is far tighter than the terms
    # D_i = -sum_fluid F V_j v_ij . x_ij = tr(B L_i^fluid): the SPH divergence of a linear field
    wantD = np.einsum("ab,iba->i", B, Lf)
    errD = np.abs(gd[f, 3] - wantD[f])
    assert np.all(errD <= slack_f[f] + 1e-10 * np.abs(wantD[f]).max())


# ---------------------------------------------------------------- artificial viscosity
def test_artificial_viscosity_dissipates():
    c = jitter(cases.dambreak3d(dx=0.04, lx=0.6, wy=0.32, lz=0.4, cw=0.3, ch=0.3, layers=0), seed=5, amp=0.2,
               drho=0.02)

    def no_bulk(gd):
        gd[:, 3] = 0.0  # D = 0: the viscous part is the alpha term alone
        return gd

    o = staged(c, gd_hook=no_bulk)
    posr, velp, typ = o.get(A_POSR, 1), o.get(A_VELP, 1), o.get(A_TYPE, 1)
    assert np.all(typ == FLUID)
    av = o.get(A_TERMS, 1)[:, 5:8]
    cpu = c.cpu
    m = float(cpu.m)
    power = float(m * np.einsum("ij,ij->", velp[:, :3], av))
    # the same power pair by pair
    x = posr[:, :3]
    rc2 = float(c.params.rc2)
    pairs = cKDTree(x).query_pairs(math.sqrt(rc2) * (1 + 1e-9), output_type="ndarray")
    a, b = pairs[:, 0], pairs[:, 1]
    d = x[a] - x[b]
    r2 = np.einsum("ij,ij->i", d, d)
    ok = r2 <= rc2
    a, b, d, r2 = a[ok], b[ok], d[ok], r2[ok]
    F = wendland_F(cpu, np.sqrt(r2))
    vx = np.einsum("ij,ij->i", velp[a, :3] - velp[b, :3], d)
    hcrm = float(c.params.h_cs_rho_m)
    eta2 = float(c.params.eta2)
    pair_power = float(np.sum(m * hcrm * float(cpu.alpha) * F * vx * vx / ((r2 + eta2) * posr[a, 3] * posr[b, 3])))
    assert pair_power < 0.0
    assert power == pytest.approx(pair_power, rel=1e-9)
    # and it is the alpha term that carries it: without it (alpha = 0) the viscous part is zero
    c0 = jitter(cases.dambreak3d(dx=0.04, lx=0.6, wy=0.32, lz=0.4, cw=0.3, ch=0.3, layers=0, alpha=0.0), seed=5,
                amp=0.2, drho=0.02)
    o0 = staged(c0, gd_hook=no_bulk)
    assert np.all(o0.get(A_TERMS, 1)[:, 5:8] == 0.0)
