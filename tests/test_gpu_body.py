"""GPU: floating body and piston against the FP64 oracle and closed forms.

Body force/torque after one governing pass, the coupled trajectory over a
few steps (body state, particles), free fall and the piston's closed form,
Newton's third law between fluid and body, graph replay = staged calls, and
the error paths of the set-up calls.
"""
import sys

import numpy as np
import pytest

import oracle
import zsph_amd as z
from zsph_amd import A_ACC, A_POSR, A_TYPE, A_VELP, BODY, FLUID, PISTON, cases
from tests.helpers import assert_close_scaled, p_scale
from tests.test_body import HI, LO, by_original, subset_case, tank

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def setup(c, body=True, piston=True, A=0.02, f=1.5):
    say(f"[body case] N={c.n} grid={c.cpu.grid_xdim}x{c.cpu.grid_ydim}x{c.cpu.grid_zdim} "
        f"body={body} piston={piston}")
    s = z.Solver(c)
    s.upload(c.host)
    o = oracle.Oracle(c.params)
    o.upload(c.host)
    if body:
        r = cases.box_body(LO, HI, 500.0)
        s.set_body(r)
        o.set_body(r)
    if piston:
        s.set_piston(A, f, 0.0, 0)
        o.set_piston(A, f, 0.0, 0)
    return s, o


def test_body_force_matches_oracle():
    c = tank(0.02)
    s, o = setup(c, piston=False)
    for x in (s, o):
        x.hash(0); x.sort_index(); x.sort_data(0, 1); x.boundary(1); x.governing(1)
    Fg, Tg = s.body_reduce(1)
    Fo, To = o.body_reduce(1)
    say(f"  F gpu {Fg} oracle {Fo}")
    scale = c.params.m * np.abs(o.get(A_ACC, 1)[o.get(A_TYPE, 1) == BODY, :3]).sum()
    np.testing.assert_allclose(Fg, Fo, atol=1e-5 * scale)
    np.testing.assert_allclose(Tg, To, atol=1e-5 * scale * 0.3)
    V = np.prod(np.subtract(HI, LO)) * 0.12 / (HI[2] - LO[2])
    assert Fg[2] == pytest.approx(1000.0 * 9.81 * V, rel=0.03)
    s.close()


def test_coupled_trajectory_matches_oracle():
    c = tank(0.04)
    s, o = setup(c)
    n = 30
    s.step(n)
    s.sync()
    o.step(n)
    assert s.time() == o.time()
    bg, bo = s.get_body(), o.get_body()
    say(f"  c gpu {bg.c[:]} oracle {bo.c[:]}  v gpu {bg.vel[:]} oracle {bo.vel[:]}")
    np.testing.assert_allclose(bg.c[:], bo.c[:], rtol=0, atol=1e-6)
    vs = np.abs(bo.vel[:]).max() + 1e-3
    np.testing.assert_allclose(bg.vel[:], bo.vel[:], rtol=0, atol=1e-3 * vs)
    np.testing.assert_allclose(bg.omega[:], bo.omega[:], rtol=0, atol=1e-3 * (np.abs(bo.omega[:]).max() + 1e-3))
    np.testing.assert_allclose(bg.q[:], bo.q[:], rtol=0, atol=1e-6)
    Fs = np.abs(bo.F[:]).max()
    np.testing.assert_allclose(bg.F[:], bo.F[:], rtol=0, atol=2e-3 * Fs)
    g, r = s.download(0), o.download(0)
    og, orr = by_original(g), by_original(r)
    assert np.array_equal(g.type[og], r.type[orr])
    assert_close_scaled(g.pos[og], r.pos[orr], np.abs(r.pos[orr]) + c.json.dx, "x after the coupled steps", rel=1e-5)
    f = r.type[orr] == FLUID
    assert_close_scaled(g.rhop[og, 1][f], r.rhop[orr, 1][f], np.maximum(np.abs(r.rhop[orr, 1][f]), p_scale(c)),
                        "p after the coupled steps", rel=1e-4)
    s.close()


def test_free_fall_and_piston_closed_forms():
    lo, hi = (0.64, 0.12, 0.48), (0.88, 0.28, 0.56)
    c = tank(0.04, lo, hi)
    say(f"[free fall] N={c.n}")
    s = z.Solver(c)
    s.upload(c.host)
    body = cases.box_body(lo, hi, 500.0)
    s.set_body(body)
    A, f = 0.02, 1.5
    s.set_piston(A, f, 0.0, 0)
    n = 40
    s.step(n)
    s.sync()
    t = s.time()
    assert t == pytest.approx(n * 2.0 * np.float64(np.float32(c.params.half_dt)), rel=1e-14)
    b = s.get_body()
    np.testing.assert_allclose(b.c[:], np.array(body.c[:]) + [0, 0, 0.5 * c.json.g * t * t], rtol=1e-13, atol=1e-14)
    np.testing.assert_allclose(b.vel[:], [0, 0, c.json.g * t], rtol=1e-12, atol=1e-15)
    assert list(b.F[:]) == [0.0, 0.0, 0.0]
    h = s.download(0)
    o = by_original(h)
    ids = np.arange(c.n)[c.host.type == BODY]
    np.testing.assert_allclose(h.pos[o[ids]], c.host.pos[ids] + np.float32([0, 0, 0.5 * c.json.g * t * t]), atol=2e-7)
    pid = np.arange(c.n)[c.host.type == PISTON]
    np.testing.assert_allclose(h.pos[o[pid], 0], c.host.pos[pid, 0] + A * np.sin(2 * np.pi * f * t), atol=1e-7)
    np.testing.assert_allclose(h.vel[o[pid], 0], A * 2 * np.pi * f * np.cos(2 * np.pi * f * t), rtol=1e-6)
    s.close()


def test_body_fluid_forces_cancel():
    src = tank(0.04)
    keep = (src.host.type == FLUID) | (src.host.type == BODY)
    c = subset_case(src, keep, g=0.0)
    fl = c.host.type == FLUID
    rng = np.random.default_rng(3)
    c.host.rhop[fl, 0] = np.float32(1000.0 * (1.0 + 0.01 * rng.random(int(fl.sum()))))
    c.host.rhop[fl, 1] = [np.float32(oracle.eos_p(c.params, float(r))) for r in c.host.rhop[fl, 0]]
    say(f"[third law] N={c.n}")
    s = z.Solver(c)
    s.upload(c.host)
    body = cases.box_body(LO, HI, 500.0)
    s.set_body(body)
    s.hash(0); s.sort_index(); s.sort_data(0, 1); s.boundary(1); s.governing(1)
    Fb, Tb = s.body_reduce(1)
    acc = s.get(A_ACC, 1)[:, :3].astype(np.float64)
    pos = s.get(A_POSR, 1)[:, :3].astype(np.float64)
    ff = s.get(A_TYPE, 1) == FLUID
    m = c.params.m
    Ff = m * acc[ff].sum(0)
    Tf = np.cross(pos[ff] - np.array(body.c[:]), m * acc[ff]).sum(0)
    scale = m * np.abs(acc[ff]).sum()
    assert np.abs(Fb).max() > 1e-3 * scale
    # FP32 pair terms: cancellation to FP32 rounding of the summed magnitudes
    np.testing.assert_allclose(Fb + Ff, 0.0, atol=1e-6 * scale)
    np.testing.assert_allclose(Tb + Tf, 0.0, atol=1e-6 * scale)
    s.close()


def test_graph_replay_equals_stages_with_body():
    c = tank(0.04)
    runs = []
    for mode in ("graph", "graph", "stages"):
        say(f"[body determinism] {mode}")
        s = z.Solver(c)
        s.upload(c.host)
        s.set_body(cases.box_body(LO, HI, 500.0))
        s.set_piston(0.02, 1.5, 0.3, 0)
        if mode == "graph":
            s.step(4)
        else:
            for _ in range(4):
                s.hash(0); s.sort_index(); s.sort_data(0, 1); s.boundary(1); s.governing(1); s.prediction(1)
                s.hash(1); s.sort_index(); s.sort_data(1, 0); s.boundary(0); s.governing(0); s.correction(0)
        s.sync()
        b = s.get_body()
        runs.append((s.get(A_POSR, 0), s.get(A_VELP, 0), np.array(b.c[:] + b.vel[:] + b.omega[:] + b.q[:]),
                     s.time()))
        s.close()
    for other in runs[1:]:
        for a, b in zip(runs[0], other):
            assert np.array_equal(a, b)


def test_setup_errors_and_upload_reset():
    c = cases.dambreak3d(dx=0.04, lx=0.8, wy=0.4, lz=0.5, cw=0.3, ch=0.3)
    say(f"[body errors] N={c.n}")
    s = z.Solver(c)
    s.upload(c.host)
    with pytest.raises(z.ZsphError) as e:
        s.get_body()
    assert e.value.code == z.abi.E_STATE
    with pytest.raises(z.ZsphError) as e:
        s.set_body(cases.box_body(LO, HI, 500.0))  # no type-2 particles
    assert e.value.code == z.abi.E_ARG
    with pytest.raises(z.ZsphError) as e:
        s.set_piston(0.01, 1.0, 0.0, 0)  # no type-3 particles
    assert e.value.code == z.abi.E_ARG
    with pytest.raises(z.ZsphError):
        s.set_piston(0.01, 1.0, 0.0, 5)
    s.close()
    c = tank(0.04)
    s = z.Solver(c)
    s.upload(c.host)
    bad = cases.box_body(LO, HI, 500.0)
    bad.mass = 0.0
    with pytest.raises(z.ZsphError):
        s.set_body(bad)
    s.set_time(0.25)
    s.set_body(cases.box_body(LO, HI, 500.0))
    s.step(2)
    s.sync()
    t = s.time()
    assert t == pytest.approx(0.25 + 4 * np.float64(np.float32(c.params.half_dt)), rel=1e-14)
    s.upload(c.host)  # new particles: the body set-up ends, the clock stays
    with pytest.raises(z.ZsphError):
        s.get_body()
    assert s.time() == t
    s.step(1)
    s.sync()
    assert s.time() == pytest.approx(t + 2 * np.float64(np.float32(c.params.half_dt)), rel=1e-14)
    s.close()
