"""Long-run accuracy of the GPU step against the FP64 oracle: mass is exact
(no particle removed, fluid count constant) and the wave front and probe
pressures stay within 1 % of the oracle's over the 2-D dam break of config 0
(2000 steps) and a coarse 3-D dam break.
"""
import sys

import numpy as np
import pytest

import oracle
import zsph_amd as z
from zsph_amd import cases

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_diagnostics_match_oracle():
    c = cases.dambreak2d()
    s = z.Solver(c)
    s.upload(c.host)
    o = oracle.Oracle(c.params)
    o.upload(c.host)
    pts = np.array([[0.0125, 0.0, 0.05], [0.2, 0.0, 0.02], [1.0, 0.0, 0.1]], np.float32)
    for step in (0, 5):
        if step:
            s.step(step)
            o.step(step)
        dg, do = s.diagnostics(0, 0.05), o.diagnostics(0, 0.05)
        assert dg["n_fluid"] == do["n_fluid"] == 4800
        assert dg["n_killed"] == do["n_killed"] == 0
        assert dg["mass_fluid"] == pytest.approx(do["mass_fluid"], rel=1e-12)
        assert dg["front_x"] == pytest.approx(do["front_x"], rel=1e-5)
        assert dg["kinetic"] == pytest.approx(do["kinetic"], rel=1e-4, abs=1e-9)
        pg, po = s.probe_pressure(pts), o.probe_pressure(pts)
        np.testing.assert_allclose(pg, po, rtol=1e-5, atol=1e-5 * 1000 * 9.81 * 0.3)
        assert pg[2] == 0.0  # dry point


@pytest.mark.parametrize("name,steps", [("2d", 2000), ("3d-coarse", 600)])
def test_long_run_front_and_probes(name, steps):
    if name == "2d":
        c = cases.dambreak2d()
        pts = np.array([[0.0125, 0.0, 0.05], [0.0125, 0.0, 0.15], [0.2, 0.0, 0.02]], np.float32)
    else:
        c = cases.dambreak3d(dx=0.04)
        pts = np.array([[0.06, 0.5, 0.06], [0.3, 0.5, 0.06], [0.06, 0.5, 0.3]], np.float32)
    say(f"[long run {name}] N={c.n}, {steps} steps")
    s = z.Solver(c)
    s.upload(c.host)
    o = oracle.Oracle(c.params)
    o.upload(c.host)
    pscale = c.json.rho0 * abs(c.json.g) * 0.3
    chunk = steps // 4
    for k in range(4):
        s.step(chunk)
        o.step(chunk)
        dg, do = s.diagnostics(0, 3 * c.json.dx), o.diagnostics(0, 3 * c.json.dx)
        pg, po = s.probe_pressure(pts), o.probe_pressure(pts)
        say(f"  step {(k + 1) * chunk}: front gpu {dg['front_x']:.5f} oracle {do['front_x']:.5f}; "
            f"probes gpu {np.round(pg, 1)} oracle {np.round(po, 1)}")
        # mass exact: every fluid particle kept, none removed
        assert dg["n_fluid"] == do["n_fluid"] == c.json.water_ptc_num
        assert dg["n_killed"] == do["n_killed"] == 0
        assert dg["front_x"] == pytest.approx(do["front_x"], rel=0.01)
        np.testing.assert_allclose(pg, po, rtol=0.01, atol=0.01 * pscale)


def test_floating_box_holds_its_equilibrium_draft():
    # a box of half the water's density, released at its equilibrium draft in
    # still water, stays there (Archimedes): heave within 0.1 dx, pitch/roll
    # within 0.05 degrees over 3000 steps (0.45 s)
    dx = 0.02
    c = cases.wavetank3d(dx=dx)
    say(f"[floating equilibrium] N={c.n}")
    s = z.Solver(c)
    s.upload(c.host)
    body = cases.box_body(*c.box, cases.BOX_DENSITY)
    s.set_body(body)
    c0 = np.array(body.c[:])
    for k in range(6):
        s.step(500)
        b = s.get_body()
        dc = np.array(b.c[:]) - c0
        angle = 2.0 * np.degrees(np.arccos(min(1.0, abs(b.q[0]))))
        say(f"  step {500 * (k + 1)}: dc {dc} angle {angle:.4f} deg F {b.F[:]}")
        assert abs(dc[2]) < 0.1 * dx
        assert np.abs(dc[:2]).max() < 0.05 * dx
        assert angle < 0.05
    assert s.killed() == 0
    s.close()


def test_floating_box_with_piston_tracks_oracle():
    # piston-driven tank with the box, 300 steps: body trajectory within 1 %
    # of its motion and particle positions within 1e-4 dx of the FP64 oracle
    c = cases.wavetank3d(dx=0.04, depth=0.6)
    say(f"[floating + piston vs oracle] N={c.n}")
    s = z.Solver(c)
    s.upload(c.host)
    o = oracle.Oracle(c.params)
    o.upload(c.host)
    body = cases.box_body(*c.box, cases.BOX_DENSITY)
    for x in (s, o):
        x.set_body(body)
        x.set_piston(cases.PISTON_AMPLITUDE, cases.PISTON_FREQUENCY, 0.0, 0)
    c0 = np.array(body.c[:])
    for k in range(3):
        s.step(100)
        o.step(100)
        bg, bo = s.get_body(), o.get_body()
        dg, do = np.array(bg.c[:]) - c0, np.array(bo.c[:]) - c0
        say(f"  step {100 * (k + 1)}: dc gpu {dg} oracle {do}")
        motion = np.abs(do).max() + 1e-6
        np.testing.assert_allclose(dg, do, rtol=0, atol=0.01 * motion)
        np.testing.assert_allclose(bg.q[:], bo.q[:], rtol=0, atol=1e-6)
    assert s.time() == o.time()
    d = s.diagnostics(0)
    assert d["n_fluid"] == int((c.host.type == z.FLUID).sum()) and d["n_killed"] == 0
    s.close()
