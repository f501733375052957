"""GPU tests at the two benchmark workloads the other files only cover at small
sizes:

* the floating-body wave tank at dx 0.006 m (4.0 x 1.0 x 1.2 m tank filled to
  0.6 m, 12.5 M particles, piston 0.025 m at 1 Hz, a 0.3 x 0.3 x 0.2 m box of
  density 500 on its FP64 rigid state): one jittered state stage by stage
  against the FP64 oracle (integer stages bit-exact, every pass within the
  1e-5 tolerances of tests/helpers.py, the body's force and torque sums), then
  50 coupled steps with the fluid count unchanged and nothing removed;
* the 512-scene 2-D ensemble (3.8 M particles, per-scene grids and dt): one
  step of a sample of scenes against the FP64 oracle, 200 steps with every
  scene's fluid count unchanged and nothing removed, and a sample of scenes
  bit for bit against the same scenes run alone.
"""
import sys

import numpy as np
import pytest

import oracle
import zsph_amd as z
from zsph_amd import A_ACC, A_CELL_START, A_INDEX, A_KEY, A_POSR, A_TABLE, A_TYPE, A_VELP, BODY, FLUID, PISTON, cases
from tests.helpers import assert_close_scaled, jitter, p_scale
from tests.test_gpu_batch import compare_scene, standalone
from tests.test_gpu_parity import compare_passes

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


# ---------------------------------------------------------------- config 2: wave tank, 12.5 M
def wavetank_full():
    c = cases.wavetank3d(dx=0.006)
    assert c.n > 12_000_000
    return c


def coupled(x, c):
    x.set_body(cases.box_body(*c.box, cases.BOX_DENSITY))
    x.set_piston(cases.PISTON_AMPLITUDE, cases.PISTON_FREQUENCY, 0.0, 0)


def test_wavetank_stages_at_full_size():
    c = jitter(wavetank_full(), seed=77)
    say(f"[wave tank stages] N={c.n} grid={c.cpu.grid_xdim}x{c.cpu.grid_ydim}x{c.cpu.grid_zdim}")
    s = z.Solver(c)
    s.upload(c.host)
    o = oracle.Oracle(c.params)
    o.upload(c.host)
    for x in (s, o):
        coupled(x, c)
    for x in (s, o):
        x.hash(0); x.sort_index(); x.sort_data(0, 1)
    for a in (A_KEY, A_INDEX, A_CELL_START):
        assert np.array_equal(s.get(a), o.get(a).astype(s.get(a).dtype)), a
    for a in (A_POSR, A_VELP, A_TYPE, A_TABLE):
        assert np.array_equal(s.get(a, 1), o.get(a, 1).astype(s.get(a, 1).dtype)), a
    say("  hash, sort, cell table, gathers: bit-exact")
    cg, dg = s.neighbor_stats(1)
    co, do = o.neighbor_stats(1)
    assert np.array_equal(cg, co) and np.array_equal(dg, do)
    say(f"  neighbour sets: bit-exact (mean {cg.mean():.1f}, max {cg.max()})")
    err_a, err_d = compare_passes(c, s, o, 1)
    say(f"  passes: max scaled err acc {err_a:.2e} drho {err_d:.2e}")
    # the moving dummies' accelerations (pressure force of their fluid neighbours) and the body sums
    typ = o.get(A_TYPE, 1)
    mv = (typ == BODY) | (typ == PISTON)
    ga, oa = s.get(A_ACC), o.get(A_ACC)
    asc = o.get(oracle.A_ACC_SCALE)
    assert_close_scaled(ga[mv, :3], oa[mv, :3], np.broadcast_to(asc[mv, 0:1], ga[mv, :3].shape), "dummy acc")
    Fg, Tg = s.body_reduce(1)
    Fo, To = o.body_reduce(1)
    scale = c.params.m * np.abs(oa[typ == BODY, :3]).sum()
    say(f"  body force gpu {Fg} oracle {Fo}")
    np.testing.assert_allclose(Fg, Fo, rtol=0, atol=1e-5 * scale)
    np.testing.assert_allclose(Tg, To, rtol=0, atol=1e-5 * scale * 0.3)
    o.close()
    s.close()


def test_wavetank_coupled_steps_at_full_size():
    c = wavetank_full()
    n_fluid = int((c.host.type == FLUID).sum())
    say(f"[wave tank run] N={c.n} fluid={n_fluid}")
    s = z.Solver(c)
    s.upload(c.host)
    coupled(s, c)
    b0 = s.get_body()
    s.step(50)
    s.sync()
    d = s.diagnostics(0)
    b = s.get_body()
    say(f"  50 steps: t {s.time():.5f} s, body c {b.c[:]}, F {b.F[:]}")
    assert d["n_fluid"] == n_fluid and d["n_killed"] == 0 and s.killed() == 0
    assert d["mass_fluid"] == pytest.approx(n_fluid * c.params.m, rel=1e-12)
    assert np.all(np.isfinite(b.c[:])) and np.all(np.isfinite(b.q[:]))
    # the box starts at its equilibrium draft: over 50 steps (2.3 ms) it barely moves
    assert np.abs(np.array(b.c[:]) - np.array(b0.c[:])).max() < 0.1 * c.json.dx
    assert s.time() == pytest.approx(50 * 2 * c.params.half_dt, rel=1e-6)
    s.close()


# ---------------------------------------------------------------- config 3: 512-scene ensemble
def ensemble_full():
    b = cases.ensemble2d(n_scenes=512)
    sizes = np.array([c.n for c in b.cases])
    say(f"[ensemble] 512 scenes, N={b.n}, sizes {sizes.min()}..{sizes.max()}")
    return b, sizes


def sample_scenes(sizes, k=8):
    """The smallest and largest scenes and k - 2 spread over the batch."""
    pick = {int(np.argmin(sizes)), int(np.argmax(sizes))}
    pick |= {int(i) for i in np.linspace(1, len(sizes) - 2, k - 2).round()}
    return sorted(pick)


def test_ensemble512_one_step_against_oracle():
    b, sizes = ensemble_full()
    sb = z.Solver(b)
    sb.upload(b.host)
    sb.step(1)
    sb.sync()
    g = sb.download(0)
    orig = np.arange(b.n) + g.table  # original index of each downloaded particle
    scene = b.scene_of(orig)
    for k in sample_scenes(sizes):
        c = b.cases[k]
        o = oracle.Oracle(c.params)
        o.upload(c.host)
        o.step(1)
        r = o.download(0)
        sel = np.where(scene == k)[0]
        order_g = sel[np.argsort(orig[sel] - b.offsets[k])]
        order_r = np.argsort(np.arange(c.n) + r.table)
        f = c.host.type == FLUID
        assert np.array_equal(g.type[order_g], r.type[order_r])
        assert_close_scaled(g.rhop[order_g, 0][f], r.rhop[order_r, 0][f], np.abs(r.rhop[order_r, 0][f]),
                            f"scene {k}: rho after one step")
        assert_close_scaled(g.pos[order_g][f], r.pos[order_r][f], np.abs(r.pos[order_r][f]) + c.json.dx,
                            f"scene {k}: x after one step")
        o.close()
        say(f"  scene {k} ({c.n} particles): within tolerance of the oracle")
    sb.close()


def test_ensemble512_run_and_scenes_alone():
    b, sizes = ensemble_full()
    steps = 200
    sb = z.Solver(b)
    sb.upload(b.host)
    sb.step(steps)
    sb.sync()
    per = sb.scene_diagnostics(0)
    for k, c in enumerate(b.cases):
        nf = int((c.host.type == FLUID).sum())
        assert per[k]["n_fluid"] == nf and per[k]["n_killed"] == 0, k
        assert per[k]["mass_fluid"] == pytest.approx(nf * c.params.m, rel=1e-12)
    assert sb.killed() == 0
    say(f"  {steps} steps: every scene keeps its fluid, nothing removed")
    for k in sample_scenes(sizes, 5):
        ss = standalone(b.cases[k], steps)
        compare_scene(sb, b, k, ss)
        ss.close()
        say(f"  scene {k}: bit-identical to the scene run alone")
    sb.close()
