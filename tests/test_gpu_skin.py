"""GPU: zsph_step in skin mode (zsph_tuning.skin > 0): the neighbour structures
are rebuilt only after a particle has moved more than s/2 since the last
rebuild (or left the grid), on a grid of cells (2h + s)/k with rows trimmed
to 2h + s/2.

* One step against the FP64 oracle within the one-step tolerances, with the
  second half step on the first half step's structures (no rebuild).
* Neighbour sets after skipped rebuilds: bit-exact against the oracle's cell
  walk of the same positions (counts and digests over original ids).
* Over many steps the run tracks the reference schedule (skin 0): the 2-D
  dam break of config 0 keeps its front and probe pressures within 1 % of the
  FP64 oracle over 2000 steps; fewer rebuilds than half steps.
* A particle leaving the grid is removed at the same step as with skin 0.
* Deterministic: graph replay equals direct launches bit for bit; a restart
  from a checkpoint equals the writing run continued.
* The recorded skin walk (ZSPH_WALK_SKIN: skin masks recorded at a rebuild,
  replayed with the exact predicate until the next) matches the oracle after
  one step and tracks the tested walk over ten steps.
* With a floating body and a piston (moving dummies) the coupled state stays
  within tolerance of the reference schedule.
"""
import sys

import numpy as np
import pytest

import oracle
import zsph_amd as z
from zsph_amd import A_POSR, A_VELP, FLUID, WALL, cases
from tests.helpers import assert_close_scaled, jitter, p_scale
from tests.test_gpu_parity import by_original_id

pytestmark = pytest.mark.gpu
SKIN = 0.2


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


CASES = {
    "2d": lambda: jitter(cases.dambreak2d(), seed=12345),
    "3d-small": lambda: jitter(cases.dambreak3d(dx=0.04, lx=0.8, wy=0.4, lz=0.5, cw=0.3, ch=0.3), seed=12345),
    "3d-k2": lambda: jitter(cases.dambreak3d(dx=0.03, lx=0.9, wy=0.5, lz=0.6, cw=0.4, ch=0.35,
                                             grid_size_factor=2.0), seed=99, amp=0.3),
}


def skin_solver(c, skin=SKIN, walk=None):
    s = z.Solver(c)
    s.upload(c.host)
    t = s.get_tuning()
    s.set_tuning(t["walk"] if walk is None else walk, t["radix_bits"], skin=skin)
    return s


@pytest.mark.parametrize("walk", [z.WALK_HITS, z.WALK_SKIN])
@pytest.mark.parametrize("name", ["2d", "3d-small", "3d-k2"])
def test_skin_one_step_matches_oracle(name, walk):
    # WALK_SKIN: the first half step records the skin masks (and tests), the second
    # replays them; "3d-small" (cells of 2h) has rows beyond the record, whose
    # particles take the tested walk
    c = CASES[name]()
    s = skin_solver(c, walk=walk)
    s.step(1)
    s.sync()
    assert s.rebuilds() == 1, "the second half step should run on the first one's structures"
    o = oracle.Oracle(c.params)
    o.upload(c.host)
    o.step(1)
    g, r = s.download(0), o.download(0)
    og, orr = by_original_id(g), by_original_id(r)
    f = r.type[orr] == FLUID
    assert np.array_equal(g.type[og], r.type[orr])
    rho_g, rho_r = g.rhop[og, 0][f], r.rhop[orr, 0][f]
    assert_close_scaled(rho_g, rho_r, np.abs(rho_r), "rho after one skin-mode step")
    assert_close_scaled(g.pos[og][f], r.pos[orr][f], np.abs(r.pos[orr][f]) + c.json.dx, "x after one skin-mode step")
    # pressure: 1e-5 plus each particle's density difference through the EOS slope
    slope = c.params.cs2 * (rho_r / c.params.rho0) ** (c.params.gamma - 1.0)
    dp = np.abs(g.rhop[og, 1][f] - r.rhop[orr, 1][f]) - slope * np.abs(rho_g - rho_r)
    assert np.all(dp <= 1e-5 * np.maximum(np.abs(r.rhop[orr, 1][f]), p_scale(c)))
    sc = o.get(oracle.A_ACC_SCALE)[orr, 0][f]
    assert_close_scaled(g.acc[og, :3][f], r.acc[orr, :3][f], np.broadcast_to(sc[:, None], (f.sum(), 3)),
                        "acc after one skin-mode step", rel=3e-5)
    assert s.killed() == o.killed() == 0
    s.close()


@pytest.mark.parametrize("name", ["2d", "3d-k2"])
def test_skin_neighbour_sets_after_skipped_rebuilds(name):
    c = CASES[name]()
    s = skin_solver(c)
    steps = 8
    s.step(steps)
    s.sync()
    nreb = s.rebuilds()
    say(f"  {name}: {nreb} rebuilds in {2 * steps} half steps")
    assert nreb < 2 * steps
    cg, dg = s.neighbor_stats(0)  # the walk of the last rebuild's structures, current positions
    h = s.download(0)
    oid = np.arange(h.n) + h.table
    o = oracle.Oracle(c.params)
    o.upload(h)
    o.hash(0); o.sort_index(); o.sort_data(0, 1)
    co, do = o.neighbor_stats(1)
    otab = o.get(z.A_TABLE, 1).astype(np.int64)
    ooid = np.arange(h.n) + otab
    cnt_g = np.empty(h.n, np.int64); cnt_g[oid] = cg
    cnt_o = np.empty(h.n, np.int64); cnt_o[ooid] = co
    dig_g = np.empty(h.n, np.uint64); dig_g[oid] = dg
    dig_o = np.empty(h.n, np.uint64); dig_o[ooid] = do
    assert np.array_equal(cnt_g, cnt_o) and np.array_equal(dig_g, dig_o)
    say(f"  neighbour sets bit-exact (mean {cg.mean():.1f})")
    s.close()


def test_skin_long_run_2d_front_and_probes():
    c = cases.dambreak2d()
    pts = np.array([[0.0125, 0.0, 0.05], [0.0125, 0.0, 0.15], [0.2, 0.0, 0.02]], np.float32)
    steps = 2000
    say(f"[skin long run 2d] N={c.n}, {steps} steps")
    s = skin_solver(c)
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
            f"probes gpu {np.round(pg, 1)} oracle {np.round(po, 1)}; rebuilds {s.rebuilds()}")
        assert dg["n_fluid"] == do["n_fluid"] == c.json.water_ptc_num
        assert dg["n_killed"] == do["n_killed"] == 0
        assert dg["front_x"] == pytest.approx(do["front_x"], rel=0.01)
        np.testing.assert_allclose(pg, po, rtol=0.01, atol=0.01 * pscale)
    assert s.rebuilds() < steps  # fewer than one per step (two half steps each)
    s.close()


def test_skin_removal_at_the_reference_step():
    # a fast particle crossing the end of the grid is removed at the same step as with skin 0
    c = CASES["3d-small"]()
    f = np.where(c.host.type == FLUID)[0]
    k = f[np.argmax(c.host.pos[f, 0])]
    gx1 = c.cpu.grid_xmin + c.cpu.grid_xdim * c.cpu.grid_size
    c.host.pos[k] = (gx1 - 0.3 * c.json.dx, c.host.pos[k, 1], c.host.pos[k, 2])
    c.host.vel[k] = (8.0, 0.0, 0.0)
    runs = {}
    for skin in (0.0, SKIN):
        s = skin_solver(c, skin)
        seq = []
        for _ in range(12):
            s.step(1)
            seq.append(s.killed())
        runs[skin] = seq
        s.close()
    say(f"  removed counts: {runs}")
    assert runs[0.0] == runs[SKIN] and runs[0.0][-1] >= 1


def test_skin_walk_tracks_the_tested_walk():
    # the recorded skin walk visits the same neighbours as the tested walks (other order):
    # over 10 skin-mode steps both stay within FP32 summation-order noise of each other
    c = CASES["3d-k2"]()
    out = {}
    for walk in (z.WALK_HITS, z.WALK_SKIN):
        s = skin_solver(c, walk=walk)
        s.step(10)
        s.sync()
        h = s.download(0)
        order = by_original_id(h)
        out[walk] = (h.pos[order], h.rhop[order], s.rebuilds())
        s.close()
    say(f"  rebuilds: hits {out[z.WALK_HITS][2]}, skin walk {out[z.WALK_SKIN][2]} in 20 half steps")
    assert out[z.WALK_SKIN][2] < 20
    dx = np.abs(out[z.WALK_SKIN][0] - out[z.WALK_HITS][0]).max()
    drho = np.abs(out[z.WALK_SKIN][1][:, 0] - out[z.WALK_HITS][1][:, 0]).max()
    say(f"  max |dx| {dx:.3e} m ({dx / c.json.dx:.2e} dx), max |drho| {drho:.3e}")
    assert dx < 1e-5 * c.json.dx and drho < 1e-5 * c.json.rho0


def test_skin_graph_replay_equals_direct_launches_and_restart(tmp_path):
    c = CASES["3d-k2"]()
    out = []
    for graph in (True, False):
        s = skin_solver(c, walk=z.WALK_SKIN)
        s.set_graph(graph)
        s.step(6)
        s.sync()
        out.append((s.get(A_POSR, 0), s.get(A_VELP, 0), s.rebuilds()))
        s.close()
    assert out[0][2] == out[1][2]
    for a, b in zip(out[0][:2], out[1][:2]):
        assert np.array_equal(a, b)
    # restart: a checkpoint makes the writer's next step rebuild, as the restored context's first
    path = str(tmp_path / "skin.bin")
    s = skin_solver(c, walk=z.WALK_SKIN)
    s.step(5)
    s.checkpoint(path, 5)
    s.step(7)
    s.sync()
    ref = (s.get(A_POSR, 0), s.get(A_VELP, 0))
    s.close()
    t = skin_solver(c, walk=z.WALK_SKIN)
    t.restore(path)
    t.step(7)
    t.sync()
    for a, b in zip(ref, (t.get(A_POSR, 0), t.get(A_VELP, 0))):
        assert np.array_equal(a, b)
    t.close()


def test_skin_with_body_and_piston_tracks_reference_schedule():
    c = cases.wavetank3d(dx=0.04, depth=0.6)
    body = cases.box_body(*c.box, cases.BOX_DENSITY)
    res = {}
    for skin in (0.0, SKIN):
        s = skin_solver(c, skin)
        s.set_body(body)
        s.set_piston(cases.PISTON_AMPLITUDE, cases.PISTON_FREQUENCY, 0.0, 0)
        s.step(200)
        s.sync()
        b = s.get_body()
        res[skin] = (np.array(b.c[:]), np.array(b.q[:]), s.diagnostics(0), s.rebuilds())
        s.close()
    c0 = np.array(body.c[:])
    motion = np.abs(res[0.0][0] - c0).max() + 1e-6
    say(f"  body dc ref {res[0.0][0] - c0} skin {res[SKIN][0] - c0}; rebuilds {res[SKIN][3]} in 400 half steps")
    np.testing.assert_allclose(res[SKIN][0], res[0.0][0], rtol=0, atol=0.01 * motion)
    np.testing.assert_allclose(res[SKIN][1], res[0.0][1], rtol=0, atol=1e-6)
    assert res[SKIN][2]["n_fluid"] == res[0.0][2]["n_fluid"] and res[SKIN][2]["n_killed"] == 0
    assert res[SKIN][3] < 400
