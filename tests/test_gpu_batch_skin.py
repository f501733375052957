"""GPU: skin mode for batches (zsph_tuning.skin > 0 in a zsph_create_batch
context).  Each scene has its own skin grid, cells (2h + s)/k of its own h,
laid out scene after scene in the key space; a particle of any scene that
moves more than its scene's s/2 since the last rebuild triggers the rebuild
of all scenes.

* One step against the FP64 oracle within the one-step tolerances (density,
  position, pressure), for the largest and the smallest scene, with the second
  half step on the first one's structures.
* One step equal bit for bit to each scene's own skin-mode run (same rebuild
  schedule, same skin grid, same particle order).
* Neighbour sets after skipped rebuilds equal those of the same state sorted on
  the case grids (counts and digests over original ids, every scene at once).
* Over a run, every scene tracks the reference schedule (skin 0): fluid kept,
  fronts within 1 %, fewer rebuilds than half steps; per-scene diagnostics read
  the skin grids' cell table after skin-mode steps.
* Graph replay equals direct launches bit for bit; the recorded skin walk is
  refused for batches.
"""
import sys

import numpy as np
import pytest

import oracle
import zsph_amd as z
from zsph_amd import FLUID, cases
from tests.helpers import assert_close_scaled, p_scale

pytestmark = pytest.mark.gpu
SKIN = 0.2


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def batch(n, seed):
    b = cases.ensemble2d(n_scenes=n, seed=seed)
    sizes = [c.n for c in b.cases]
    say(f"[batch skin] {n} scenes, N={b.n}, sizes {min(sizes)}..{max(sizes)}")
    return b


def solver(b, skin=SKIN, walk=z.WALK_HITS, graph=True):
    s = z.Solver(b)
    s.upload(b.host)
    s.set_tuning(walk, s.get_tuning()["radix_bits"], skin=skin)
    s.set_graph(graph)
    return s


def test_batch_skin_one_step_matches_oracle():
    b = batch(4, seed=5)
    s = solver(b)
    s.step(1)
    s.sync()
    assert s.rebuilds() == 1, "the second half step should run on the first one's structures"
    g = s.download(0)
    orig = np.arange(b.n) + g.table
    sizes = [c.n for c in b.cases]
    for k in (int(np.argmax(sizes)), int(np.argmin(sizes))):
        c = b.cases[k]
        o = oracle.Oracle(c.params)
        o.upload(c.host)
        o.step(1)
        r = o.download(0)
        sel = np.where(b.scene_of(orig) == k)[0]
        order_g = sel[np.argsort(orig[sel] - b.offsets[k])]
        order_r = np.argsort(np.arange(c.n) + r.table)
        f = c.host.type == FLUID
        assert np.array_equal(g.type[order_g], r.type[order_r])
        rho_g, rho_r = g.rhop[order_g, 0][f], r.rhop[order_r, 0][f]
        assert_close_scaled(rho_g, rho_r, np.abs(rho_r), f"scene {k}: rho after one skin-mode step")
        assert_close_scaled(g.pos[order_g][f], r.pos[order_r][f], np.abs(r.pos[order_r][f]) + c.json.dx,
                            f"scene {k}: x after one skin-mode step")
        slope = c.params.cs2 * (rho_r / c.params.rho0) ** (c.params.gamma - 1.0)
        dp = np.abs(g.rhop[order_g, 1][f] - r.rhop[order_r, 1][f]) - slope * np.abs(rho_g - rho_r)
        assert np.all(dp <= 1e-5 * np.maximum(np.abs(r.rhop[order_r, 1][f]), p_scale(c)))
        say(f"  scene {k} ({c.n} particles) within tolerance of the oracle")
    assert s.killed() == 0
    s.close()


def test_batch_skin_step_equals_standalone_skin_scenes():
    # Over one step a batch and each of its scenes alone rebuild on the same
    # schedule (the first half step, on each scene's identical skin grid), so
    # every scene's particles come out in the same order with the same bits;
    # the standalone skin path is checked against the oracle in test_gpu_skin.py.
    b = batch(5, seed=13)
    s = solver(b)
    s.step(1)
    s.sync()
    assert s.rebuilds() == 1
    g = s.download(0)
    orig = np.arange(b.n) + g.table
    s.close()
    for k, c in enumerate(b.cases):
        a = solver(c)
        a.step(1)
        a.sync()
        assert a.rebuilds() == 1
        h = a.download(0)
        a.close()
        sel = np.where(b.scene_of(orig) == k)[0]
        order_g = sel[np.argsort(orig[sel] - b.offsets[k])]
        order_h = np.argsort(np.arange(c.n) + h.table)
        # the same particle order: scene k's rows of the batch (fluid, then dummies) hold the
        # standalone rows' particles in the standalone order
        assert np.array_equal(orig[sel] - b.offsets[k], np.arange(c.n) + h.table)
        for f in ("pos", "vel", "rhop", "acc", "type"):
            assert np.array_equal(getattr(g, f)[order_g], getattr(h, f)[order_h]), f"scene {k}: {f}"
    say(f"  {b.nscenes} scenes bit-identical to their standalone skin-mode runs")


def test_batch_skin_neighbour_sets_after_skipped_rebuilds():
    b = batch(6, seed=21)
    s = solver(b)
    steps = 10
    s.step(steps)
    s.sync()
    nreb = s.rebuilds()
    say(f"  {nreb} rebuilds in {2 * steps} half steps")
    assert nreb < 2 * steps
    cg, dg = s.neighbor_stats(0)  # the last rebuild's skin structures, current positions
    h = s.download(0)
    s.close()
    # the same state sorted on the case grids (the reference schedule's structures)
    r = z.Solver(b)
    r.upload(h)
    r.hash(0)
    r.sort_index()
    r.sort_data(0, 1)
    cr, dr = r.neighbor_stats(1)
    tr = r.download(1).table.astype(np.int64)
    r.close()
    oid_g = np.arange(b.n) + h.table.astype(np.int64)
    oid_r = np.arange(b.n) + tr
    cnt_g = np.empty(b.n, np.int64); cnt_g[oid_g] = cg
    cnt_r = np.empty(b.n, np.int64); cnt_r[oid_r] = cr
    dig_g = np.empty(b.n, np.uint64); dig_g[oid_g] = dg
    dig_r = np.empty(b.n, np.uint64); dig_r[oid_r] = dr
    assert np.array_equal(cnt_g, cnt_r) and np.array_equal(dig_g, dig_r)
    say(f"  neighbour sets bit-exact over {b.n} particles (mean {cg.mean():.1f})")


def test_batch_skin_run_tracks_reference_schedule():
    b = batch(8, seed=3)
    steps = 600
    runs = {}
    for skin in (0.0, 0.1):
        s = solver(b, skin=skin)
        s.step(steps)
        s.sync()
        d = s.scene_diagnostics(0, 0.05)
        runs[skin] = (d, s.rebuilds(), s.killed())
        s.close()
    (d0, _, k0), (d1, nreb, k1) = runs[0.0], runs[0.1]
    say(f"  skin 0.1: {nreb} rebuilds in {2 * steps} half steps")
    assert 0 < nreb < steps
    assert k0 == k1 == 0
    for k, (a, c) in enumerate(zip(d0, d1)):
        assert a["n_fluid"] == c["n_fluid"] == b.cases[k].json.water_ptc_num
        assert a["mass_fluid"] == c["mass_fluid"]
        assert c["front_x"] == pytest.approx(a["front_x"], rel=0.01), f"scene {k}"
    fr = [abs(c["front_x"] / a["front_x"] - 1.0) for a, c in zip(d0, d1)]
    say(f"  fronts: largest relative difference {max(fr):.2e}")


def test_batch_skin_graph_equals_direct_launches():
    b = batch(5, seed=8)
    out = []
    for graph in (True, False):
        s = solver(b, graph=graph)
        s.step(30)
        s.sync()
        out.append((s.download(0), s.rebuilds()))
        s.close()
    (g, ng), (d, nd) = out
    assert ng == nd
    for f in ("pos", "vel", "rhop", "type", "table"):
        assert np.array_equal(getattr(g, f), getattr(d, f)), f


def test_batch_refuses_the_recorded_skin_walk():
    b = batch(2, seed=9)
    s = z.Solver(b)
    with pytest.raises(z.ZsphError) as e:
        s.set_tuning(z.WALK_SKIN, s.get_tuning()["radix_bits"], skin=0.1)
    assert e.value.code == z.abi.E_ARG
    s.close()
