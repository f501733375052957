"""GPU: batched scenes equal the same scenes run one by one, bit for bit.

A batch concatenates the scenes' cell grids in one key space, so after the
sort each scene's fluid (and its dummies) is a contiguous, identically
ordered copy of the standalone run's; every kernel then computes the same
FP32 operations in the same order.  Checked on an uneven ensemble (smallest
and largest scenes included): cell tables, sorted arrays, the state after
graph-replayed steps, per-scene diagnostics, and independence from the
launch tuning.  The standalone runs themselves are pinned against the FP64
oracle by test_gpu_parity.py; one batched scene is also checked against the
oracle directly.
"""
import sys

import numpy as np
import pytest

import oracle
import zsph_amd as z
from zsph_amd import A_CELL_START, A_POSR, A_TYPE, A_VELP, FLUID, cases
from tests.helpers import assert_close_scaled, p_scale

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def uneven_batch(n=10, seed=11):
    b = cases.ensemble2d(n_scenes=n, seed=seed)
    sizes = [c.n for c in b.cases]
    say(f"[batch] {n} scenes, N={b.n}, sizes {min(sizes)}..{max(sizes)}, cells {b.grid_num}")
    return b


def scene_ranges(cs, G, k_base, Gk):
    """(fluid range, dummy range) of one scene in a cell table with G cells."""
    return (cs[k_base], cs[k_base + Gk]), (cs[G + 1 + k_base], cs[G + 1 + k_base + Gk])


def standalone(case, steps=0, walk=None):
    s = z.Solver(case)
    s.upload(case.host)
    if walk is not None:
        s.set_tuning(walk, s.get_tuning()["radix_bits"])
    if steps:
        s.step(steps)
        s.sync()
    return s


def compare_scene(sb, b, k, ss, buf=0):
    """Bitwise comparison of scene k of batch solver sb with standalone solver ss (both in cell order)."""
    G = b.grid_num
    bases = np.cumsum([0] + [c.cpu.grid_num for c in b.cases])
    csb, css = sb.get(A_CELL_START), ss.get(A_CELL_START)
    Gk = b.cases[k].cpu.grid_num
    (fb0, fb1), (db0, db1) = scene_ranges(csb, G, bases[k], Gk)
    (fs0, fs1), (ds0, ds1) = scene_ranges(css, Gk, 0, Gk)
    assert fb1 - fb0 == fs1 - fs0 and db1 - db0 == ds1 - ds0, f"scene {k}: segment sizes differ"
    np.testing.assert_array_equal(csb[bases[k]:bases[k] + Gk + 1] - fb0, css[:Gk + 1] - fs0)
    np.testing.assert_array_equal(csb[G + 1 + bases[k]:G + 2 + bases[k] + Gk] - db0, css[Gk + 1:] - ds0)
    for a in (A_POSR, A_VELP, A_TYPE):
        xb, xs = sb.get(a, buf), ss.get(a, buf)
        assert np.array_equal(xb[fb0:fb1], xs[fs0:fs1]), f"scene {k}: fluid array {a} differs"
        assert np.array_equal(xb[db0:db1], xs[ds0:ds1]), f"scene {k}: dummy array {a} differs"


def test_batch_sort_matches_standalone_scenes():
    b = uneven_batch()
    sb = z.Solver(b)
    sb.upload(b.host)
    sb.hash(0); sb.sort_index(); sb.sort_data(0, 1)
    for k, c in enumerate(b.cases):
        ss = standalone(c)
        ss.hash(0); ss.sort_index(); ss.sort_data(0, 1)
        compare_scene(sb, b, k, ss, buf=1)
        ss.close()
    sb.close()


@pytest.mark.parametrize("walk", [z.WALK_MASK, z.WALK_HITS, z.WALK_LIST])
def test_batch_steps_match_standalone_scenes(walk):
    b = uneven_batch()
    steps = 12
    sb = z.Solver(b)
    sb.upload(b.host)
    sb.set_tuning(walk, sb.get_tuning()["radix_bits"])
    sb.step(steps)
    sb.sync()
    dsb = sb.scene_diagnostics(0, 0.05)
    for k, c in enumerate(b.cases):
        ss = standalone(c, steps, walk)
        compare_scene(sb, b, k, ss)
        dss = ss.scene_diagnostics(0, 0.05)[0]
        assert dsb[k] == dss, f"scene {k}: diagnostics differ"
        assert dss["n_fluid"] == int((c.host.type == FLUID).sum())
        ss.close()
    tot = sb.diagnostics(0, 0.05)
    assert tot["n_fluid"] == sum(d["n_fluid"] for d in dsb) and tot["n_killed"] == 0
    sb.close()


@pytest.mark.parametrize("walk", [z.WALK_MASK, z.WALK_HITS, z.WALK_LIST])
def test_mixed_diffusion_batch_matches_standalone_scenes(walk):
    # Antuono and Molteni-Colagrossi scenes in one batch: the batched kernels
    # read each particle's diffusion form from its scene's parameters, the
    # standalone ones have it fixed at compile time; the bits must agree
    anto = cases.ensemble2d(n_scenes=3, seed=5).cases
    mc = cases.ensemble2d(n_scenes=3, seed=6, diffusion=z.DIFF_MC).cases
    b = z.Batch([c for pair in zip(anto, mc) for c in pair])
    assert {c.params.diffusion for c in b.cases} == {z.DIFF_ANTUONO, z.DIFF_MC}
    say(f"[mixed diffusion batch] {len(b.cases)} scenes, N={b.n}")
    steps = 8
    sb = z.Solver(b)
    sb.upload(b.host)
    sb.set_tuning(walk, sb.get_tuning()["radix_bits"])
    sb.step(steps)
    sb.sync()
    for k, c in enumerate(b.cases):
        ss = standalone(c, steps, walk)
        compare_scene(sb, b, k, ss)
        ss.close()
    sb.close()


def test_batch_graph_replay_is_deterministic_and_tuning_independent():
    b = uneven_batch(6, seed=3)
    runs = []
    for walk, block, tile in ((z.WALK_MASK, 256, 16), (z.WALK_MASK, 256, 16), (z.WALK_HITS, 512, 8),
                              (z.WALK_LIST, 128, 16)):
        s = z.Solver(b)
        s.upload(b.host)
        s.set_tuning(walk, 9, block, tile)
        s.step(5)
        s.sync()
        runs.append((s.get(A_POSR), s.get(A_VELP)))
        s.close()
    for r in runs[1:]:
        assert np.array_equal(r[0], runs[0][0]) and np.array_equal(r[1], runs[0][1])


def test_batched_scene_matches_oracle_after_one_step():
    b = uneven_batch(4, seed=5)
    sb = z.Solver(b)
    sb.upload(b.host)
    sb.step(1)
    sb.sync()
    g = sb.download(0)
    orig = np.arange(b.n) + g.table  # original index of each downloaded particle
    k = int(np.argmax([c.n for c in b.cases]))
    c = b.cases[k]
    o = oracle.Oracle(c.params)
    o.upload(c.host)
    o.step(1)
    r = o.download(0)
    sel = np.where(b.scene_of(orig) == k)[0]
    gi = orig[sel] - b.offsets[k]  # standalone original index
    order_g = sel[np.argsort(gi)]
    order_r = np.argsort(np.arange(c.n) + r.table)
    f = c.host.type == FLUID
    assert np.array_equal(g.type[order_g], r.type[order_r])
    assert_close_scaled(g.pos[order_g][f], r.pos[order_r][f], np.abs(r.pos[order_r][f]) + c.json.dx,
                        "x after one batched step")
    assert_close_scaled(g.rhop[order_g, 1][f], r.rhop[order_r, 1][f],
                        np.maximum(np.abs(r.rhop[order_r, 1][f]), p_scale(c)), "p after one batched step")
    sb.close()


def test_batch_reupload_of_a_sorted_state_keeps_scenes():
    # A state downloaded in cell order and uploaded again: each particle's scene
    # follows its original index (row + table), so the run continues bit for bit.
    b = uneven_batch(5, seed=17)
    s1 = z.Solver(b)
    s1.upload(b.host)
    s1.step(5)
    s1.sync()
    h = s1.download(0)
    assert not np.array_equal(h.table, np.zeros_like(h.table))  # in cell order, not upload order
    s2 = z.Solver(b)
    s2.upload(h)
    for s in (s1, s2):
        s.step(5)
        s.sync()
    g1, g2 = s1.download(0), s2.download(0)
    for f in ("pos", "vel", "rhop", "type", "table"):
        assert np.array_equal(getattr(g1, f), getattr(g2, f)), f
    d1, d2 = s1.scene_diagnostics(0), s2.scene_diagnostics(0)
    assert [d["n_fluid"] for d in d1] == [d["n_fluid"] for d in d2] == [c.json.water_ptc_num for c in b.cases]
    bad = h.table.copy()
    bad[0] = b.n  # original index outside [0, N)
    h.table[:] = bad
    with pytest.raises(z.ZsphError) as e:
        s2.upload(h)
    assert e.value.code == z.abi.E_ARG
    s1.close()
    s2.close()


def test_batch_rejects_single_scene_calls():
    b = uneven_batch(2, seed=9)
    s = z.Solver(b)
    s.upload(b.host)
    with pytest.raises(z.ZsphError):
        s.set_piston(0.01, 1.0)
    with pytest.raises(z.ZsphError):
        s.probe_pressure([[0.1, 0.0, 0.1]])
    with pytest.raises(z.ZsphError) as e:
        s.scene_diagnostics(0)  # not in cell order yet
    assert e.value.code == z.abi.E_STATE
    s.close()
