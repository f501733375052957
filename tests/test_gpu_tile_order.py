"""GPU: skin mode's wave order (zsph_tuning.tile_rows).  The gradient and
continuity+momentum passes take their waves in units of tile_rows cell rows
of one plane, tile by tile (k_wave_table), instead of in cell order.  A
particle's arithmetic and neighbour order do not depend on which wave runs
it, so every order gives the same state bit for bit:

* 3-D dam breaks over 150 steps with rebuilds, for every walk (mask, hit
  masks, lists, recorded skin walk), at 0 (cell order), 1, 7 and 16 rows per
  unit;
* graph replay and direct launches with the tiled order;
* a wave tank with a floating body and a piston (moving dummies, body
  coupling), cell order against 16 rows;
* the learned closure's passes, cell order against 16 rows;
* the tuning's range and one-scene checks;
* the stand-alone driver's frames under --skin with --tile-rows 16 and 0.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

import zsph_amd as z
from zsph_amd import cases
from tests.helpers import jitter

pytestmark = pytest.mark.gpu
SKIN = 0.2
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def dam():
    return jitter(cases.dambreak3d(dx=0.03, lx=0.9, wy=0.5, lz=0.6, cw=0.4, ch=0.35, grid_size_factor=2.0),
                  seed=99, amp=0.3)


def state(s):
    h = s.download(0)
    return [h.pos, h.vel, h.rhop, h.acc, h.table]


def run(c, walk, rows, steps, graph=True, setup=None):
    s = z.Solver(c)
    s.upload(c.host)
    t = s.get_tuning()
    s.set_tuning(walk, t["radix_bits"], skin=SKIN, tile_rows=rows)
    assert s.get_tuning()["tile_rows"] == rows
    s.set_graph(graph)
    if setup:
        setup(s)
    s.step(steps)
    s.sync()
    out = state(s), s.rebuilds(), s
    return out


@pytest.mark.parametrize("walk", [z.WALK_MASK, z.WALK_HITS, z.WALK_LIST, z.WALK_SKIN])
def test_wave_orders_give_identical_states(walk):
    c = dam()
    ref, nreb, s = run(c, walk, 0, 150)
    s.close()
    say(f"  walk {walk}: {nreb} rebuilds in 300 half steps, N = {c.n}")
    assert 1 < nreb < 300
    for rows in (1, 7, 16):
        got, n2, s = run(c, walk, rows, 150)
        s.close()
        assert n2 == nreb
        for a, b in zip(ref, got):
            assert np.array_equal(a, b), f"tile_rows {rows} differs from cell order"


def test_tiled_order_graph_equals_direct():
    c = dam()
    a, _, s = run(c, z.WALK_HITS, 16, 40, graph=True)
    s.close()
    b, _, s = run(c, z.WALK_HITS, 16, 40, graph=False)
    s.close()
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_tiled_order_with_body_and_piston():
    c = cases.wavetank3d(dx=0.04, depth=0.6)
    body = cases.box_body(*c.box, cases.BOX_DENSITY)

    def setup(s):
        s.set_body(body)
        s.set_piston(cases.PISTON_AMPLITUDE, cases.PISTON_FREQUENCY, 0.0, 0)

    res = []
    for rows in (0, 16):
        st, nreb, s = run(c, z.WALK_HITS, rows, 120, setup=setup)
        b = s.get_body()
        res.append((st, np.array(b.c[:]), np.array(b.q[:]), nreb))
        s.close()
    assert res[0][3] == res[1][3]
    for x, y in zip(res[0][0], res[1][0]):
        assert np.array_equal(x, y)
    assert np.array_equal(res[0][1], res[1][1]) and np.array_equal(res[0][2], res[1][2])


def test_tiled_order_with_learned_closure():
    # the closure's feature-writing gradient pass and scaled continuity+momentum
    # pass take their waves from the same table
    import os
    from zsph_amd import closure
    net = closure.load_network(os.path.join(os.path.dirname(closure.__file__), "data", "closure_dambreak3d.npz"))
    c = dam()
    res = []
    for rows in (0, 16):
        st, nreb, s = run(c, z.WALK_HITS, rows, 60, setup=lambda x: x.set_closure(net))
        res.append((st, nreb))
        s.close()
    assert res[0][1] == res[1][1]
    for x, y in zip(res[0][0], res[1][0]):
        assert np.array_equal(x, y)


def test_tile_rows_tuning_checks():
    c = dam()
    s = z.Solver(c)
    s.upload(c.host)
    assert s.get_tuning()["tile_rows"] == 16  # the default of a one-scene context
    t = s.get_tuning()
    with pytest.raises(z.ZsphError):
        s.set_tuning(t["walk"], t["radix_bits"], skin=SKIN, tile_rows=5000)
    s.set_tuning(t["walk"], t["radix_bits"], skin=SKIN, tile_rows=-1)  # -1 keeps the current order
    assert s.get_tuning()["tile_rows"] == 16
    s.close()
    b = z.Batch([cases.dambreak2d(), cases.dambreak2d()])
    s = z.Solver(b)
    s.upload(b.host)
    assert s.get_tuning()["tile_rows"] == 0
    t = s.get_tuning()
    with pytest.raises(z.ZsphError):
        s.set_tuning(t["walk"], t["radix_bits"], skin=SKIN, tile_rows=16)
    s.close()


def test_driver_tile_rows_frames_identical(tmp_path):
    exe = os.path.join(ROOT, "zsph_amd", "zsph_run")
    base = [exe, "--dambreak", "3", "--dx", "0.03", "--output-step", "20", "--steps", "60", "--skin", str(SKIN)]
    outs = {}
    for rows in (0, 16):
        d = tmp_path / f"r{rows}"
        d.mkdir()
        say(f"[driver tile rows] {rows}")
        subprocess.run(base + ["--tile-rows", str(rows), "--out", str(d)], check=True, capture_output=True,
                       text=True, timeout=120)
        outs[rows] = d
    for k in (0, 20, 40):
        a = (outs[0] / f"zsph{k}.vtk").read_bytes()
        b = (outs[16] / f"zsph{k}.vtk").read_bytes()
        assert a == b, f"frame {k} differs between cell order and 16-row units"
    with pytest.raises(subprocess.CalledProcessError):  # an out-of-range order is refused before any step
        subprocess.run(base + ["--tile-rows", "5000", "--out", str(tmp_path)], check=True, capture_output=True,
                       text=True, timeout=120)
