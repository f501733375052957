"""GPU: the stage clocks of the neighbour passes (zsph_clock_stage).

The first thread of the first block of each pass stamps the device wall
clock through the kernels' parameter block: the wall pass opens its stage,
the gradient pass closes it and opens its own, the continuity+momentum pass
closes that and opens its own, and the integrator closes it.  No clock
kernels run in the stream.

* Every half step closes each clocked stage once (launches = 2 x steps), in
  the captured graph and in direct launches, in skin mode and with the
  reference schedule.
* The clocked times are positive, and together stay below the wall-clock
  time of the steps.
* Clocks change nothing else: a run with the clocks on equals the same run
  with them off bit for bit.
* Only the clocked stages count (the continuity+momentum clock alone).
"""
import sys
import time

import numpy as np
import pytest

import zsph_amd as z
from zsph_amd import cases

pytestmark = pytest.mark.gpu
STAGES = (z.ST_WALL, z.ST_LDG, z.ST_GOVERNING)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def run(case, steps, clocks, graph=True, skin=0.0):
    s = z.Solver(case, device=0)
    s.upload(case.host)
    if skin:
        s.set_skin(skin)
    s.set_graph(graph)
    for st in clocks:
        s.clock_stage(st, True)
    s.step(2)  # warm-up (and the graph's capture)
    s.sync()
    for st in clocks:
        s.clock_read_stage(st)  # reset
    t0 = time.perf_counter()
    s.step(steps)
    s.sync()
    wall_ms = (time.perf_counter() - t0) * 1e3
    got = {st: s.clock_read_stage(st) for st in clocks}
    state = s.download(0)
    s.close()
    return got, wall_ms, state


@pytest.mark.parametrize("graph,skin", [(True, 0.0), (True, 0.1), (False, 0.1)])
def test_stage_clocks_count_every_half_step(graph, skin):
    case = cases.dambreak3d(dx=0.04, grid_size_factor=2.0)
    say(f"[clock] 3-D dam break N={case.n}, graph={graph}, skin={skin}")
    steps = 20
    got, wall_ms, st_on = run(case, steps, STAGES, graph=graph, skin=skin)
    for st, (ms, n) in got.items():
        say(f"  stage {st}: {ms:.4f} ms over {n} launches")
        assert n == 2 * steps
        assert ms > 0.0
    assert sum(ms for ms, _ in got.values()) < wall_ms
    _, _, st_off = run(case, steps, (), graph=graph, skin=skin)
    for f in ("pos", "vel", "rhop", "type", "table"):
        assert np.array_equal(getattr(st_on, f), getattr(st_off, f)), f


def test_only_the_clocked_stage_counts():
    case = cases.dambreak2d(dx=0.01)
    s = z.Solver(case, device=0)
    s.upload(case.host)
    s.clock_governing(True)
    s.step(3)
    s.sync()
    ms, n = s.clock_read()
    assert n == 6 and ms > 0.0
    s.step(4)
    s.sync()
    ms, n = s.clock_read()
    assert n == 8 and ms > 0.0
    s.clock_governing(False)
    s.step(2)
    s.sync()
    ms, n = s.clock_read()  # nothing counted while it was off
    assert n == 0 and ms == 0.0
    s.close()
