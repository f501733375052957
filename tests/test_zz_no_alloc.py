"""zsph_probe_pressure_noalloc (the probe entry point that allocates nothing)
against zsph_probe_pressure through the C ABI: two calls in a row, the second
with fewer points, bit for bit; more than ZSPH_PROBE_NOALLOC_MAX points is
refused.  Runs alone in a fresh process and after the other tests."""
import sys

import numpy as np
import pytest

import zsph_amd as z
from zsph_amd import cases

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def probe_points(c, n, seed):
    rng = np.random.default_rng(seed)
    lo = np.array([c.cpu.xmin, c.cpu.ymin, c.cpu.zmin])
    hi = np.array([c.cpu.xmax, c.cpu.ymax, c.cpu.zmax])
    return (lo + rng.random((n, 3)) * (hi - lo)).astype(np.float32)


def test_noalloc_probes_equal_the_allocating_entry_point():
    c = cases.dambreak3d(dx=0.02)
    say(f"[no-alloc probes] N={c.n}")
    s = z.Solver(c)
    s.upload(c.host)
    s.step(20)
    s.sync()
    for npts, seed in ((3000, 1), (17, 2)):  # the second call is smaller: stale staging must not leak in
        pts = probe_points(c, npts, seed)
        want = s.probe_pressure(pts, 0)
        got = s.probe_pressure_noalloc(pts, 0)
        assert got.shape == want.shape
        assert np.array_equal(got, want), (npts, np.abs(got - want).max())
        assert np.count_nonzero(want) > npts // 10  # many points lie in the water
        say(f"  {npts} points: identical, {np.count_nonzero(want)} inside the fluid")
    # the limit is enforced before any device work
    with pytest.raises(z.ZsphError):
        s.probe_pressure_noalloc(probe_points(c, z.abi.PROBE_NOALLOC_MAX + 1, 3), 0)
    assert s.probe_pressure_noalloc(np.zeros((0, 3), np.float32), 0).shape == (0,)
    # a second context on the same GPU shares the array: ordered calls stay correct
    s2 = z.Solver(c)
    s2.upload(c.host)
    s2.step(5)
    s2.sync()
    pts = probe_points(c, 200, 4)
    assert np.array_equal(s2.probe_pressure_noalloc(pts, 0), s2.probe_pressure(pts, 0))
    assert np.array_equal(s.probe_pressure_noalloc(pts, 0), s.probe_pressure(pts, 0))
    s2.close()
    s.close()
