"""GPU: every neighbour walk gives the same bits with moving dummies and with the closure.

The walks (mask, hit masks recorded by the gradient pass and replayed by the
continuity+momentum pass, neighbour lists) visit the same neighbours in the
same order, so whole steps agree bit for bit; the plain dam breaks are checked
in test_gpu_parity.py.  Here: a floating body and a piston (the FP64 body
state included), and the learned closure (its gradient pass also computes
the network's features).
"""
import sys

import numpy as np
import pytest

import zsph_amd as z
from zsph_amd import A_ACC, A_POSR, A_TABLE, A_VELP, cases, closure
from tests.helpers import jitter
from tests.test_body import HI, LO, tank
from tests.test_closure_oracle import random_net

pytestmark = pytest.mark.gpu

WALKS = (z.WALK_MASK, z.WALK_HITS, z.WALK_LIST)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def run(case, walk, setup, steps):
    s = z.Solver(case)
    s.upload(case.host)
    s.set_tuning(walk, 8)
    setup(s)
    s.step(steps)
    s.sync()
    out = [s.get(a, 0) for a in (A_POSR, A_VELP, A_TABLE)] + [s.get(A_ACC)]
    return s, out


def assert_walks_agree(case, setup, steps, extra=None):
    ref = None
    for walk in WALKS:
        s, got = run(case, walk, setup, steps)
        if extra:
            got.append(extra(s))
        s.close()
        if ref is None:
            ref = got
            continue
        for a, b in zip(ref, got):
            assert np.array_equal(np.asarray(a), np.asarray(b)), f"walk {walk} differs"


def test_walks_agree_with_body_and_piston():
    c = tank()
    say(f"[walks body+piston] N={c.n}")

    def setup(s):
        s.set_body(cases.box_body(LO, HI, 500.0))
        s.set_piston(0.02, 1.5, 0.3, 0)

    assert_walks_agree(c, setup, 6, extra=lambda s: np.frombuffer(bytes(s.get_body()), np.uint8))


def test_walks_agree_with_closure():
    c = jitter(cases.dambreak3d(dx=0.04, lx=0.8, wy=0.4, lz=0.5, cw=0.3, ch=0.3), seed=12345)
    net = random_net(6)
    net.params()[-130:] *= 0.3
    s = z.Solver(c)
    s.upload(c.host)
    s.hash(0); s.sort_index(); s.sort_data(0, 1)
    f, _ = s.closure_sample(1)
    s.close()
    closure.set_normalisation(net, f[f.any(1)], margin=0.5)
    say(f"[walks closure] N={c.n}")
    assert_walks_agree(c, lambda s: s.set_closure(net), 4)
