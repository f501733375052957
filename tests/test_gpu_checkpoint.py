"""GPU: a run restarted from a checkpoint continues bit for bit.

For each kind of run -- 2-D and 3-D dam breaks, a wave tank with a piston
and a floating body (FP64 device state), a batch of scenes, the learned
closure -- a run of k1 + k2 steps is compared with one that writes a
checkpoint after k1 steps and continues in a fresh context restored from
it: every particle array, the time, the removal count and the body state
must be identical.  Also: restoring into the writing context rewinds it; a
checkpoint is refused by a context with other parameters, which keeps its
state; and the stand-alone driver's --checkpoint-step / --restart reproduce
the uninterrupted run's output frames byte for byte.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

import zsph_amd as z
from zsph_amd import A_POSR, A_TABLE, A_TPOSR, A_TVELP, A_TYPE, A_VELP, cases
from zsph_amd import _abi as abi
from zsph_amd import closure
from tests.helpers import jitter
from tests.test_body import HI, LO, subset_case, tank
from tests.test_closure_oracle import random_net

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def small3d():
    return jitter(cases.dambreak3d(dx=0.04, lx=0.8, wy=0.4, lz=0.5, cw=0.3, ch=0.3), seed=12345)


def closure_setup(c):
    """A random network normalised on the case's own features (margin 0.5)."""
    net = random_net(6)
    net.params()[-130:] *= 0.3
    s = z.Solver(c)
    s.upload(c.host)
    s.hash(0); s.sort_index(); s.sort_data(0, 1)
    f, _ = s.closure_sample(1)
    s.close()
    closure.set_normalisation(net, f[f.any(1)], margin=0.5)
    return lambda s: s.set_closure(net)


def body_piston_setup(s):
    s.set_body(cases.box_body(LO, HI, 500.0))
    s.set_piston(0.02, 1.5, 0.3, 0)


# name -> (case factory, set-up of a new context or None, k1, k2)
RUNS = {
    "dambreak2d": (lambda: jitter(cases.dambreak2d(dx=0.01), seed=7), None, 7, 9),
    "dambreak3d": (small3d, None, 5, 6),
    "body_piston": (tank, body_piston_setup, 6, 5),
    "batch": (lambda: cases.ensemble2d(n_scenes=5, seed=5), None, 6, 6),
    "closure": (small3d, "closure", 4, 5),
}


def state(s):
    st = {a: s.get(a, 0) for a in (A_POSR, A_VELP, A_TPOSR, A_TVELP, A_TYPE, A_TABLE)}
    st["killed"] = s.killed()
    if s.nscenes == 1:
        st["time"] = s.time()
    return st


def assert_same(a, b, what):
    assert a.keys() == b.keys()
    for k in a:
        assert np.array_equal(a[k], b[k]), f"{what}: {k} differs"


@pytest.mark.parametrize("name", list(RUNS))
def test_restart_continues_bit_for_bit(name, tmp_path):
    make, setup, k1, k2 = RUNS[name]
    c = make()
    if setup == "closure":
        setup = closure_setup(c)
    say(f"[checkpoint {name}] N={c.n} k1={k1} k2={k2}")
    path = str(tmp_path / f"{name}.bin")
    s = z.Solver(c)
    s.upload(c.host)
    if setup:
        setup(s)
    s.step(k1)
    s.sync()
    s.checkpoint(path, k1)
    at_ckpt = state(s)
    s.step(k2)
    s.sync()
    ref = state(s)
    body_ref = bytes(s.get_body()) if name == "body_piston" else None
    info = z.checkpoint_info(path)
    assert info["step"] == k1 and info["ptc_num"] == c.n and info["scenes"] == s.nscenes
    assert info["has_body"] == (name == "body_piston") and info["closure"] == (name == "closure")

    # a fresh context, nothing uploaded or set up: everything comes from the file
    r = z.Solver(c)
    assert r.restore(path) == k1
    assert_same(state(r), at_ckpt, f"{name} restored")
    r.step(k2)
    r.sync()
    assert_same(state(r), ref, f"{name} continued")
    if body_ref is not None:
        assert bytes(r.get_body()) == body_ref  # FP64 rigid state, byte for byte
    r.close()

    # the writing context rewinds to the checkpoint and replays the same steps
    s.restore(path)
    assert_same(state(s), at_ckpt, f"{name} rewound")
    s.step(k2)
    s.sync()
    assert_same(state(s), ref, f"{name} replayed")
    s.close()


def test_restore_rejects_other_parameters_and_keeps_state(tmp_path):
    c = jitter(cases.dambreak2d(dx=0.01), seed=7)
    path = str(tmp_path / "a.bin")
    s = z.Solver(c)
    s.upload(c.host)
    s.step(2)
    s.checkpoint(path, 2)
    s.close()
    # same particles and grid, other viscosity: only the parameter hash differs
    other = subset_case(c, np.ones(c.n, bool), alpha=c.json.alpha * 2.0)
    t = z.Solver(other)
    t.upload(other.host)
    before = state(t)
    with pytest.raises(z.ZsphError) as e:
        t.restore(path)
    assert e.value.code == abi.E_ARG
    assert_same(state(t), before, "refused restore")
    t.close()
    # other particle count
    u_case = jitter(cases.dambreak2d(dx=0.0125), seed=7)
    u = z.Solver(u_case)
    u.upload(u_case.host)
    with pytest.raises(z.ZsphError) as e:
        u.restore(path)
    assert e.value.code == abi.E_ARG
    u.close()


def test_driver_restart_reproduces_frames(tmp_path):
    exe = os.path.join(ROOT, "zsph_amd", "zsph_run")
    full, part = tmp_path / "full", tmp_path / "part"
    full.mkdir()
    part.mkdir()
    base = [exe, "--dambreak", "2", "--dx", "0.01", "--output-step", "10"]
    run = lambda args: subprocess.run(base + args, check=True, capture_output=True, text=True, timeout=120)
    say("[checkpoint driver] uninterrupted 40 steps")
    run(["--steps", "40", "--out", str(full)])
    say("[checkpoint driver] 20 steps with a checkpoint, then restart to 40")
    run(["--steps", "20", "--checkpoint-step", "20", "--out", str(part)])
    ck = part / "zsph_ckpt20.bin"
    assert ck.exists() and z.checkpoint_info(str(ck))["step"] == 20
    (part / "zsph10.vtk").unlink()
    out = run(["--steps", "40", "--restart", str(ck), "--out", str(part)]).stdout
    assert "restart from" in out and "at step 20" in out
    for k in (20, 30):
        a = (full / f"zsph{k}.vtk").read_bytes()
        b = (part / f"zsph{k}.vtk").read_bytes()
        assert a == b, f"frame {k} differs after restart"
    assert not (part / "zsph10.vtk").exists()  # the restarted run starts at step 20
