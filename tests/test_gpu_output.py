"""GPU: asynchronous output (zsph_download_async / zsph_frame_wait), the
solver-side half of the reference's output path (lib/gmemcpy.cu:3-36 feeding
io/mul_thread.cu:50-80).

* Frames started between steps equal the blocking zsph_download at the same
  point, bit for bit, with every frame slot in flight while the solver keeps
  stepping, and when a frame is waited for from another thread.
* The driver's frames with asynchronous output equal those of its blocking
  path (--sync-output 1), byte for byte.
"""
import os
import subprocess
import sys
import threading

import numpy as np
import pytest

import zsph_amd as z
from zsph_amd import cases
from tests.helpers import jitter

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FIELDS = ("pos", "vel", "rhop", "acc", "type", "table", "hash")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def same(a, b, what):
    for k in FIELDS:
        assert np.array_equal(getattr(a, k), getattr(b, k)), f"{what}: {k} differs"


def test_async_frames_equal_blocking_downloads():
    c = jitter(cases.dambreak3d(dx=0.02), seed=3)
    say(f"[async frames] N={c.n}")
    s = z.Solver(c)
    s.upload(c.host)
    refs, frames = [], []
    for k in range(z.abi.FRAME_SLOTS):  # every slot in flight
        s.step(4)
        refs.append(s.download(0))
        frames.append(s.download_async(0))
    s.step(10)  # the solver keeps going while the copies are in flight
    s.sync()
    # one frame waited for from another thread (the driver's writers do this)
    box = {}
    t = threading.Thread(target=lambda: box.setdefault("h", frames[0].wait()))
    t.start()
    t.join(60)
    assert "h" in box
    same(box["h"], refs[0], "frame 0 (other thread)")
    for k in range(1, len(frames)):
        same(frames[k].wait(), refs[k], f"frame {k}")
    # slots are reusable, and a field subset leaves the other host arrays untouched
    f = s.download_async(0, z.F_POS | z.F_TYPE)
    h = f.wait()
    ref = s.download(0)
    assert np.array_equal(h.pos, ref.pos) and np.array_equal(h.type, ref.type)
    assert not h.vel.any() and not h.acc.any()
    with pytest.raises(z.ZsphError):
        f.wait()
    s.close()


def test_driver_async_frames_equal_blocking_frames(tmp_path):
    exe = os.path.join(ROOT, "zsph_amd", "zsph_run")
    base = [exe, "--dambreak", "2", "--dx", "0.01", "--output-step", "10", "--steps", "40"]
    outs = {}
    for mode in ("async", "sync"):
        d = tmp_path / mode
        d.mkdir()
        extra = ["--sync-output", "1"] if mode == "sync" else []
        say(f"[driver output] {mode}")
        subprocess.run(base + ["--out", str(d)] + extra, check=True, capture_output=True, text=True, timeout=120)
        outs[mode] = d
    for k in (0, 10, 20, 30):
        a = (outs["async"] / f"zsph{k}.vtk").read_bytes()
        b = (outs["sync"] / f"zsph{k}.vtk").read_bytes()
        assert a == b, f"frame {k} differs between asynchronous and blocking output"
