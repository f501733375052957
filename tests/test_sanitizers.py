"""CPU: host code under AddressSanitizer + UndefinedBehaviorSanitizer.

* The library's host sources (case generators, parameters, JSON / VTK /
  checkpoint readers and writers, rigid body) are built with the sanitizers
  together with tests/native/host_checks.cpp, which also feeds the readers
  truncated, bit-flipped and hostile files; any report aborts the run.
* The oracle is built the same way and steps a 2-D dam break, a 3-D dam break
  and a wave tank with a floating body and a piston, in a Python process with
  the ASan runtime preloaded.
Both builds use g++ only (no device code).
"""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "zsph_amd", "csrc")
SAN_ENV = {"ASAN_OPTIONS": "halt_on_error=1:abort_on_error=1", "UBSAN_OPTIONS": "halt_on_error=1:print_stacktrace=1"}


def runtime(name):
    p = subprocess.run(["g++", f"-print-file-name={name}"], capture_output=True, text=True).stdout.strip()
    return p if os.path.isabs(p) and os.path.exists(p) else None


def make(target_dir, target):
    r = subprocess.run(["make", "-C", target_dir, target], capture_output=True, text=True, timeout=900)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]


@pytest.mark.skipif(runtime("libasan.so") is None, reason="no ASan runtime for g++")
def test_host_code_under_sanitizers(tmp_path):
    make(CSRC, "host-asan")
    exe = os.path.join(CSRC, "_build", "asan", "host_checks")
    r = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, timeout=600,
                       env={**os.environ, **SAN_ENV})
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-6000:]
    assert "host checks: 0 failed" in r.stdout


ORACLE_RUN = """
import sys
sys.path.insert(0, {root!r})
import numpy as np
import oracle
from zsph_amd import cases
from tests.helpers import jitter
from tests.test_body import HI, LO, tank
assert oracle.LIB_PATH.endswith("_asan/liboracle.so"), oracle.LIB_PATH
for c in (jitter(cases.dambreak2d(dx=0.02), seed=3),
          jitter(cases.dambreak3d(dx=0.05, lx=0.6, wy=0.3, lz=0.4, cw=0.2, ch=0.2), seed=4)):
    o = oracle.Oracle(c.params)
    o.upload(c.host)
    o.step(2)
    assert np.isfinite(o.download(0).pos).all()
    o.close()
c = tank()
o = oracle.Oracle(c.params)
o.upload(c.host)
o.set_body(cases.box_body(LO, HI, 500.0))
o.set_piston(0.02, 1.5, 0.0, 0)
o.step(2)
o.close()
print("oracle ok")
"""


@pytest.mark.skipif(runtime("libasan.so") is None, reason="no ASan runtime for g++")
def test_oracle_under_sanitizers():
    make(os.path.join(ROOT, "oracle"), "asan")
    env = {**os.environ, **SAN_ENV,
           "LD_PRELOAD": runtime("libasan.so"),
           "ZSPH_ORACLE_LIB": os.path.join(ROOT, "oracle", "_asan", "liboracle.so"),
           # the interpreter and the HIP runtime keep allocations until exit
           "ASAN_OPTIONS": SAN_ENV["ASAN_OPTIONS"] + ":detect_leaks=0",
           "OMP_NUM_THREADS": "4"}
    r = subprocess.run([sys.executable, "-c", ORACLE_RUN.format(root=ROOT)], capture_output=True, text=True,
                       timeout=900, env=env, cwd=ROOT)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-6000:]
    assert "oracle ok" in r.stdout
