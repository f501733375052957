"""Host side of the C ABI, CPU only: exported symbols, struct layouts,
parameter derivation (lib/init.cu:153-251), dt check (lib/check.cu:3-18),
case files (io/read_json.cu) and legacy VTK (io/write_vtk.cu, io/read_vkt.cu).
No GPU compute calls.
"""
import ctypes as C
import os
import re

import numpy as np
import pytest

import zsph_amd as z
from zsph_amd import _abi as abi
from zsph_amd import cases

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "zsph.h")).read()
    declared = set(re.findall(r"\b(zsph_[a-z0-9_]+)\s*\(", hdr))
    assert declared == set(abi.EXPORTS)
    L = C.CDLL(z.LIB_PATH)
    for name in declared:
        assert hasattr(L, name), name


def test_abi_struct_sizes_match():
    z.lib()  # raises on a mismatch
    assert z.lib().zsph_sizeof(2) == C.sizeof(abi.Params)


def test_set_cpu_param_reference_formulas():
    j = z.Case.defaults()
    j.dim, j.dx, j.h_factor, j.r_factor, j.eta_factor = 3, 0.01, 1.3, 2.0, 0.1
    j.cs_factor, j.g, j.rho0 = 10.0, -9.81, 1000.0
    j.xmin, j.xmax, j.ymin, j.ymax, j.zmin, j.zmax = 0.0, 1.0, 0.0, 0.5, 0.0, 0.6
    j.grid_layer_factor, j.grid_size_factor = 1.0, 1.0
    j.dt, j.ptc_num = 1e-5, 10
    cp = abi.CpuParam()
    assert z.lib().zsph_set_cpu_param(C.byref(cp), C.byref(j)) == abi.OK
    f = np.float32
    assert cp.m == f(f(f(f(1000.0) * f(0.01)) * f(0.01)) * f(0.01))          # rho0 dx^3 (init.cu:157)
    assert cp.h == f(f(1.3) * f(0.01))                                          # init.cu:160
    assert cp.cs == pytest.approx(10.0 * np.sqrt(9.81 * 0.6), rel=1e-6)          # init.cu:163
    assert cp.adh == pytest.approx(21.0 / (16.0 * np.pi * cp.h ** 3), rel=1e-6)   # init.cu:166
    gs = f(f(f(2.0) * f(1.3)) * f(0.01))
    assert cp.grid_size == gs                                                   # init.cu:177
    assert cp.grid_xmin == f(f(0.0) - gs)
    assert cp.grid_xdim == int(f(f(f(1.0) + gs) - f(f(0.0) - gs)) / gs)        # init.cu:185
    assert cp.grid_num == cp.grid_xdim * cp.grid_ydim * cp.grid_zdim
    assert cp.grid_hash_max == cp.grid_num - 1
    gp = abi.Params()
    z.lib().zsph_set_gpu_param(C.byref(gp), C.byref(cp))
    assert gp.half_dt == f(cp.dt / f(2.0))
    assert gp.h2 == f(cp.h * cp.h) and gp.cs2 == f(cp.cs * cp.cs)
    assert gp.delta_h_cs == f(f(cp.delta * cp.h) * cp.cs)                       # init.cu:233
    assert gp.h_cs_rho_m == f(f(f(cp.h * cp.cs) * cp.rho0) * cp.m)              # init.cu:234
    assert gp.rc2 == f(4.0) * gp.h2


def test_dt_check_returns_error_instead_of_exit():
    j = z.Case.defaults()
    j.dx, j.xmax, j.ymax, j.zmax, j.ptc_num = 0.01, 1.0, 1.0, 1.0, 1
    j.dt = 1.0  # far above 0.25 h / sqrt(|g| H)
    cp = abi.CpuParam()
    assert z.lib().zsph_set_cpu_param(C.byref(cp), C.byref(j)) == abi.E_DT
    j.dt = 1e-5
    assert z.lib().zsph_set_cpu_param(C.byref(cp), C.byref(j)) == abi.OK


def test_two_d_normalisation_and_grid():
    c = cases.dambreak2d()
    assert c.cpu.adh == pytest.approx(7.0 / (4.0 * np.pi * c.cpu.h ** 2), rel=1e-6)
    assert c.cpu.m == pytest.approx(1000.0 * 0.005 ** 2, rel=1e-6)
    assert c.cpu.grid_ydim == 1
    assert np.all(c.host.pos[:, 1] == 0.0)


def test_case_counts_match_configurations():
    c0 = cases.dambreak2d()
    assert (c0.json.water_ptc_num, c0.json.wall_ptc_num) == (4800, 1944)
    assert c0.cpu.cs == pytest.approx(10 * np.sqrt(2 * 9.81 * 0.3), rel=1e-6)
    assert c0.cpu.dt == pytest.approx(0.2 * c0.cpu.h / c0.cpu.cs, rel=1e-6)
    c1 = z.Case.dambreak(3, 0.008, 3.22, 1.0, 1.0, 1.2, 0.55, layers=3)
    assert c1.json.water_ptc_num == 1293750
    # hydrostatic start: p = rho0 |g| (H - z) in the fluid
    f = c1.host.type == z.FLUID
    H = 69 * 0.008
    np.testing.assert_allclose(c1.host.rhop[f, 1], 1000 * 9.81 * (H - c1.host.pos[f, 2]), rtol=1e-5, atol=1e-2)


def test_json_reader(tmp_path):
    p = tmp_path / "input.json"
    p.write_text('{"g": -9.81, "rho0": 998.0, "dx": 0.01, "h_factor": 1.3, "r_factor": 2, "eta_factor": 0.1,'
                 ' "cs_factor": 10, "delta": 0.1, "alpha": 0.02, "xmin": 0, "xmax": 1, "ymin": 0, "ymax": 1,'
                 ' "zmin": 0, "zmax": 1, "grid_layer_factor": 1, "grid_size_factor": 1, "dt": 1e-5,'
                 ' "start_step": 0, "current_step": 0, "end_step": 10, "output_step": 5, "ptc_num": 3,'
                 ' "water_ptc_num": 2, "air_ptc_num": 0, "wall_ptc_num": 1, "rigid_ptc_num": 0,'
                 ' "thread_num": 2, "gpu_id": 0, "input_path": "/in", "output_path": "/out",'
                 ' "git_hash": "abc", "gamma": 7, "diffusion": "MC", "extra": {"nested": [1, 2]}}')
    j = abi.Json()
    assert z.lib().zsph_read_json(str(p).encode(), C.byref(j)) == abi.OK
    assert j.rho0 == np.float32(998.0) and j.ptc_num == 3 and j.thread_num == 2
    assert j.input_path == b"/in" and j.output_path == b"/out" and j.git_hash == b"abc"
    assert j.gamma == 7.0 and j.diffusion == abi.DIFF_MC and j.dim == 3
    assert z.lib().zsph_read_json(b"/nonexistent.json", C.byref(j)) == abi.E_IO
    bad = tmp_path / "bad.json"
    bad.write_text('{"g": ')
    assert z.lib().zsph_read_json(str(bad).encode(), C.byref(j)) == abi.E_IO


@pytest.mark.parametrize("binary", [0, 1])
def test_vtk_roundtrip(tmp_path, binary):
    c = cases.dambreak2d(dx=0.02)
    h = c.host
    rng = np.random.default_rng(0)
    h.vel[:] = rng.normal(size=h.vel.shape).astype(np.float32)
    h.acc[:, :3] = rng.normal(size=(h.n, 3)).astype(np.float32)
    h.table[:] = rng.integers(-5, 5, h.n)
    path = str(tmp_path / "f.vtk").encode()
    st = h.state()
    assert z.lib().zsph_write_vtk(path, C.byref(st), h.n, abi.F_ALL, binary) == abi.OK
    n = C.c_int()
    assert z.lib().zsph_vtk_count(path, C.byref(n)) == abi.OK and n.value == h.n
    g = z.HostArrays(h.n)
    gs = g.state()
    assert z.lib().zsph_read_vtk(path, C.byref(gs), h.n, 1000.0) == abi.OK
    for k in ("pos", "vel", "rhop", "type", "table"):
        assert np.array_equal(getattr(g, k), getattr(h, k)), k
    assert np.array_equal(g.acc[:, :3], h.acc[:, :3])
    # particle count mismatch is an error (read_vkt.cu:14 asserts it)
    assert z.lib().zsph_read_vtk(path, C.byref(gs), h.n + 1, 1000.0) == abi.E_ARG


def test_vtk_reader_defaults_and_big_endian(tmp_path):
    # a file with only POINTS and type: rho defaults to rho0, p/vel to 0 (read_vkt.cu:261-267)
    n = 2
    pts = np.array([[0.1, 0.2, 0.3], [1.0, 2.0, 3.0]], ">f4")
    typ = np.array([1, 0], ">i4")
    body = (b"# vtk DataFile Version 3.0\nzsph data\nBINARY\nDATASET UNSTRUCTURED_GRID\nPOINTS 2 float\n"
            + pts.tobytes() + b"POINT_DATA 2\nSCALARS type int 1\nLOOKUP_TABLE default\n" + typ.tobytes())
    path = tmp_path / "in.vtk"
    path.write_bytes(body)
    g = z.HostArrays(n)
    gs = g.state()
    assert z.lib().zsph_read_vtk(str(path).encode(), C.byref(gs), n, 1000.0) == abi.OK
    np.testing.assert_array_equal(g.pos, pts.astype(np.float32))
    assert list(g.type) == [1, 0]
    assert np.all(g.rhop[:, 0] == 1000.0) and np.all(g.rhop[:, 1] == 0.0) and np.all(g.vel == 0.0)


def test_invalid_params_rejected_without_gpu():
    p = abi.Params()
    ctx = C.c_void_p()
    assert z.lib().zsph_create(C.byref(p), 0, C.byref(ctx)) == abi.E_ARG
    # sizes: an empty context and one at the addressing limit (2^28 - 64
    # particles) are refused before any device call, alone and in a batch
    good = cases.dambreak2d(dx=0.02).params
    for n in (0, -1, (1 << 28) - 64, 1 << 28, (1 << 31) - 1):
        q = abi.Params.from_buffer_copy(good)
        q.ptc_num = n
        assert z.lib().zsph_create(C.byref(q), 0, C.byref(ctx)) == abi.E_ARG, n
        scenes = (abi.Params * 2)(abi.Params.from_buffer_copy(good), q)
        assert z.lib().zsph_create_batch(scenes, 2, 0, C.byref(ctx)) == abi.E_ARG, n
        assert not ctx.value
    big = (abi.Params * 2)(abi.Params.from_buffer_copy(good), abi.Params.from_buffer_copy(good))
    big[0].ptc_num = big[1].ptc_num = (1 << 27) - 32
    assert z.lib().zsph_create_batch(big, 2, 0, C.byref(ctx)) == abi.E_ARG
    # grids: at most 2^29 cells, alone and summed over a batch
    q = abi.Params.from_buffer_copy(good)
    q.grid_xdim, q.grid_ydim, q.grid_zdim = 1024, 1024, 513
    q.grid_hash_max = 1024 * 1024 * 513 - 1
    assert z.lib().zsph_create(C.byref(q), 0, C.byref(ctx)) == abi.E_ARG
    half = abi.Params.from_buffer_copy(good)
    half.grid_xdim, half.grid_ydim, half.grid_zdim = 1024, 1024, 257
    half.grid_hash_max = 1024 * 1024 * 257 - 1
    scenes = (abi.Params * 2)(half, abi.Params.from_buffer_copy(half))
    assert z.lib().zsph_create_batch(scenes, 2, 0, C.byref(ctx)) == abi.E_ARG
