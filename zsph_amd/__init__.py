"""zsph_amd: MI355X (gfx950) delta-SPH step behind the reference's host interface.

Python host mirror of the reference driver's entry points
(single_gpu_sph_src/deltasph/main.cu:58-90, header/lib.cuh:7-36,
header/equation.cuh:8-17) over the C ABI of ``libzsph.so`` (include/zsph.h).
Every compute call runs the hand-written HIP kernels of the library; there is
no CPU fallback: without the built library or a GPU the calls raise.
"""
import ctypes as C
import os

import numpy as np

from . import _abi as abi
from ._abi import (A_ACC, A_CELL_START, A_GD, A_INDEX, A_KEY, A_POSR, A_TABLE, A_TPOSR, A_TVELP,
                   A_TYPE, A_VELP, BODY, DIFF_ANTUONO, DIFF_MC, F_ACC, F_ALL, F_HASH, F_POS, F_RHOP,
                   F_TABLE, F_TYPE, F_VEL, FLUID, KILLED, PISTON, WALL, WALK_LIST, WALK_MASK, WALK_HITS, WALK_SKIN,
                   ST_GOVERNING, ST_LDG, ST_WALL)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libzsph.so")
_LIB = None


def lib():
    """The loaded libzsph.so (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        _LIB = abi.load(LIB_PATH)
    return _LIB


class ZsphError(RuntimeError):
    def __init__(self, code, msg=""):
        super().__init__(f"zsph error {code} ({abi.ERRORS.get(code, 'unknown')}){': ' + msg if msg else ''}")
        self.code = code


def _ptr(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype)) if a is not None else None


class HostArrays:
    """Host particle arrays in the reference's cpu_output_t layout (data_structure.cuh:146-156)."""

    def __init__(self, n):
        self.n = n
        self.pos = np.zeros((n, 3), np.float32)
        self.vel = np.zeros((n, 3), np.float32)
        self.rhop = np.zeros((n, 2), np.float32)
        self.acc = np.zeros((n, 4), np.float32)
        self.type = np.zeros(n, np.int32)
        self.table = np.zeros(n, np.int32)
        self.hash = np.zeros(n, np.int32)

    def state(self):
        return abi.HostState(_ptr(self.pos, C.c_float), _ptr(self.vel, C.c_float), _ptr(self.rhop, C.c_float),
                             _ptr(self.acc, C.c_float), _ptr(self.type, C.c_int), _ptr(self.table, C.c_int),
                             _ptr(self.hash, C.c_int))

    def copy(self):
        h = HostArrays(self.n)
        for k in ("pos", "vel", "rhop", "acc", "type", "table", "hash"):
            getattr(h, k)[...] = getattr(self, k)
        return h


class Case:
    """Case parameters (cpu_json_t -> cpu_param_t -> gpu_param_t, lib/init.cu:153-251) and initial particles."""

    def __init__(self, json, host):
        self.json = json
        self.host = host
        self.cpu = abi.CpuParam()
        self.params = abi.Params()
        self.derive()

    def derive(self):
        rc = lib().zsph_set_cpu_param(C.byref(self.cpu), C.byref(self.json))
        if rc != abi.OK:
            raise ZsphError(rc, "zsph_set_cpu_param")
        lib().zsph_set_gpu_param(C.byref(self.params), C.byref(self.cpu))
        return self

    @property
    def n(self):
        return self.host.n

    @staticmethod
    def defaults():
        j = abi.Json()
        lib().zsph_json_defaults(C.byref(j))
        return j

    @classmethod
    def dambreak(cls, dim, dx, lx, wy, lz, cw, ch, layers=3, wall_height=0.0, gamma=7.0, cfl=0.2,
                 diffusion=DIFF_ANTUONO, **overrides):
        """Lattice dam break (zsph_case_dambreak).  Extra keyword arguments set zsph_json fields."""
        j = cls.defaults()
        j.gamma, j.cfl, j.diffusion = gamma, cfl, diffusion
        for k, v in overrides.items():
            setattr(j, k, v)
        d = abi.DamBreak(dim, dx, lx, wy, lz, cw, ch, layers, wall_height)
        n = C.c_int(0)
        rc = lib().zsph_case_dambreak(C.byref(d), C.byref(j), None, C.byref(n))
        if rc != abi.OK:
            raise ZsphError(rc, "zsph_case_dambreak")
        host = HostArrays(n.value)
        st = host.state()
        rc = lib().zsph_case_dambreak(C.byref(d), C.byref(j), C.byref(st), C.byref(n))
        if rc != abi.OK:
            raise ZsphError(rc, "zsph_case_dambreak")
        return cls(j, host)

    @classmethod
    def wavetank(cls, dx, lx, wy, lz, depth, box_lo=(0.0, 0.0, 0.0), box_hi=(0.0, 0.0, 0.0), layers=3,
                 piston_travel=0.0, gamma=7.0, cfl=0.2, diffusion=DIFF_ANTUONO, **overrides):
        """Wave tank with a piston and a floating box (zsph_case_wavetank).  box_hi <= box_lo: no box.
        The box is snapped to lattice faces; ``case.box`` holds the snapped (lo, hi)."""
        j = cls.defaults()
        j.gamma, j.cfl, j.diffusion = gamma, cfl, diffusion
        for k, v in overrides.items():
            setattr(j, k, v)
        w = abi.WaveTank(dx, lx, wy, lz, depth, (C.c_float * 3)(*box_lo), (C.c_float * 3)(*box_hi), layers,
                         piston_travel)
        n = C.c_int(0)
        rc = lib().zsph_case_wavetank(C.byref(w), C.byref(j), None, C.byref(n))
        if rc != abi.OK:
            raise ZsphError(rc, "zsph_case_wavetank")
        host = HostArrays(n.value)
        st = host.state()
        rc = lib().zsph_case_wavetank(C.byref(w), C.byref(j), C.byref(st), C.byref(n))
        if rc != abi.OK:
            raise ZsphError(rc, "zsph_case_wavetank")
        case = cls(j, host)
        case.box = (tuple(w.box_lo[:]), tuple(w.box_hi[:]))  # snapped to the lattice
        return case

    @classmethod
    def from_json(cls, path):
        """Case file + <input_path>/input.vtk, like the reference's main (deltasph/main.cu:13-20)."""
        j = abi.Json()
        rc = lib().zsph_read_json(path.encode(), C.byref(j))
        if rc != abi.OK:
            raise ZsphError(rc, f"zsph_read_json({path})")
        host = HostArrays(j.ptc_num)
        vtk = os.path.join(j.input_path.decode(), "input.vtk")
        st = host.state()
        rc = lib().zsph_read_vtk(vtk.encode(), C.byref(st), j.ptc_num, j.rho0)
        if rc != abi.OK:
            raise ZsphError(rc, f"zsph_read_vtk({vtk})")
        return cls(j, host)


_ARRAY_SHAPES = {A_POSR: (4, np.float32), A_VELP: (4, np.float32), A_TPOSR: (4, np.float32),
                 A_TVELP: (4, np.float32), A_TYPE: (1, np.int32), A_TABLE: (1, np.int32),
                 A_KEY: (1, np.uint32), A_INDEX: (1, np.int32), A_ACC: (4, np.float32),
                 A_GD: (4, np.float32)}


class Batch:
    """Independent scenes advanced together in one context (zsph_create_batch).

    Particles are concatenated scene by scene; ``offsets[k]`` is scene k's
    first particle in that upload order (``table + index`` of a downloaded
    particle gives it back)."""

    def __init__(self, cases):
        self.cases = list(cases)
        if not self.cases:
            raise ValueError("a batch needs at least one scene")
        self.offsets = np.cumsum([0] + [c.n for c in self.cases]).astype(np.int64)
        self.host = HostArrays(int(self.offsets[-1]))
        for c, o in zip(self.cases, self.offsets):
            for k in ("pos", "vel", "rhop", "acc", "type", "table", "hash"):
                getattr(self.host, k)[o:o + c.n] = getattr(c.host, k)
        self.params = (abi.Params * len(self.cases))(*[c.params for c in self.cases])
        self.grid_num = sum(c.params.grid_xdim * c.params.grid_ydim * c.params.grid_zdim for c in self.cases)

    @property
    def n(self):
        return self.host.n

    @property
    def nscenes(self):
        return len(self.cases)

    def scene_of(self, original_index):
        """Scene of each original particle index."""
        return np.searchsorted(self.offsets, original_index, side="right") - 1


def checkpoint_info(path, verify=True):
    """Header of checkpoint `path` as a dict (host only; verify: check the payload checksum)."""
    info = abi.CheckpointInfo()
    rc = lib().zsph_checkpoint_inspect(os.fsencode(path), 1 if verify else 0, C.byref(info))
    if rc != abi.OK:
        raise ZsphError(rc, f"zsph_checkpoint_inspect({path})")
    d = {k: getattr(info, k) for k, _ in abi.CheckpointInfo._fields_}
    d["grid"] = tuple(info.grid)
    return d


class Frame:
    """An asynchronous download in flight (zsph_download_async); wait() once."""

    def __init__(self, handle, n):
        self._h, self.n = handle, n

    def wait(self):
        if self._h is None:
            raise ZsphError(abi.E_STATE, "frame already waited for")
        h = HostArrays(self.n)
        st = h.state()
        rc = lib().zsph_frame_wait(self._h, C.byref(st))
        self._h = None
        if rc != abi.OK:
            raise ZsphError(rc, "zsph_frame_wait")
        return h


class Solver:
    """One device context (zsph_ctx): the reference's device buffers and kernels.
    `case` is a Case (one scene) or a Batch (independent scenes in one context)."""

    def __init__(self, case, device=0):
        self.case = case
        self._ctx = C.c_void_p()
        if isinstance(case, Batch):
            self.params = case.params
            self.n = case.n
            self.grid_num = case.grid_num
            self.nscenes = case.nscenes
            rc = lib().zsph_create_batch(case.params, case.nscenes, device, C.byref(self._ctx))
        else:
            self.params = case.params if isinstance(case, Case) else case
            self.n = self.params.ptc_num
            self.grid_num = self.params.grid_xdim * self.params.grid_ydim * self.params.grid_zdim
            self.nscenes = 1
            rc = lib().zsph_create(C.byref(self.params), device, C.byref(self._ctx))
        if rc != abi.OK:
            raise ZsphError(rc, "zsph_create")

    def _chk(self, rc, what):
        if rc != abi.OK:
            raise ZsphError(rc, f"{what}: {lib().zsph_last_error(self._ctx).decode()}")

    def close(self):
        if self._ctx:
            lib().zsph_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # lib/init.cu:136-151 / lib/gmemcpy.cu:3-36
    def upload(self, host=None, buf=0):
        host = host if host is not None else self.case.host
        st = host.state()
        self._chk(lib().zsph_upload(self._ctx, buf, C.byref(st)), "zsph_upload")

    def download(self, buf=0, fields=F_ALL):
        h = HostArrays(self.n)
        st = h.state()
        self._chk(lib().zsph_download(self._ctx, buf, C.byref(st), fields), "zsph_download")
        return h

    def download_async(self, buf=0, fields=F_ALL):
        """Start an asynchronous download of buffer `buf` (zsph_download_async): returns a
        Frame whose wait() gives the HostArrays, equal to download() at this point of the
        stream.  At most abi.FRAME_SLOTS frames may be in flight per context."""
        h = C.c_void_p()
        self._chk(lib().zsph_download_async(self._ctx, buf, fields, C.byref(h)), "zsph_download_async")
        return Frame(h, self.n)

    # the reference's per-stage entry points
    def hash(self, buf=0):
        self._chk(lib().zsph_hash(self._ctx, buf), "zsph_hash")

    def sort_index(self):
        self._chk(lib().zsph_sort_index(self._ctx), "zsph_sort_index")

    def sort_data(self, src=0, dst=1):
        self._chk(lib().zsph_sort_data(self._ctx, src, dst), "zsph_sort_data")

    def boundary(self, buf=1):
        self._chk(lib().zsph_boundary(self._ctx, buf), "zsph_boundary")

    def governing(self, buf=1):
        self._chk(lib().zsph_governing(self._ctx, buf), "zsph_governing")

    def prediction(self, buf=1):
        self._chk(lib().zsph_prediction(self._ctx, buf), "zsph_prediction")

    def correction(self, buf=0):
        self._chk(lib().zsph_correction(self._ctx, buf), "zsph_correction")

    def step(self, nsteps=1):
        self._chk(lib().zsph_step(self._ctx, nsteps), "zsph_step")

    def sync(self):
        self._chk(lib().zsph_sync(self._ctx), "zsph_sync")

    # raw arrays
    def get(self, which, buf=0):
        if which == A_CELL_START:
            out = np.empty(2 * (self.grid_num + 1), np.int32)
        else:
            w, dt = _ARRAY_SHAPES[which]
            out = np.empty((self.n, w) if w > 1 else self.n, dt)
        self._chk(lib().zsph_array_get(self._ctx, which, buf, out.ctypes.data, out.nbytes), "zsph_array_get")
        return out

    def set(self, which, arr, buf=0):
        w, dt = _ARRAY_SHAPES[which]
        a = np.ascontiguousarray(arr, dtype=dt)
        self._chk(lib().zsph_array_set(self._ctx, which, buf, a.ctypes.data, a.nbytes), "zsph_array_set")

    def neighbor_stats(self, buf=1):
        cnt = np.empty(self.n, np.int32)
        dig = np.empty(self.n, np.uint64)
        self._chk(lib().zsph_neighbor_stats(self._ctx, buf, _ptr(cnt, C.c_int), _ptr(dig, C.c_uint64)),
                  "zsph_neighbor_stats")
        return cnt, dig

    def set_tuning(self, walk, radix_bits, block=None, sort_tile=None, skin=None, tile_rows=None):
        """Neighbour walk, sort digit width, neighbour-pass block size, sort tile, zsph_step's
        neighbour skin in h and skin mode's wave order (cell rows per unit, 0: cell order);
        None keeps the current block size / tile / skin / wave order."""
        cur = self.get_tuning()
        t = abi.Tuning(walk, radix_bits, cur["block"] if block is None else block,
                       cur["sort_tile"] if sort_tile is None else sort_tile,
                       cur["skin"] if skin is None else skin,
                       cur["tile_rows"] if tile_rows is None else tile_rows)
        self._chk(lib().zsph_set_tuning(self._ctx, C.byref(t)), "zsph_set_tuning")

    def set_skin(self, skin):
        """zsph_step's neighbour skin in units of h (0: rebuild every half step, the reference schedule)."""
        t = self.get_tuning()
        self.set_tuning(t["walk"], t["radix_bits"], skin=skin)

    def get_tuning(self):
        t = abi.Tuning()
        self._chk(lib().zsph_get_tuning(self._ctx, C.byref(t)), "zsph_get_tuning")
        return {"walk": t.walk, "radix_bits": t.radix_bits, "block": t.block, "sort_tile": t.sort_tile,
                "skin": round(t.skin, 6), "tile_rows": t.tile_rows}

    def autotune(self, steps=5):
        """Time tuning candidates from the current buffer-0 state (restored afterwards); keep the fastest."""
        t = abi.Tuning()
        ms = C.c_float(0.0)
        self._chk(lib().zsph_autotune(self._ctx, steps, C.byref(t), C.byref(ms)), "zsph_autotune")
        return {"walk": t.walk, "radix_bits": t.radix_bits, "block": t.block, "sort_tile": t.sort_tile,
                "skin": round(t.skin, 6), "tile_rows": t.tile_rows, "ms_per_step": ms.value / steps}

    def rebuilds(self):
        """Neighbour-structure rebuilds done by zsph_step in skin mode since creation."""
        n = C.c_longlong(0)
        self._chk(lib().zsph_rebuilds(self._ctx, C.byref(n)), "zsph_rebuilds")
        return n.value

    def time_steps(self, nsteps):
        """Milliseconds of `nsteps` graph-replayed steps, between two HIP events on the solver's stream."""
        ms = C.c_float(0.0)
        self._chk(lib().zsph_time_steps(self._ctx, nsteps, C.byref(ms)), "zsph_time_steps")
        return ms.value

    def profile_steps(self, nsteps):
        """Per-stage milliseconds (dict over abi.STAGES) of `nsteps` steps issued stage by stage."""
        out = (C.c_float * len(abi.STAGES))()
        self._chk(lib().zsph_profile_steps(self._ctx, nsteps, out), "zsph_profile_steps")
        return dict(zip(abi.STAGES, list(out)))

    def diagnostics(self, buf=0, front_zmax=float("inf")):
        d = abi.Diag()
        self._chk(lib().zsph_diagnostics(self._ctx, buf, front_zmax, C.byref(d)), "zsph_diagnostics")
        return d.as_dict()

    def scene_diagnostics(self, buf=0, front_zmax=float("inf")):
        """Per-scene diagnostics (list of dicts); buffer `buf` must be in cell order."""
        d = (abi.Diag * self.nscenes)()
        self._chk(lib().zsph_scene_diagnostics(self._ctx, buf, front_zmax, d), "zsph_scene_diagnostics")
        return [x.as_dict() for x in d]

    def probe_pressure(self, points, buf=0):
        pts = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
        out = np.zeros(len(pts), np.float32)
        self._chk(lib().zsph_probe_pressure(self._ctx, buf, _ptr(pts, C.c_float), len(pts), _ptr(out, C.c_float)),
                  "zsph_probe_pressure")
        return out

    def probe_pressure_noalloc(self, points, buf=0):
        """probe_pressure staged in the library's fixed device array: no device allocation,
        at most abi.PROBE_NOALLOC_MAX points; calls overlapping on one GPU race."""
        pts = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
        out = np.zeros(len(pts), np.float32)
        self._chk(lib().zsph_probe_pressure_noalloc(self._ctx, buf, _ptr(pts, C.c_float), len(pts),
                                                    _ptr(out, C.c_float)), "zsph_probe_pressure_noalloc")
        return out

    def killed(self):
        k = C.c_int(0)
        self._chk(lib().zsph_killed(self._ctx, C.byref(k)), "zsph_killed")
        return k.value

    # ---- checkpoint / restart (io/read_json.cu:32-33 start_step/current_step) --
    def checkpoint(self, path, step):
        """Write the state between steps to `path` (step = the loop index to continue from)."""
        self._chk(lib().zsph_checkpoint_write(self._ctx, os.fsencode(path), int(step)), "zsph_checkpoint_write")

    def restore(self, path):
        """Take the state of checkpoint `path` (same parameters); returns its step."""
        step = C.c_longlong(0)
        self._chk(lib().zsph_checkpoint_read(self._ctx, os.fsencode(path), C.byref(step)), "zsph_checkpoint_read")
        return step.value

    # ---- learned closure ------------------------------------------------------
    def set_closure(self, net):
        """Run the step with the closure network `net` (abi.Closure); None: plain terms."""
        self._chk(lib().zsph_set_closure(self._ctx, C.byref(net) if net is not None else None), "zsph_set_closure")

    def closure_sample(self, buf=1):
        """(features [N, NF], terms [N, 8]) of buffer `buf` (in cell order), rows in that buffer's order."""
        f = np.zeros((self.n, abi.NF), np.float32)
        t = np.zeros((self.n, 8), np.float32)
        self._chk(lib().zsph_closure_sample(self._ctx, buf, _ptr(f, C.c_float), _ptr(t, C.c_float)),
                  "zsph_closure_sample")
        return f, t

    def set_graph(self, on=True):
        """Replay the captured step graph (default) or launch each step's kernels directly."""
        self._chk(lib().zsph_set_graph(self._ctx, 1 if on else 0), "zsph_set_graph")

    def clock_governing(self, on=True):
        """Device clock around every continuity+momentum launch (also inside the captured step)."""
        self._chk(lib().zsph_clock_governing(self._ctx, 1 if on else 0), "zsph_clock_governing")

    def clock_read(self):
        """(milliseconds, launches) of the clocked kernel since the last read; resets."""
        ms = C.c_double(0.0)
        n = C.c_longlong(0)
        self._chk(lib().zsph_clock_read(self._ctx, C.byref(ms), C.byref(n)), "zsph_clock_read")
        return ms.value, n.value

    def clock_stage(self, stage, on=True):
        """Device clock around every launch of one neighbour pass (ST_WALL, ST_LDG, ST_GOVERNING)."""
        self._chk(lib().zsph_clock_stage(self._ctx, stage, 1 if on else 0), "zsph_clock_stage")

    def clock_read_stage(self, stage):
        """(milliseconds, launches) of that pass since the last read; resets."""
        ms = C.c_double(0.0)
        n = C.c_longlong(0)
        self._chk(lib().zsph_clock_read_stage(self._ctx, stage, C.byref(ms), C.byref(n)), "zsph_clock_read_stage")
        return ms.value, n.value

    def filter_fields(self, rows, radius, hc, buf=1):
        """Shepard-filtered (weight sum, rho, v, a, drho/dt) around particles `rows` of buffer `buf`."""
        rows = np.ascontiguousarray(rows, np.int32)
        out = np.zeros((len(rows), 9), np.float32)
        self._chk(lib().zsph_filter_fields(self._ctx, buf, radius, hc, _ptr(rows, C.c_int), len(rows),
                                           _ptr(out, C.c_float)), "zsph_filter_fields")
        return out

    # ---- moving dummies -------------------------------------------------------
    def set_body(self, rigid):
        """Make the type-2 particles of buffer 0 one rigid body with state `rigid` (abi.Rigid)."""
        self._chk(lib().zsph_set_body(self._ctx, C.byref(rigid)), "zsph_set_body")

    def get_body(self):
        r = abi.Rigid()
        self._chk(lib().zsph_get_body(self._ctx, C.byref(r)), "zsph_get_body")
        return r

    def set_piston(self, amplitude, frequency, phase=0.0, axis=0):
        p = abi.Piston(amplitude, frequency, phase, axis)
        self._chk(lib().zsph_set_piston(self._ctx, C.byref(p)), "zsph_set_piston")

    def set_time(self, t):
        self._chk(lib().zsph_set_time(self._ctx, t), "zsph_set_time")

    def time(self):
        t = C.c_double(0.0)
        self._chk(lib().zsph_get_time(self._ctx, C.byref(t)), "zsph_get_time")
        return t.value

    def body_reduce(self, buf=1):
        """Fluid force and torque on the body from the last governing pass on `buf`."""
        F = (C.c_double * 3)()
        T = (C.c_double * 3)()
        self._chk(lib().zsph_body_reduce(self._ctx, buf, F, T), "zsph_body_reduce")
        return np.array(F[:]), np.array(T[:])


def rigid_state(c, mass, inertia, vel=(0.0, 0.0, 0.0), omega=(0.0, 0.0, 0.0), q=(1.0, 0.0, 0.0, 0.0)):
    """An abi.Rigid with centre c, mass, 3x3 inertia about c (body frame), velocities and orientation."""
    r = abi.Rigid()
    r.c[:] = [float(x) for x in c]
    r.vel[:] = [float(x) for x in vel]
    r.omega[:] = [float(x) for x in omega]
    r.q[:] = [float(x) for x in q]
    r.I[:] = [float(x) for x in np.asarray(inertia, np.float64).reshape(9)]
    r.mass = float(mass)
    return r


def rigid_dict(r):
    return {k: np.array(getattr(r, k)[:]) for k in ("c", "vel", "omega", "q", "F", "T", "acc", "alpha", "I")} | \
        {"mass": r.mass}
