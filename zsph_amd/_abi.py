"""ctypes mirror of include/zsph.h (the C ABI of libzsph.so).

Struct layouts are checked against the library's own ``zsph_sizeof`` at
load time, so a header change that is not mirrored here fails loudly.
"""
import ctypes as C
import os

PATH_MAX = 256

OK = 0
E_ARG, E_HIP, E_NOMEM, E_DT, E_IO, E_STATE, E_NODEV = -1, -2, -3, -4, -5, -6, -7
ERRORS = {E_ARG: "invalid argument", E_HIP: "HIP runtime error", E_NOMEM: "out of memory",
          E_DT: "dt above the stability bound", E_IO: "I/O error", E_STATE: "call out of order",
          E_NODEV: "no HIP device"}

FLUID, WALL, BODY, PISTON, KILLED = 1, 0, 2, 3, 100
DIFF_ANTUONO, DIFF_MC = 0, 1

F_POS, F_VEL, F_RHOP, F_ACC, F_TYPE, F_TABLE, F_HASH = 1, 2, 4, 8, 16, 32, 64
F_ALL = 127

A_POSR, A_VELP, A_TPOSR, A_TVELP, A_TYPE, A_TABLE = 0, 1, 2, 3, 4, 5
A_KEY, A_INDEX, A_CELL_START, A_ACC, A_GD = 6, 7, 8, 9, 10

_f, _i = C.c_float, C.c_int


class Json(C.Structure):
    _fields_ = [(n, _f) for n in ("g", "rho0", "dx", "h_factor", "r_factor", "eta_factor",
                                  "cs_factor", "delta", "alpha")] + \
               [(n, _f) for n in ("xmin", "xmax", "ymin", "ymax", "zmin", "zmax")] + \
               [("grid_layer_factor", _f), ("grid_size_factor", _f), ("dt", _f)] + \
               [(n, _i) for n in ("start_step", "current_step", "end_step", "output_step",
                                  "ptc_num", "water_ptc_num", "air_ptc_num", "wall_ptc_num",
                                  "rigid_ptc_num", "thread_num", "gpu_id")] + \
               [("input_path", C.c_char * PATH_MAX), ("output_path", C.c_char * PATH_MAX),
                ("git_hash", C.c_char * 64)] + \
               [("dim", _i), ("gamma", _f), ("c0", _f), ("cfl", _f), ("diffusion", _i),
                ("output_fields", _i), ("binary", _i)]


class CpuParam(C.Structure):
    _fields_ = [(n, _f) for n in ("g", "m", "rho0", "dx", "h", "r", "eta", "cs", "adh", "delta",
                                  "alpha", "dt")] + \
               [(n, _i) for n in ("start_step", "current_step", "end_step", "output_step")] + \
               [(n, _f) for n in ("xmin", "xmax", "ymin", "ymax", "zmin", "zmax")] + \
               [("grid_factor", _i), ("grid_size", _f)] + \
               [(n, _f) for n in ("grid_xmin", "grid_ymin", "grid_zmin", "grid_xmax", "grid_ymax",
                                  "grid_zmax")] + \
               [(n, _i) for n in ("grid_xdim", "grid_ydim", "grid_zdim", "grid_num", "grid_hash_min",
                                  "grid_hash_max", "ptc_num", "water_ptc_num", "air_ptc_num",
                                  "wall_ptc_num", "rigid_ptc_num", "thread_num", "gpu_id")] + \
               [("input_path", C.c_char * PATH_MAX), ("output_path", C.c_char * PATH_MAX),
                ("git_hash", C.c_char * 64)] + \
               [("dim", _i), ("gamma", _f), ("diffusion", _i), ("output_fields", _i), ("binary", _i)]


class Params(C.Structure):
    _fields_ = [(n, _f) for n in ("half_dt", "h", "g", "m", "rho0", "adh", "h2", "eta2", "cs2",
                                  "alpha", "delta_h_cs", "h_cs_rho_m", "grid_size")] + \
               [("grid_factor", _i)] + \
               [(n, _f) for n in ("grid_xmin", "grid_ymin", "grid_zmin")] + \
               [(n, _i) for n in ("grid_xdim", "grid_ydim", "grid_zdim", "grid_hash_min",
                                  "grid_hash_max", "ptc_num", "dim")] + \
               [("gamma", _f), ("diffusion", _i), ("rc2", _f)]


class HostState(C.Structure):
    _fields_ = [("pos", C.POINTER(_f)), ("vel", C.POINTER(_f)), ("rhop", C.POINTER(_f)),
                ("acc_drhodt", C.POINTER(_f)), ("type", C.POINTER(_i)), ("table", C.POINTER(_i)),
                ("hash", C.POINTER(_i))]


class Diag(C.Structure):
    _fields_ = [("n_fluid", _i), ("n_killed", _i), ("mass_fluid", C.c_double), ("front_x", C.c_double),
                ("zmax_fluid", C.c_double), ("kinetic", C.c_double), ("momentum", C.c_double * 3)]

    def as_dict(self):
        return {"n_fluid": self.n_fluid, "n_killed": self.n_killed, "mass_fluid": self.mass_fluid,
                "front_x": self.front_x, "zmax_fluid": self.zmax_fluid, "kinetic": self.kinetic,
                "momentum": list(self.momentum)}


class Rigid(C.Structure):
    _fields_ = [(n, C.c_double * 3) for n in ("c", "vel", "omega")] + [("q", C.c_double * 4)] + \
               [(n, C.c_double * 3) for n in ("F", "T", "acc", "alpha")] + \
               [("I", C.c_double * 9), ("mass", C.c_double)]


class Tuning(C.Structure):
    _fields_ = [("walk", _i), ("radix_bits", _i), ("block", _i), ("sort_tile", _i), ("skin", _f),
                ("tile_rows", _i)]


WALK_MASK, WALK_HITS, WALK_LIST, WALK_SKIN = 0, 1, 2, 3

NF, NH, NT = 24, 64, 12
PROBE_NOALLOC_MAX = 4096  # ZSPH_PROBE_NOALLOC_MAX
FRAME_SLOTS = 3  # ZSPH_FRAME_SLOTS
NPARAM = NF * NH + NH + 2 * (NH * NH + NH) + NH * 2 + 2


class Closure(C.Structure):
    _fields_ = [("w0", _f * (NF * NH)), ("b0", _f * NH), ("w1", _f * (NH * NH)), ("b1", _f * NH),
                ("w2", _f * (NH * NH)), ("b2", _f * NH), ("w3", _f * (NH * 2)), ("b3", _f * 2),
                ("mean", _f * NF), ("inv_std", _f * NF), ("lo", _f * NF), ("hi", _f * NF), ("out_scale", _f)]

    def params(self):
        """The ZSPH_NPARAM parameters (w0 .. b3) as a writable numpy view."""
        import numpy as np
        return np.frombuffer(self, dtype=np.float32, count=NPARAM)


class DamBreak(C.Structure):
    _fields_ = [("dim", _i), ("dx", _f), ("lx", _f), ("wy", _f), ("lz", _f), ("cw", _f), ("ch", _f),
                ("layers", _i), ("wall_height", _f)]


class Piston(C.Structure):
    _fields_ = [("amplitude", C.c_double), ("frequency", C.c_double), ("phase", C.c_double), ("axis", _i)]


class WaveTank(C.Structure):
    _fields_ = [("dx", _f), ("lx", _f), ("wy", _f), ("lz", _f), ("depth", _f), ("box_lo", _f * 3),
                ("box_hi", _f * 3), ("layers", _i), ("piston_travel", _f)]


class CheckpointInfo(C.Structure):
    _fields_ = [("step", C.c_longlong), ("time", C.c_double), ("ptc_num", _i), ("cells", _i), ("scenes", _i),
                ("dim", _i), ("grid", _i * 3), ("killed", _i), ("has_body", _i), ("has_piston", _i),
                ("closure", _i), ("payload_bytes", C.c_longlong)]


def _sig(lib, name, res, args):
    fn = getattr(lib, name)
    fn.restype = res
    fn.argtypes = args
    return fn


def load(path):
    """Load libzsph.so at `path` and declare every exported entry point."""
    if not os.path.exists(path):
        raise RuntimeError(
            f"libzsph.so not found at {path}: build it first (python -c 'import __graft_entry__ as g; g.build()')")
    lib = C.CDLL(path)
    P = C.POINTER
    vp = C.c_void_p
    _sig(lib, "zsph_sizeof", C.c_size_t, [_i])
    for which, cls in enumerate((Json, CpuParam, Params, HostState, DamBreak, Rigid, Diag, Tuning, Piston,
                                 WaveTank, Closure, CheckpointInfo)):
        got = lib.zsph_sizeof(which)
        if got != C.sizeof(cls):
            raise RuntimeError(f"ABI mismatch: sizeof({cls.__name__}) = {C.sizeof(cls)} here, {got} in {path}")
    _sig(lib, "zsph_json_defaults", None, [P(Json)])
    _sig(lib, "zsph_read_json", _i, [C.c_char_p, P(Json)])
    _sig(lib, "zsph_set_cpu_param", _i, [P(CpuParam), P(Json)])
    _sig(lib, "zsph_set_gpu_param", None, [P(Params), P(CpuParam)])
    _sig(lib, "zsph_case_dambreak", _i, [P(DamBreak), P(Json), P(HostState), P(_i)])
    _sig(lib, "zsph_write_vtk", _i, [C.c_char_p, P(HostState), _i, _i, _i])
    _sig(lib, "zsph_read_vtk", _i, [C.c_char_p, P(HostState), _i, _f])
    _sig(lib, "zsph_vtk_count", _i, [C.c_char_p, P(_i)])
    _sig(lib, "zsph_rigid_mass_properties", _i, [P(_f), _i, C.c_double, P(Rigid)])
    _sig(lib, "zsph_rigid_acc_alpha", None, [P(Rigid)])
    _sig(lib, "zsph_rigid_euler_step", None, [P(Rigid), P(_f), _i, C.c_double])
    _sig(lib, "zsph_create", _i, [P(Params), _i, P(vp)])
    _sig(lib, "zsph_destroy", None, [vp])
    _sig(lib, "zsph_sync", _i, [vp])
    _sig(lib, "zsph_last_error", C.c_char_p, [vp])
    _sig(lib, "zsph_upload", _i, [vp, _i, P(HostState)])
    _sig(lib, "zsph_download", _i, [vp, _i, P(HostState), _i])
    for name in ("zsph_hash", "zsph_boundary", "zsph_governing", "zsph_prediction", "zsph_correction",
                 "zsph_step"):
        _sig(lib, name, _i, [vp, _i])
    _sig(lib, "zsph_sort_index", _i, [vp])
    _sig(lib, "zsph_sort_data", _i, [vp, _i, _i])
    _sig(lib, "zsph_array_get", _i, [vp, _i, _i, vp, C.c_size_t])
    _sig(lib, "zsph_array_set", _i, [vp, _i, _i, vp, C.c_size_t])
    _sig(lib, "zsph_neighbor_stats", _i, [vp, _i, P(_i), P(C.c_uint64)])
    _sig(lib, "zsph_killed", _i, [vp, P(_i)])
    _sig(lib, "zsph_time_steps", _i, [vp, _i, P(_f)])
    _sig(lib, "zsph_diagnostics", _i, [vp, _i, _f, P(Diag)])
    _sig(lib, "zsph_probe_pressure", _i, [vp, _i, P(_f), _i, P(_f)])
    _sig(lib, "zsph_probe_pressure_noalloc", _i, [vp, _i, P(_f), _i, P(_f)])
    _sig(lib, "zsph_download_async", _i, [vp, _i, _i, P(vp)])
    _sig(lib, "zsph_frame_wait", _i, [vp, P(HostState)])
    _sig(lib, "zsph_profile_steps", _i, [vp, _i, P(_f)])
    _sig(lib, "zsph_set_tuning", _i, [vp, P(Tuning)])
    _sig(lib, "zsph_get_tuning", _i, [vp, P(Tuning)])
    _sig(lib, "zsph_autotune", _i, [vp, _i, P(Tuning), P(_f)])
    _sig(lib, "zsph_rebuilds", _i, [vp, P(C.c_longlong)])
    _sig(lib, "zsph_set_body", _i, [vp, P(Rigid)])
    _sig(lib, "zsph_get_body", _i, [vp, P(Rigid)])
    _sig(lib, "zsph_set_piston", _i, [vp, P(Piston)])
    _sig(lib, "zsph_set_time", _i, [vp, C.c_double])
    _sig(lib, "zsph_get_time", _i, [vp, P(C.c_double)])
    _sig(lib, "zsph_body_reduce", _i, [vp, _i, P(C.c_double), P(C.c_double)])
    _sig(lib, "zsph_case_wavetank", _i, [P(WaveTank), P(Json), P(HostState), P(_i)])
    _sig(lib, "zsph_create_batch", _i, [P(Params), _i, _i, P(vp)])
    _sig(lib, "zsph_scene_diagnostics", _i, [vp, _i, _f, P(Diag)])
    Pf = P(_f)
    _sig(lib, "zsph_mlp_create", _i, [_i, P(vp)])
    _sig(lib, "zsph_mlp_destroy", None, [vp])
    _sig(lib, "zsph_mlp_last_error", C.c_char_p, [vp])
    _sig(lib, "zsph_mlp_set", _i, [vp, P(Closure)])
    _sig(lib, "zsph_mlp_get", _i, [vp, P(Closure)])
    _sig(lib, "zsph_mlp_data", _i, [vp, Pf, Pf, _i, _f, _f])
    _sig(lib, "zsph_mlp_eval", _i, [vp, _i, _i, Pf, Pf])
    _sig(lib, "zsph_mlp_adam", _i, [vp, _i, _i, _f, _f, _f, _f, Pf])
    _sig(lib, "zsph_mlp_infer", _i, [vp, Pf, _i, Pf])
    _sig(lib, "zsph_set_closure", _i, [vp, P(Closure)])
    _sig(lib, "zsph_closure_sample", _i, [vp, _i, Pf, Pf])
    _sig(lib, "zsph_filter_fields", _i, [vp, _i, _f, _f, P(_i), _i, Pf])
    _sig(lib, "zsph_clock_governing", _i, [vp, _i])
    _sig(lib, "zsph_set_graph", _i, [vp, _i])
    _sig(lib, "zsph_clock_read", _i, [vp, P(C.c_double), P(C.c_longlong)])
    _sig(lib, "zsph_clock_stage", _i, [vp, _i, _i])
    _sig(lib, "zsph_clock_read_stage", _i, [vp, _i, P(C.c_double), P(C.c_longlong)])
    _sig(lib, "zsph_checkpoint_write", _i, [vp, C.c_char_p, C.c_longlong])
    _sig(lib, "zsph_checkpoint_read", _i, [vp, C.c_char_p, P(C.c_longlong)])
    _sig(lib, "zsph_checkpoint_inspect", _i, [C.c_char_p, _i, P(CheckpointInfo)])
    return lib


# Every symbol include/zsph.h declares (checked by the CPU test suite).
EXPORTS = (
    "zsph_json_defaults", "zsph_read_json", "zsph_set_cpu_param", "zsph_set_gpu_param",
    "zsph_case_dambreak", "zsph_write_vtk", "zsph_read_vtk", "zsph_vtk_count", "zsph_create",
    "zsph_destroy", "zsph_sync", "zsph_last_error", "zsph_upload", "zsph_download", "zsph_hash",
    "zsph_sort_index", "zsph_sort_data", "zsph_boundary", "zsph_governing", "zsph_prediction",
    "zsph_correction", "zsph_step", "zsph_array_get", "zsph_array_set", "zsph_neighbor_stats",
    "zsph_killed", "zsph_sizeof", "zsph_time_steps", "zsph_profile_steps", "zsph_diagnostics",
    "zsph_probe_pressure", "zsph_probe_pressure_noalloc", "zsph_download_async", "zsph_frame_wait", "zsph_rigid_mass_properties", "zsph_rigid_acc_alpha", "zsph_rigid_euler_step",
    "zsph_set_tuning", "zsph_get_tuning", "zsph_autotune", "zsph_set_body", "zsph_get_body",
    "zsph_set_piston", "zsph_set_time", "zsph_get_time", "zsph_body_reduce", "zsph_case_wavetank",
    "zsph_create_batch", "zsph_scene_diagnostics", "zsph_mlp_create", "zsph_mlp_destroy", "zsph_mlp_last_error",
    "zsph_mlp_set", "zsph_mlp_get", "zsph_mlp_data", "zsph_mlp_eval", "zsph_mlp_adam", "zsph_mlp_infer",
    "zsph_set_closure", "zsph_closure_sample", "zsph_filter_fields", "zsph_clock_governing", "zsph_clock_read", "zsph_rebuilds", "zsph_clock_stage", "zsph_clock_read_stage",
    "zsph_set_graph", "zsph_checkpoint_write", "zsph_checkpoint_read", "zsph_checkpoint_inspect",
)

STAGES = ("hash", "sort", "cells", "reorder", "wall", "ldg", "governing", "integrate")
ST_HASH, ST_SORT, ST_CELLS, ST_REORDER, ST_WALL, ST_LDG, ST_GOVERNING, ST_INTEGRATE = range(8)
