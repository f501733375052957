"""CPU oracle of the delta-SPH step -- TEST INFRASTRUCTURE ONLY.

ctypes front end of oracle/liboracle.so (built from zsph_oracle.hpp by
oracle/Makefile).  Only tests/, the smoke check and the bench's CPU-baseline
leg use it, as the checker; the product package never imports it.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# ZSPH_ORACLE_LIB: another build of the same library (the sanitizer build of
# tests/test_sanitizers.py)
LIB_PATH = os.environ.get("ZSPH_ORACLE_LIB") or os.path.join(_HERE, "liboracle.so")
_LIB = None

A_ACC_SCALE = 100   # per particle (sum |acc terms|, sum |drho terms|)
A_GD_SCALE = 101    # per particle (sum |grad rho terms|, sum |div v terms|)
A_FEATURES = 102    # per particle, closure features (after features() or a closure governing pass)
A_TERMS = 103       # per particle: cont, diff, pressure + g (3), viscous (3) of the last governing pass
A_SCALE = 104       # per particle closure scale factors (empty without the closure)


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: run make -C oracle")
        from zsph_amd import _abi as abi
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        P = C.POINTER

        def sig(name, res, args):
            f = getattr(L, name)
            f.restype, f.argtypes = res, args

        sig("orc_create", vp, [P(abi.Params), C.c_int])
        sig("orc_destroy", None, [vp])
        sig("orc_upload", C.c_int, [vp, C.c_int, P(abi.HostState)])
        sig("orc_download", C.c_int, [vp, C.c_int, P(abi.HostState), C.c_int])
        for n in ("orc_hash", "orc_boundary", "orc_wall", "orc_ldg", "orc_governing", "orc_prediction",
                  "orc_correction", "orc_step"):
            sig(n, C.c_int, [vp, C.c_int])
        sig("orc_sort_index", C.c_int, [vp])
        sig("orc_set_threads", C.c_int, [C.c_int])
        sig("orc_sort_data", C.c_int, [vp, C.c_int, C.c_int])
        sig("orc_killed", C.c_int, [vp, P(C.c_longlong)])
        sig("orc_array_get", C.c_int, [vp, C.c_int, C.c_int, P(C.c_double), C.c_size_t])
        sig("orc_array_set", C.c_int, [vp, C.c_int, C.c_int, P(C.c_double), C.c_size_t])
        sig("orc_neighbor_stats", C.c_int, [vp, C.c_int, C.c_int, P(C.c_int), P(C.c_uint64)])
        for n in ("orc_kernel_W", "orc_kernel_F", "orc_eos_p", "orc_eos_rho"):
            sig(n, C.c_double, [P(abi.Params), C.c_double])
        sig("orc_cell", C.c_int, [C.c_float, C.c_float, C.c_float, C.c_int])
        sig("orc_diagnostics", C.c_int, [vp, C.c_int, C.c_float, P(abi.Diag)])
        sig("orc_probe_pressure", C.c_int, [vp, C.c_int, P(C.c_float), C.c_int, P(C.c_double)])
        sig("orc_set_body", C.c_int, [vp, P(abi.Rigid)])
        sig("orc_get_body", C.c_int, [vp, P(abi.Rigid)])
        sig("orc_set_piston", C.c_int, [vp, C.c_double, C.c_double, C.c_double, C.c_int])
        sig("orc_set_time", C.c_int, [vp, C.c_double])
        sig("orc_get_time", C.c_int, [vp, P(C.c_double)])
        sig("orc_body_reduce", C.c_int, [vp, C.c_int, P(C.c_double), P(C.c_double)])
        sig("orc_rigid_acc_alpha", None, [P(abi.Rigid), C.c_double])
        sig("orc_set_closure", C.c_int, [vp, P(abi.Closure), C.c_int])
        sig("orc_features", C.c_int, [vp, C.c_int])
        sig("orc_mlp_loss_grad", C.c_double, [P(abi.Closure), C.c_int, P(C.c_float), P(C.c_float), C.c_int,
                                              C.c_double, C.c_double, P(C.c_double)])
        sig("orc_mlp_infer", None, [P(abi.Closure), C.c_int, P(C.c_float), C.c_int, P(C.c_double)])
        sig("orc_bf16_round", C.c_double, [C.c_double])
        sig("orc_rigid_advance", None, [P(abi.Rigid), P(abi.Rigid), C.c_double, P(abi.Rigid)])
        _LIB = L
    return _LIB


class Oracle:
    """Same call sequence as zsph_amd.Solver, on the CPU (fp64=True: parity oracle)."""

    def __init__(self, params, fp64=True):
        self.params = params
        self.n = params.ptc_num
        self.grid_num = params.grid_xdim * params.grid_ydim * params.grid_zdim
        self._h = lib().orc_create(C.byref(params), 1 if fp64 else 0)

    def close(self):
        if self._h:
            lib().orc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _ok(self, rc, what):
        if rc != 0:
            raise RuntimeError(f"oracle {what} failed ({rc})")

    def upload(self, host, buf=0):
        st = host.state()
        self._ok(lib().orc_upload(self._h, buf, C.byref(st)), "upload")

    def download(self, buf=0, fields=127):
        from zsph_amd import HostArrays
        h = HostArrays(self.n)
        st = h.state()
        self._ok(lib().orc_download(self._h, buf, C.byref(st), fields), "download")
        return h

    def hash(self, buf=0): self._ok(lib().orc_hash(self._h, buf), "hash")
    def sort_index(self): self._ok(lib().orc_sort_index(self._h), "sort_index")
    def sort_data(self, src=0, dst=1): self._ok(lib().orc_sort_data(self._h, src, dst), "sort_data")
    def boundary(self, buf=1): self._ok(lib().orc_boundary(self._h, buf), "boundary")
    def wall(self, buf=1): self._ok(lib().orc_wall(self._h, buf), "wall")
    def ldg(self, buf=1): self._ok(lib().orc_ldg(self._h, buf), "ldg")
    def governing(self, buf=1): self._ok(lib().orc_governing(self._h, buf), "governing")
    def prediction(self, buf=1): self._ok(lib().orc_prediction(self._h, buf), "prediction")
    def correction(self, buf=0): self._ok(lib().orc_correction(self._h, buf), "correction")
    def step(self, n=1): self._ok(lib().orc_step(self._h, n), "step")

    def killed(self):
        k = C.c_longlong(0)
        self._ok(lib().orc_killed(self._h, C.byref(k)), "killed")
        return k.value

    def get(self, which, buf=0):
        from zsph_amd import A_CELL_START
        if which == A_CELL_START:
            count = 2 * (self.grid_num + 1)
            shape = (count,)
        elif which in (A_ACC_SCALE, A_GD_SCALE, A_SCALE):
            count, shape = 2 * self.n, (self.n, 2)
        elif which == A_FEATURES:
            count, shape = 24 * self.n, (self.n, 24)
        elif which == A_TERMS:
            count, shape = 8 * self.n, (self.n, 8)
        elif which in (0, 1, 2, 3, 9, 10):
            count, shape = 4 * self.n, (self.n, 4)
        else:
            count, shape = self.n, (self.n,)
        out = np.empty(count, np.float64)
        self._ok(lib().orc_array_get(self._h, which, buf, out.ctypes.data_as(C.POINTER(C.c_double)), count),
                 "array_get")
        return out.reshape(shape)

    def set(self, which, arr, buf=0):
        a = np.ascontiguousarray(arr, dtype=np.float64).ravel()
        self._ok(lib().orc_array_set(self._h, which, buf, a.ctypes.data_as(C.POINTER(C.c_double)), a.size),
                 "array_set")

    def diagnostics(self, buf=0, front_zmax=float("inf")):
        from zsph_amd import _abi as abi
        d = abi.Diag()
        self._ok(lib().orc_diagnostics(self._h, buf, front_zmax, C.byref(d)), "diagnostics")
        return d.as_dict()

    def probe_pressure(self, points, buf=0):
        pts = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
        out = np.zeros(len(pts), np.float64)
        self._ok(lib().orc_probe_pressure(self._h, buf, pts.ctypes.data_as(C.POINTER(C.c_float)), len(pts),
                                          out.ctypes.data_as(C.POINTER(C.c_double))), "probe_pressure")
        return out

    def neighbor_stats(self, buf=1, brute=False):
        cnt = np.empty(self.n, np.int32)
        dig = np.empty(self.n, np.uint64)
        self._ok(lib().orc_neighbor_stats(self._h, buf, 1 if brute else 0, cnt.ctypes.data_as(C.POINTER(C.c_int)),
                                          dig.ctypes.data_as(C.POINTER(C.c_uint64))), "neighbor_stats")
        return cnt, dig

    # ---- closure ---------------------------------------------------------------
    def set_closure(self, net, emulate_bf16=True):
        self._ok(lib().orc_set_closure(self._h, C.byref(net) if net is not None else None,
                                       1 if emulate_bf16 else 0), "set_closure")

    def features(self, buf=1):
        self._ok(lib().orc_features(self._h, buf), "features")
        return self.get(A_FEATURES, buf)

    # ---- moving dummies (buffer 0) -------------------------------------------
    def set_body(self, rigid):
        self._ok(lib().orc_set_body(self._h, C.byref(rigid)), "set_body")

    def get_body(self):
        from zsph_amd import _abi as abi
        r = abi.Rigid()
        self._ok(lib().orc_get_body(self._h, C.byref(r)), "get_body")
        return r

    def set_piston(self, amplitude, frequency, phase=0.0, axis=0):
        self._ok(lib().orc_set_piston(self._h, amplitude, frequency, phase, axis), "set_piston")

    def set_time(self, t):
        self._ok(lib().orc_set_time(self._h, t), "set_time")

    def time(self):
        t = C.c_double(0.0)
        self._ok(lib().orc_get_time(self._h, C.byref(t)), "get_time")
        return t.value

    def body_reduce(self, buf=1):
        F = (C.c_double * 3)()
        T = (C.c_double * 3)()
        self._ok(lib().orc_body_reduce(self._h, buf, F, T), "body_reduce")
        return np.array(F[:]), np.array(T[:])


def set_threads(n):
    """OpenMP threads of the oracle's later steps (0: keep); returns the previous count."""
    return lib().orc_set_threads(int(n))


def kernel_W(params, r): return lib().orc_kernel_W(C.byref(params), r)
def kernel_F(params, r): return lib().orc_kernel_F(C.byref(params), r)
def eos_p(params, rho): return lib().orc_eos_p(C.byref(params), rho)
def eos_rho(params, p): return lib().orc_eos_rho(C.byref(params), p)
def cell(x, o, size, dim): return lib().orc_cell(x, o, size, dim)



def rigid_acc_alpha(rigid, g=0.0):
    """The oracle's FP64 accelerations of `rigid` (abi.Rigid, in place), gravity g along z."""
    lib().orc_rigid_acc_alpha(C.byref(rigid), g)


def rigid_advance(s0, rates, dt):
    """(c, v, w, q) of s0 advanced by dt with the rates (v, a, w, alpha, q) of `rates`."""
    from zsph_amd import _abi as abi
    out = abi.Rigid()
    lib().orc_rigid_advance(C.byref(s0), C.byref(rates), dt, C.byref(out))
    return out


def mlp_loss_grad(net, x, rec, w_rho=1.0, w_acc=1.0, emulate_bf16=False, grad=True):
    """(mean loss, gradient or None) of the closure network `net` (abi.Closure) in FP64."""
    x = np.ascontiguousarray(x, np.float32)
    rec = np.ascontiguousarray(rec, np.float32)
    g = np.zeros(net_nparam(), np.float64) if grad else None
    loss = lib().orc_mlp_loss_grad(C.byref(net), 1 if emulate_bf16 else 0, x.ctypes.data_as(C.POINTER(C.c_float)),
                                   rec.ctypes.data_as(C.POINTER(C.c_float)), len(x), w_rho, w_acc,
                                   g.ctypes.data_as(C.POINTER(C.c_double)) if grad else None)
    return loss, g


def mlp_infer(net, x, emulate_bf16=False):
    x = np.ascontiguousarray(x, np.float32)
    out = np.zeros((len(x), 2), np.float64)
    lib().orc_mlp_infer(C.byref(net), 1 if emulate_bf16 else 0, x.ctypes.data_as(C.POINTER(C.c_float)), len(x),
                        out.ctypes.data_as(C.POINTER(C.c_double)))
    return out


def bf16_round(v):
    return lib().orc_bf16_round(float(v))


def net_nparam():
    from zsph_amd import _abi as abi
    return abi.NPARAM
