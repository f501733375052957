"""Learned closure: the per-particle network that scales the density-diffusion
and viscous terms (C ABI ``zsph_mlp_*`` / ``zsph_set_closure``).

The network maps ZSPH_NF = 24 neighbour-moment features of a fluid particle
(normalised by the training set's mean and spread) through three SiLU layers
of 64 to two outputs o, giving the scale factors s = 1 + out_scale tanh(o) of
the Antuono density-diffusion term (s_d) and of the artificial viscosity
(s_v).  A zero output layer is the plain method (s = 1); so is any particle
whose normalised features leave the training range.
"""
import ctypes as C

import numpy as np

from . import ZsphError, _abi as abi, lib

NF, NH, NT, NPARAM = abi.NF, abi.NH, abi.NT, abi.NPARAM

FEATURE_NAMES = (
    "kernel-sum deficiency 1 - sum V W", "trace(L)/dim (free-surface indicator)", "det L", "|offdiag L|",
    "grad rho_x h/rho0", "grad rho_y h/rho0", "grad rho_z h/rho0", "|grad rho| h/rho0",
    "div v h/c0", "curl_x h/c0", "curl_y h/c0", "curl_z h/c0", "|curl v| h/c0", "|dev strain| h/c0",
    "v_x/c0", "v_y/c0", "v_z/c0", "|v|/c0", "(rho-rho0)/rho0", "p/(rho0 c0^2)",
    "neighbours/nominal", "velocity spread/c0", "smoothed density contrast", "dummy-neighbour fraction",
)
assert len(FEATURE_NAMES) == NF


def init_network(seed=0, out_scale=0.9, zero_output=True):
    """He-initialised network; a zero output layer starts as the plain method."""
    rng = np.random.default_rng(seed)
    net = abi.Closure()
    P = net.params()
    sizes = [(NF, NH), (NH, NH), (NH, NH), (NH, 2)]
    off = 0
    for k, (fin, fout) in enumerate(sizes):
        w = rng.normal(0.0, np.sqrt(2.0 / fin), (fin, fout)).astype(np.float32)
        if k == 3 and zero_output:
            w[:] = 0.0
        P[off:off + fin * fout] = w.ravel()
        off += fin * fout
        P[off:off + fout] = 0.0
        off += fout
    assert off == NPARAM
    net.mean[:] = [0.0] * NF
    net.inv_std[:] = [1.0] * NF
    net.lo[:] = [-1e30] * NF
    net.hi[:] = [1e30] * NF
    net.out_scale = out_scale
    return net


def set_normalisation(net, x, margin=0.5):
    """Input statistics from training features x (n x NF): mean, 1/std and the
    normalised range [min - margin, max + margin] outside which the plain terms apply."""
    x = np.asarray(x, np.float64)
    mean = x.mean(0)
    std = x.std(0)
    inv = np.where(std > 1e-12, 1.0 / np.maximum(std, 1e-12), 1.0)
    z = (x - mean) * inv
    net.mean[:] = mean.astype(np.float32).tolist()
    net.inv_std[:] = inv.astype(np.float32).tolist()
    net.lo[:] = (z.min(0) - margin).astype(np.float32).tolist()
    net.hi[:] = (z.max(0) + margin).astype(np.float32).tolist()
    return net


def copy_network(net):
    out = abi.Closure()
    C.memmove(C.byref(out), C.byref(net), C.sizeof(abi.Closure))
    return out


class Mlp:
    """Device-resident network + training set (zsph_mlp)."""

    def __init__(self, net=None, device=0):
        self._h = C.c_void_p()
        rc = lib().zsph_mlp_create(device, C.byref(self._h))
        if rc != abi.OK:
            raise ZsphError(rc, "zsph_mlp_create")
        self.n = 0
        if net is not None:
            self.set(net)

    def _chk(self, rc, what):
        if rc != abi.OK:
            raise ZsphError(rc, f"{what}: {lib().zsph_mlp_last_error(self._h).decode()}")

    def close(self):
        if self._h:
            lib().zsph_mlp_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set(self, net):
        self._chk(lib().zsph_mlp_set(self._h, C.byref(net)), "zsph_mlp_set")

    def get(self):
        net = abi.Closure()
        self._chk(lib().zsph_mlp_get(self._h, C.byref(net)), "zsph_mlp_get")
        return net

    def data(self, x, rec, w_rho=1.0, w_acc=1.0):
        x = np.ascontiguousarray(x, np.float32)
        rec = np.ascontiguousarray(rec, np.float32)
        assert x.shape[1] == NF and rec.shape[1] == NT and len(x) == len(rec)
        fp = C.POINTER(C.c_float)
        self._chk(lib().zsph_mlp_data(self._h, x.ctypes.data_as(fp), rec.ctypes.data_as(fp), len(x), w_rho, w_acc),
                  "zsph_mlp_data")
        self.n = len(x)

    def eval(self, first=0, count=None, grad=True):
        count = self.n - first if count is None else count
        loss = C.c_float(0.0)
        g = np.zeros(NPARAM, np.float32) if grad else None
        self._chk(lib().zsph_mlp_eval(self._h, first, count, C.byref(loss),
                                      g.ctypes.data_as(C.POINTER(C.c_float)) if grad else None), "zsph_mlp_eval")
        return loss.value, g

    def adam(self, first, count, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        loss = C.c_float(0.0)
        self._chk(lib().zsph_mlp_adam(self._h, first, count, lr, beta1, beta2, eps, C.byref(loss)), "zsph_mlp_adam")
        return loss.value

    def infer(self, x):
        x = np.ascontiguousarray(x, np.float32)
        out = np.zeros((len(x), 2), np.float32)
        fp = C.POINTER(C.c_float)
        self._chk(lib().zsph_mlp_infer(self._h, x.ctypes.data_as(fp), len(x), out.ctypes.data_as(fp)),
                  "zsph_mlp_infer")
        return out
