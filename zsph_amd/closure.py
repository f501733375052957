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

from . import FLUID, ZsphError, _abi as abi, lib

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


# ---------------------------------------------------------------------------
# Training data: coarse-grained snapshots of fine plain runs
# ---------------------------------------------------------------------------
# The coarse particles follow the fine particles of the even lattice sites
# (one per 2 x 2 x 2 block), displaced by half a fine spacing along each
# lattice axis to the block centre -- at t = 0 exactly the coarse lattice, so
# the coarse fluid keeps its half-spacing gap to the walls -- and move with
# the fine flow, so the coarse configuration stays as regular as the fine one.  At a snapshot the coarse state at those
# particles is the fine (rho, v) Shepard-filtered with the coarse kernel
# (zsph_filter_fields, h_c = 2 h_f, radius 2 h_c), and the targets are the
# equally filtered fine (a, drho/dt).  The coarse solver evaluates its plain
# terms at that state (zsph_closure_sample); the network learns the scale
# factors that bring coarse diffusion and viscosity closest to the targets.

def fluid_lattice(case):
    """(nx, ny, nz) of the generator's fluid lattice (x fastest, then y, then z)."""
    p = case.host.pos[case.host.type == FLUID].astype(np.float64)
    dx = float(case.json.dx)
    n = [int(round((p[:, a].max() - p[:, a].min()) / dx)) + 1 for a in range(3)]
    assert n[0] * n[1] * n[2] == len(p), "fluid is not one lattice block"
    return n


def block_map(fine, coarse):
    """Fine fluid indices (n_coarse_fluid x 2^dim) of each coarse fluid particle's
    lattice block; column 0 is the block's even site."""
    nf, nc = fluid_lattice(fine), fluid_lattice(coarse)
    dim = fine.cpu.dim
    span = [2, 2 if dim == 3 else 1, 2]
    for a in range(3):
        assert nc[a] * span[a] <= nf[a], "the coarse column must fit in the fine one"
    I, J, K = np.meshgrid(np.arange(nc[0]), np.arange(nc[1]), np.arange(nc[2]), indexing="ij")
    I, J, K = (v.transpose(2, 1, 0).ravel() for v in (I, J, K))  # coarse index order: x fastest
    cols = []
    for dk in range(span[2]):
        for dj in range(span[1]):
            for di in range(span[0]):
                fi = (span[0] * I + di) + nf[0] * ((span[1] * J + dj) + nf[1] * (span[2] * K + dk))
                cols.append(fi)
    return np.stack(cols, 1)


def _tait_p(rho, case):
    g = case.params.gamma
    rho0 = case.params.rho0
    if g == 1.0:
        return case.params.cs2 * (rho - rho0)
    B = rho0 * case.params.cs2 / g
    return B * ((rho / rho0) ** g - 1.0)


def coarse_samples(s_fine, fine, s_coarse, coarse, subset):
    """Features and training records of the coarse-grained state of the fine
    solver's buffer 0 (left unchanged: the evaluation runs into buffer 1).
    subset: fine original indices of the coarse particles, in coarse order."""
    from . import A_ACC, F_POS, F_RHOP, F_TABLE, F_TYPE, F_VEL, KILLED
    s_fine.hash(0); s_fine.sort_index(); s_fine.sort_data(0, 1); s_fine.boundary(1); s_fine.governing(1)
    h = s_fine.download(1, F_POS | F_VEL | F_RHOP | F_TYPE | F_TABLE)
    row = np.empty(fine.n, np.int64)
    row[np.arange(fine.n) + h.table] = np.arange(fine.n)  # current row of each original particle
    r = row[subset]
    alive = h.type[r] == FLUID
    hc = float(coarse.params.h)
    flt = s_fine.filter_fields(r, 2.0 * hc, hc, buf=1).astype(np.float64)
    nfc = len(subset)
    host = coarse.host.copy()
    shift = 0.5 * float(fine.json.dx) * np.array([1.0, 1.0 if fine.cpu.dim == 3 else 0.0, 1.0])
    host.pos[:nfc] = (h.pos[r] + shift).astype(np.float32)
    host.vel[:nfc] = flt[:, 2:5].astype(np.float32)
    rho = flt[:, 1]
    host.rhop[:nfc, 0] = rho.astype(np.float32)
    host.rhop[:nfc, 1] = _tait_p(rho, coarse).astype(np.float32)
    host.type[:nfc][~alive] = KILLED
    s_coarse.upload(host)
    s_coarse.hash(0); s_coarse.sort_index(); s_coarse.sort_data(0, 1)
    f, t = s_coarse.closure_sample(1)
    orig = np.arange(coarse.n) + s_coarse.get(5, 1)  # original index of each row of buffer 1
    sel = (orig < nfc) & (s_coarse.get(4, 1) == FLUID)
    ids = orig[sel]
    X = f[sel]
    T = t[sel]
    rec = np.zeros((len(ids), NT), np.float32)
    rec[:, 0:2] = T[:, 0:2]
    rec[:, 2] = flt[ids, 8]
    rec[:, 3:9] = T[:, 2:8]
    rec[:, 9:12] = flt[ids, 5:8]
    return X, rec


def column_of_seed(seed, lo=(1.0, 0.45), hi=(1.4, 0.65), quantum=0.016):
    """Seeded column (width, height) on a `quantum` grid (SplitMix64 draws)."""
    from .cases import splitmix64
    s, u = (seed * 0x632BE59BD9B4E019) & ((1 << 64) - 1), []
    for _ in range(2):
        s, z = splitmix64(s)
        u.append((z >> 11) * 2.0 ** -53)
    w = lo[0] + (hi[0] - lo[0]) * u[0]
    h = lo[1] + (hi[1] - lo[1]) * u[1]
    return round(w / quantum) * quantum, round(h / quantum) * quantum


def collect(columns, dx_fine=0.008, t_end=0.5, n_snap=10, log=print, **case_kw):
    """Training set (X, rec) from fine plain 3-D dam breaks of the given
    (width, height) columns, coarse-grained to 2 dx_fine at n_snap times."""
    from . import Solver
    from .cases import dambreak3d
    Xs, Rs = [], []
    for (w, hgt) in columns:
        fine = dambreak3d(dx=dx_fine, cw=w, ch=hgt, grid_size_factor=2.0, **case_kw)
        coarse = dambreak3d(dx=2 * dx_fine, cw=w, ch=hgt, grid_size_factor=2.0, **case_kw)
        subset = block_map(fine, coarse)[:, 0]
        sf, sc = Solver(fine), Solver(coarse)
        sf.upload(fine.host)
        dt = 2.0 * float(np.float32(fine.params.half_dt))
        done = 0
        for k in range(1, n_snap + 1):
            target = int(round(k * t_end / n_snap / dt))
            sf.step(target - done)
            done = target
            X, rec = coarse_samples(sf, fine, sc, coarse, subset)
            Xs.append(X)
            Rs.append(rec)
            log(f"[collect] column {w:.3f}x{hgt:.3f}: t={done * dt:.3f} s, {len(X)} samples")
        sf.close()
        sc.close()
    return np.concatenate(Xs), np.concatenate(Rs)


def plain_loss(rec, w_rho, w_acc):
    """The training loss with s = 1 (the plain coarse method)."""
    r = rec.astype(np.float64)
    res_rho = r[:, 0] + r[:, 1] - r[:, 2]
    res_a = r[:, 3:6] + r[:, 6:9] - r[:, 9:12]
    return float(np.mean(w_rho * res_rho ** 2 + w_acc * (res_a ** 2).sum(1)))


def loss_weights(rec):
    """Weights that make both parts of the loss O(1): inverse target variances."""
    r = rec.astype(np.float64)
    return 1.0 / max(np.var(r[:, 2]), 1e-30), 1.0 / max(3.0 * np.var(r[:, 9:12]), 1e-30)


def train(X, rec, val_frac=0.1, iters=3000, batch=65536, lr=2e-3, seed=0, margin=0.5, log=print):
    """Fit the network on the device (zsph_mlp): shuffled split, Adam with a
    cosine learning-rate decay; returns (network, report)."""
    rng = np.random.default_rng(seed)
    perm = rng.permutation(len(X))
    X, rec = X[perm], rec[perm]
    nv = max(1, int(len(X) * val_frac))
    Xt, Rt, Xv, Rv = X[nv:], rec[nv:], X[:nv], rec[:nv]
    w_rho, w_acc = loss_weights(Rt)
    net = set_normalisation(init_network(seed), Xt, margin=margin)
    m = Mlp(net)
    m.data(Xv, Rv, w_rho, w_acc)
    val0, _ = m.eval(grad=False)
    m.data(Xt, Rt, w_rho, w_acc)
    n = len(Xt)
    b = min(batch, n)
    nb = max(1, n // b)
    for it in range(iters):
        cur = lr * 0.5 * (1.0 + np.cos(np.pi * it / iters))
        loss = m.adam((it % nb) * b, b, lr=cur)
        if it % 500 == 0 or it == iters - 1:
            log(f"[train] iter {it}: batch loss {loss:.5f} (lr {cur:.2e})")
    trained = m.get()
    m.data(Xv, Rv, w_rho, w_acc)
    val1, _ = m.eval(grad=False)
    m.close()
    report = {"samples": int(len(X)), "train": int(n), "validation": int(nv), "w_rho": w_rho, "w_acc": w_acc,
              "validation_loss_plain": float(val0), "validation_loss_trained": float(val1),
              "validation_loss_plain_check": plain_loss(Rv, w_rho, w_acc)}
    log(f"[train] validation loss: plain {val0:.5f} -> trained {val1:.5f}")
    return trained, report


def save_network(net, path, **meta):
    np.savez(path, params=net.params().copy(), mean=np.array(net.mean[:], np.float32),
             inv_std=np.array(net.inv_std[:], np.float32), lo=np.array(net.lo[:], np.float32),
             hi=np.array(net.hi[:], np.float32), out_scale=np.float32(net.out_scale),
             meta=np.array(repr(meta)))


def load_network(path):
    d = np.load(path, allow_pickle=False)
    net = abi.Closure()
    net.params()[:] = d["params"]
    net.mean[:] = d["mean"].tolist()
    net.inv_std[:] = d["inv_std"].tolist()
    net.lo[:] = d["lo"].tolist()
    net.hi[:] = d["hi"].tolist()
    net.out_scale = float(d["out_scale"])
    return net


# ---------------------------------------------------------------------------
# A-posteriori evaluation: front position and far-wall impact pressure
# ---------------------------------------------------------------------------

FAR_WALL_PROBES_Z = (0.01, 0.05, 0.10, 0.20)


def far_wall(case):
    """x of the far wall's face: half a spacing before its first dummy layer
    (dummies above the floor, inside the tank's width, beyond the column)."""
    pos = case.host.pos.astype(np.float64)
    dx = float(case.json.dx)
    dummy = case.host.type != FLUID
    xf = float(pos[~dummy, 0].max())
    y0, y1 = pos[~dummy, 1].min(), pos[~dummy, 1].max()
    far = dummy & (pos[:, 2] > 0.0) & (pos[:, 1] >= y0) & (pos[:, 1] <= y1) & (pos[:, 0] > xf)
    return float(pos[far, 0].min()) - 0.5 * dx


def far_wall_probes(case):
    """Probe points on the far wall face at mid-width, heights FAR_WALL_PROBES_Z."""
    y = 0.5 * float(case.cpu.ymax + case.cpu.ymin)
    return np.array([[far_wall(case), y, z] for z in FAR_WALL_PROBES_Z], np.float32)


def run_history(case, t_end, net=None, sample_dt=2.5e-4, front_zmax=0.05, log=print, skin=0.0):
    """Front x(t) (fluid below front_zmax) and far-wall probe pressures of one run
    (skin > 0: zsph_step's skin mode, the same physics summed in another order)."""
    import time as _time
    from . import Solver
    s = Solver(case)
    s.upload(case.host)
    if skin > 0.0:
        s.set_skin(skin)
    if net is not None:
        s.set_closure(net)
    dt = 2.0 * float(np.float32(case.params.half_dt))
    every = max(1, int(round(sample_dt / dt)))
    pts = far_wall_probes(case)
    steps = int(round(t_end / dt))
    t, front, probes = [], [], []
    s.step(1)
    s.sync()
    wall0 = _time.perf_counter()
    done = 1
    while done < steps:
        k = min(every, steps - done)
        s.step(k)
        done += k
        d = s.diagnostics(0, front_zmax)
        t.append(done * dt)
        front.append(d["front_x"])
        probes.append(s.probe_pressure(pts))
    s.sync()
    wall = _time.perf_counter() - wall0
    # stepping cost alone (graph replays between HIP events), for the wall-clock comparison
    ms = s.time_steps(50) / 50.0
    s.close()
    return {"t": np.array(t), "front": np.array(front), "probes": np.array(probes), "wall_s": wall,
            "ms_per_step": ms, "step_s": ms * 1e-3 * steps, "steps": steps, "dt": dt, "probe_points": pts}


def smoothed_peak(t, p, window=0.004):
    """Peak of each probe's pressure after a centred moving average over `window`
    seconds (impact pressures of weakly compressible SPH carry acoustic spikes
    shorter than the impact itself)."""
    p = np.asarray(p, np.float64)
    n = max(1, int(round(window / (t[1] - t[0])))) if len(t) > 1 else 1
    ker = np.ones(n) / n
    return np.array([np.convolve(p[:, k], ker, mode="same").max() for k in range(p.shape[1])])


def compare(fine, coarse, x_wall):
    """Relative errors of a coarse history against the fine one: the front
    position (maximum over the samples before it reaches the far wall), the raw
    and the 4 ms-smoothed peak probe pressures, and the wall-clock ratio of
    the stepping alone."""
    tf, tc = fine["t"], coarse["t"]
    ff = np.interp(tc, tf, fine["front"])
    before = ff < x_wall - 0.05
    front_err = float(np.max(np.abs(coarse["front"][before] - ff[before]) / ff[before])) if before.any() else 0.0
    out = {"front_err_max": front_err}
    for name, pf, pc in (("peak_p", fine["probes"].max(0), coarse["probes"].max(0)),
                         ("smoothed_peak_p", smoothed_peak(tf, fine["probes"]), smoothed_peak(tc, coarse["probes"]))):
        out[name + "_fine"] = np.asarray(pf, np.float64).tolist()
        out[name + "_coarse"] = np.asarray(pc, np.float64).tolist()
        out[name + "_err"] = (np.abs(pc - pf) / np.maximum(pf, 1e-9)).tolist()
    out["step_s"] = coarse["step_s"]
    out["speedup"] = fine["step_s"] / coarse["step_s"]
    return out
