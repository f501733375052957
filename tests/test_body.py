"""Floating body and piston (CPU): the wave-tank generator and the oracle's FP64
rigid coupling.

The oracle's rigid equations are pinned by the reference's own Test_6DOF
routines (Test_6DOF/equation/acc_alpha.cpp, compiled unmodified into
oracle/_ref), by the product's independent FP64 module, and by closed forms
of the midpoint scheme (constant acceleration is integrated exactly, a
torque-free principal spin keeps w).  The coupled step is pinned by free fall
(no fluid within reach: c(t) = c0 + g t^2 / 2 exactly), the piston's closed
form x0 + d(t), Archimedes' force on a submerged box, and Newton's third law
(with no fixed walls and g = 0 the fluid's and the body's forces cancel).
"""
import ctypes as C
import os

import numpy as np
import pytest

import oracle
import zsph_amd as z
from zsph_amd import BODY, FLUID, PISTON, WALL, _abi as abi, cases

REF = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref", "libref6dof.so")

LO, HI = (0.64, 0.12, 0.24), (0.88, 0.28, 0.44)  # box on the dx = 0.04 / 0.02 lattices, draft 0.12


def tank(dx=0.04, lo=LO, hi=HI, **kw):
    return z.Case.wavetank(dx, 1.6, 0.4, 0.6, 0.36, lo, hi, layers=3, piston_travel=0.04, **kw)


def subset_case(case, keep, **json_fields):
    """A case holding only the particles `keep` of `case`, with json fields changed."""
    j = abi.Json()
    C.pointer(j)[0] = case.json
    for k, v in json_fields.items():
        setattr(j, k, v)
    host = z.HostArrays(int(keep.sum()))
    for k in ("pos", "vel", "rhop", "acc", "type", "table", "hash"):
        getattr(host, k)[...] = getattr(case.host, k)[keep]
    j.ptc_num = host.n
    j.water_ptc_num = int((host.type == FLUID).sum())
    return z.Case(j, host)


def test_wavetank_geometry():
    dx = 0.04
    c = tank(dx)
    t, p = c.host.type, c.host.pos
    nx, ny, nz, nd, L, tc = 40, 10, 15, 9, 3, 1
    box = (6, 4, 5)
    in_fluid_box = 6 * 4 * 3  # box cells below the water line
    assert (t == FLUID).sum() == nx * ny * nd - in_fluid_box
    shell = np.prod(box) - max(0, box[0] - 2 * L) * max(0, box[1] - 2 * L) * max(0, box[2] - 2 * L)
    assert (t == BODY).sum() == shell
    assert (t == PISTON).sum() == L * nz * ny
    floor = L * (ny + 2 * L) * (nx + 2 * L + tc)
    far = L * nz * (ny + 2 * L)
    sides = 2 * L * nz * (nx + L + tc)
    assert (t == WALL).sum() == floor + far + sides
    assert c.n == c.json.ptc_num and c.json.rigid_ptc_num == shell
    # box particles inside [lo, hi], no fluid there; piston behind x = 0
    b = p[t == BODY]
    assert (b >= np.float32(LO) - 1e-6).all() and (b <= np.float32(HI) + 1e-6).all()
    f = p[t == FLUID]
    inside = np.all((f > np.float32(LO)) & (f < np.float32(HI)), axis=1)
    assert not inside.any()
    assert (p[t == PISTON][:, 0] < 0).all()
    # hydrostatic start: p = rho0 |g| (H - z), rho from the Tait EOS
    H = nd * dx
    np.testing.assert_allclose(c.host.rhop[t == FLUID, 1], 1000 * 9.81 * (H - f[:, 2]), rtol=1e-5, atol=1e-2)
    assert c.json.c0 == pytest.approx(10 * np.sqrt(2 * 9.81 * H), rel=1e-6)


def random_rigid(seed):
    rng = np.random.default_rng(seed)
    A = rng.normal(size=(3, 3))
    I = A @ A.T + 3 * np.eye(3)
    q = rng.normal(size=4)
    r = z.rigid_state(rng.normal(size=3), 2.5, I, vel=rng.normal(size=3), omega=rng.normal(size=3),
                      q=q / np.linalg.norm(q))
    r.F[:] = list(rng.normal(size=3) * 10)
    r.T[:] = list(rng.normal(size=3) * 10)
    return r


def test_oracle_rigid_accelerations_match_product_module():
    for seed in range(5):
        r = random_rigid(seed)
        p = abi.Rigid.from_buffer_copy(r)
        oracle.rigid_acc_alpha(r, 0.0)
        z.lib().zsph_rigid_acc_alpha(C.byref(p))
        np.testing.assert_allclose(r.acc[:], p.acc[:], rtol=1e-13, atol=1e-13)
        np.testing.assert_allclose(r.alpha[:], p.alpha[:], rtol=1e-11, atol=1e-11)
    # gravity enters as acc_z += g only
    r = random_rigid(7)
    g0 = abi.Rigid.from_buffer_copy(r)
    oracle.rigid_acc_alpha(r, -9.81)
    oracle.rigid_acc_alpha(g0, 0.0)
    assert r.acc[2] - g0.acc[2] == pytest.approx(-9.81, abs=1e-12)


@pytest.mark.skipif(not os.path.exists(REF), reason="oracle/_ref/libref6dof.so not built (needs the reference tree)")
def test_oracle_rigid_accelerations_match_reference_test6dof():
    ref = C.CDLL(REF)
    ref.ref6_cal_acc_alpha.restype = None
    for seed in range(5):
        r = random_rigid(seed)
        st = np.zeros(35, np.float32)
        for k, name in enumerate(("c", "F", "T", "vel", "omega", "acc", "alpha")):
            st[3 * k:3 * k + 3] = getattr(r, name)[:]
        st[21:30] = r.I[:]
        st[30:34] = r.q[:]
        st[34] = r.mass
        # both from the FP32-rounded state
        for k, name in enumerate(("c", "F", "T", "vel", "omega")):
            getattr(r, name)[:] = [float(v) for v in st[3 * k:3 * k + 3]]
        r.I[:] = [float(v) for v in st[21:30]]
        r.q[:] = [float(v) for v in st[30:34]]
        ref.ref6_cal_acc_alpha(st.ctypes.data_as(C.POINTER(C.c_float)))
        oracle.rigid_acc_alpha(r, 0.0)
        np.testing.assert_allclose(r.acc[:], st[15:18], rtol=1e-5, atol=1e-5)
        np.testing.assert_allclose(r.alpha[:], st[18:21], rtol=1e-4, atol=1e-4 * (1 + np.abs(st[18:21]).max()))


def test_oracle_midpoint_closed_forms():
    # constant acceleration: predictor (dt/2) + corrector (dt) is exact
    r = z.rigid_state((0.1, 0.2, 0.3), 2.0, np.eye(3), vel=(1.0, -2.0, 0.5))
    a = np.array([0.3, -0.7, -9.81])
    dt, n = 1e-3, 200
    c0, v0 = np.array(r.c[:]), np.array(r.vel[:])
    for _ in range(n):
        r.acc[:] = list(a)
        mid = oracle.rigid_advance(r, r, dt / 2)
        mid.acc[:] = list(a)
        r = oracle.rigid_advance(r, mid, dt)
    t = n * dt
    np.testing.assert_allclose(r.c[:], c0 + v0 * t + 0.5 * a * t * t, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(r.vel[:], v0 + a * t, rtol=1e-12)
    # torque-free spin about a principal axis: w constant, q = (cos(wt/2), 0, 0, sin(wt/2)) to O(dt^2)
    r = z.rigid_state((0, 0, 0), 1.0, np.diag([1.0, 2.0, 3.0]), omega=(0.0, 0.0, 2.0))
    for _ in range(n):
        oracle.rigid_acc_alpha(r, 0.0)
        mid = oracle.rigid_advance(r, r, dt / 2)
        oracle.rigid_acc_alpha(mid, 0.0)
        r = oracle.rigid_advance(r, mid, dt)
    np.testing.assert_allclose(r.omega[:], [0.0, 0.0, 2.0], atol=1e-14)
    np.testing.assert_allclose(r.q[:], [np.cos(t), 0.0, 0.0, np.sin(t)], atol=1e-6)
    assert np.linalg.norm(r.q[:]) == pytest.approx(1.0, abs=1e-15)


def test_oracle_free_fall_and_piston_closed_forms():
    lo, hi = (0.64, 0.12, 0.48), (0.88, 0.28, 0.56)  # 0.12 above the water: beyond the kernel's reach
    c = tank(0.04, lo, hi)
    o = oracle.Oracle(c.params)
    o.upload(c.host)
    body = cases.box_body(lo, hi, 500.0)
    o.set_body(body)
    A, f = 0.02, 1.5
    o.set_piston(A, f, 0.0, 0)
    x0 = c.host.pos[c.host.type == PISTON].astype(np.float64)
    n = 40
    o.step(n)
    dt = 2.0 * np.float64(np.float32(c.params.half_dt))
    t = o.time()
    assert t == pytest.approx(n * dt, rel=1e-14)
    b = o.get_body()
    c0 = np.array(body.c[:])
    np.testing.assert_allclose(b.c[:], c0 + np.array([0, 0, 0.5 * c.json.g * t * t]), rtol=1e-13, atol=1e-14)
    np.testing.assert_allclose(b.vel[:], [0, 0, c.json.g * t], rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(b.F[:], 0.0, atol=0.0)
    np.testing.assert_allclose(b.q[:], [1, 0, 0, 0], atol=0.0)
    h = o.download(0)
    # body particles: their start positions shifted by the fall (r0 kept in FP32)
    ids = np.arange(c.n)[(c.host.type == BODY)]
    g_order = by_original(h)
    np.testing.assert_allclose(h.pos[g_order[ids]], c.host.pos[ids] + np.float32([0, 0, 0.5 * c.json.g * t * t]),
                               atol=2e-7)
    pid = np.arange(c.n)[c.host.type == PISTON]
    d = A * np.sin(2 * np.pi * f * t)
    np.testing.assert_allclose(h.pos[g_order[pid], 0], x0[:, 0] + d, atol=1e-7)
    np.testing.assert_allclose(h.vel[g_order[pid], 0], A * 2 * np.pi * f * np.cos(2 * np.pi * f * t), rtol=1e-6)
    np.testing.assert_array_equal(h.pos[g_order[pid], 1:], c.host.pos[pid, 1:])


def by_original(h):
    """Index into downloaded arrays by original particle id (the reorder keeps table = original - current)."""
    cur = np.arange(h.n)
    orig = cur + h.table
    out = np.empty(h.n, np.int64)
    out[orig] = cur
    return out


def test_oracle_archimedes_force():
    c = tank(0.02)
    o = oracle.Oracle(c.params)
    o.upload(c.host)
    o.set_body(cases.box_body(LO, HI, 500.0))
    o.hash(0); o.sort_index(); o.sort_data(0, 1); o.boundary(1); o.governing(1)
    F, T = o.body_reduce(1)
    V = np.prod(np.subtract(HI, LO)) * 0.12 / (HI[2] - LO[2])  # submerged volume
    buoy = 1000.0 * 9.81 * V
    assert F[2] == pytest.approx(buoy, rel=0.03)
    assert abs(F[0]) < 1e-4 * buoy and abs(F[1]) < 1e-4 * buoy
    assert np.abs(T).max() < 1e-4 * buoy * 0.1


def test_oracle_body_fluid_forces_cancel():
    # fluid and body only (no fixed walls, no piston), g = 0, compressed fluid:
    # every pair force is antisymmetric, so the momentum and angular momentum
    # changes of fluid and body cancel
    src = tank(0.04)
    keep = (src.host.type == FLUID) | (src.host.type == BODY)
    c = subset_case(src, keep, g=0.0)
    f = c.host.type == FLUID
    rng = np.random.default_rng(3)
    c.host.rhop[f, 0] = np.float32(1000.0 * (1.0 + 0.01 * rng.random(int(f.sum()))))
    c.host.rhop[f, 1] = [np.float32(oracle.eos_p(c.params, float(r))) for r in c.host.rhop[f, 0]]
    o = oracle.Oracle(c.params)
    o.upload(c.host)
    body = cases.box_body(LO, HI, 500.0)
    o.set_body(body)
    o.hash(0); o.sort_index(); o.sort_data(0, 1); o.boundary(1); o.governing(1)
    Fb, Tb = o.body_reduce(1)
    acc = o.get(z.A_ACC, 1)[:, :3]
    pos = o.get(z.A_POSR, 1)[:, :3]
    ff = o.get(z.A_TYPE, 1) == FLUID
    m = c.params.m
    Ff = m * acc[ff].sum(0)
    Tf = np.cross(pos[ff] - np.array(body.c[:]), m * acc[ff]).sum(0)
    scale = m * np.abs(acc[ff]).sum()
    assert np.abs(Fb).max() > 1e-3 * scale  # the body does feel the fluid
    np.testing.assert_allclose(Fb + Ff, 0.0, atol=1e-12 * scale)
    np.testing.assert_allclose(Tb + Tf, 0.0, atol=1e-12 * scale)
