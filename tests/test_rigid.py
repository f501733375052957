"""FP64 rigid-body module (Test_6DOF equations), CPU only.

Pinned by the reference's own routines (Test_6DOF/equation/*.cpp and
Test_6DOF/lib/math.cpp, compiled unmodified into oracle/_ref/libref6dof.so
by `make -C oracle ref`), by the analytic inertia of the reference driver's
21^3 lattice (Test_6DOF/main/main.cpp:18-40) and by closed-form invariants.
"""
import ctypes as C
import os

import numpy as np
import pytest

import zsph_amd as z
from zsph_amd import _abi as abi

REF = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref", "libref6dof.so")


def lattice21():
    i = np.arange(21, dtype=np.float32) * np.float32(0.1)
    X, Y, Zc = np.meshgrid(i, i, i, indexing="ij")
    return np.stack([X.ravel(), Y.ravel(), Zc.ravel()], 1).astype(np.float32)


def fptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def rigid_from(pts, mass):
    r = abi.Rigid()
    assert z.lib().zsph_rigid_mass_properties(fptr(pts), len(pts), mass, C.byref(r)) == abi.OK
    r.q[0] = 1.0
    return r


def test_lattice_mass_properties_closed_form():
    # main.cpp:18-40: c = (1,1,1), I = M * 2 * E[(0.1k - 1)^2] = 10 * 2 * 7.7/21 on the diagonal
    r = rigid_from(lattice21(), 10.0)
    np.testing.assert_allclose(list(r.c), [1.0, 1.0, 1.0], atol=1e-6)
    I = np.array(r.I).reshape(3, 3)
    np.testing.assert_allclose(np.diag(I), 10 * 2 * 7.7 / 21, rtol=1e-6)
    np.testing.assert_allclose(I - np.diag(np.diag(I)), 0.0, atol=1e-6)


def test_torque_free_principal_spin_is_steady():
    r = rigid_from(lattice21(), 10.0)
    r.omega[2] = 3.0
    for _ in range(200):
        z.lib().zsph_rigid_acc_alpha(C.byref(r))
        z.lib().zsph_rigid_euler_step(C.byref(r), None, 0, 1e-3)
    np.testing.assert_allclose(list(r.omega), [0.0, 0.0, 3.0], atol=1e-12)
    assert np.linalg.norm(list(r.q)) == pytest.approx(1.0, abs=1e-14)
    # rotation angle 3 rad/s * 0.2 s about z: q = (cos 0.3, 0, 0, sin 0.3) up to Euler error
    assert r.q[3] == pytest.approx(np.sin(0.3), abs=1e-3)


def test_constant_force_uniform_acceleration():
    r = rigid_from(lattice21(), 10.0)
    c0 = np.array(r.c)
    r.F[:] = (5.0, -2.0, 1.0)
    dt, n = 1e-3, 100
    for _ in range(n):
        z.lib().zsph_rigid_acc_alpha(C.byref(r))
        z.lib().zsph_rigid_euler_step(C.byref(r), None, 0, dt)
    t = n * dt
    a = np.array([0.5, -0.2, 0.1])
    np.testing.assert_allclose(list(r.vel), a * t, rtol=1e-12)
    # explicit Euler position: sum_{k<n} k dt a dt = a dt^2 n (n-1) / 2
    np.testing.assert_allclose(np.array(r.c) - c0, a * dt * dt * n * (n - 1) / 2, rtol=1e-9, atol=1e-15)


@pytest.mark.skipif(not os.path.exists(REF), reason="oracle/_ref/libref6dof.so not built (needs the reference tree)")
def test_matches_reference_test6dof_routines():
    ref = C.CDLL(REF)
    for n in ("ref6_gravity_center", "ref6_inertia_tensor", "ref6_cal_acc_alpha", "ref6_time_int"):
        getattr(ref, n).restype = None
    pts = lattice21()
    # mass properties
    c_ref = np.zeros(3, np.float32)
    ref.ref6_gravity_center(fptr(pts), len(pts), fptr(c_ref))
    st = np.zeros(35, np.float32)
    st[0:3] = c_ref
    st[30] = 1.0
    st[34] = 10.0
    ref.ref6_inertia_tensor(fptr(pts), len(pts), fptr(st))
    r = rigid_from(pts, 10.0)
    # the reference sums 9261 points in FP32: its mean carries ~6e-6 of rounding
    np.testing.assert_allclose(list(r.c), c_ref, rtol=2e-5)
    np.testing.assert_allclose(list(r.I), st[21:30], rtol=2e-5, atol=2e-5)
    # from here on both integrate from the reference's own c and I
    r.c[:] = [float(v) for v in c_ref]
    r.I[:] = [float(v) for v in st[21:30]]
    # the reference driver's loads (main.cpp:9-10,53-55): F, T modulated each step; dt = 1e-3
    st[3:6] = (1000.0, 1000.0, 0.0)
    st[6:9] = (0.0, 1000.0, 0.0)
    r.F[:] = (1000.0, 1000.0, 0.0)
    r.T[:] = (0.0, 1000.0, 0.0)
    ref_pts = pts.copy()
    our_pts = pts.copy()
    for i in range(300):
        fac = 0.5 - np.sin(np.float32(i) * np.float32(0.01) * np.float32(np.pi))
        facy = 0.5 - np.cos(np.float32(i) * np.float32(0.01) * np.float32(np.pi))
        st[6:9] *= np.float32(fac)
        st[3] *= np.float32(fac)
        st[4] *= np.float32(facy)
        for k in range(3):
            r.T[k] = float(st[6 + k])
        r.F[0], r.F[1], r.F[2] = float(st[3]), float(st[4]), float(st[5])
        ref.ref6_cal_acc_alpha(fptr(st))
        z.lib().zsph_rigid_acc_alpha(C.byref(r))
        np.testing.assert_allclose(list(r.alpha), st[18:21], rtol=1e-4, atol=1e-4 * (1 + np.abs(st[18:21]).max()))
        ref.ref6_time_int(fptr(ref_pts), len(ref_pts), fptr(st))
        z.lib().zsph_rigid_euler_step(C.byref(r), fptr(our_pts), len(our_pts), 1e-3)
    scale = 1 + np.abs(st[0:3]).max()
    np.testing.assert_allclose(list(r.c), st[0:3], atol=1e-4 * scale)
    np.testing.assert_allclose(list(r.vel), st[9:12], atol=1e-4 * (1 + np.abs(st[9:12]).max()))
    np.testing.assert_allclose(list(r.omega), st[12:15], atol=1e-4 * (1 + np.abs(st[12:15]).max()))
    np.testing.assert_allclose(list(r.q), st[30:34], atol=1e-4)
    np.testing.assert_allclose(our_pts, ref_pts, atol=1e-4 * (1 + np.abs(ref_pts).max()))


def test_general_inverse_handles_asymmetric_inertia():
    # the reference's inv() is the transpose of the inverse (math.cpp:245-271);
    # a non-symmetric I shows the difference: alpha must equal I^-1 T for q = 1, w = 0
    r = abi.Rigid()
    r.q[0] = 1.0
    r.mass = 1.0
    A = np.array([[2.0, 0.5, 0.0], [0.1, 3.0, 0.2], [0.0, 0.4, 1.5]])
    r.I[:] = A.ravel()
    r.T[:] = (1.0, 2.0, 3.0)
    z.lib().zsph_rigid_acc_alpha(C.byref(r))
    np.testing.assert_allclose(list(r.alpha), np.linalg.solve(A, [1.0, 2.0, 3.0]), rtol=1e-12)
