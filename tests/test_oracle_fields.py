"""Pins the oracle's field integrands (CalculateFieldsFast.cxx, CalculateFieldsSlow.cxx,
CalculateFieldsMethods.cxx) without trusting the oracle's own formulas:
  - closed forms: filament A and B via complete elliptic integrals; on-axis B of thin solenoid, pancake
    and thick coil;
  - brute force: off-axis B of thin and rectangular coils by adaptive quadrature (scipy) of the loop field;
  - consistency between the three independent integrands: B = curl A and the gradient equals the central
    finite difference of B (catches a dropped term, wrong sign or swapped index in any of them);
  - invariants: rotation covariance, linearity in current, dipole far field, div B = 0.
The quadrature orders are high enough that quadrature error sits far below the tolerances."""
import numpy as np
import pytest
from scipy import integrate
from scipy.special import ellipe, ellipk

import oracle as O

MU0 = 4e-7 * np.pi


def loop_B(R, I, r, z):
    """(B_r, B_z) of a circular filament loop of radius R (Jackson 5.37 / standard elliptic form)."""
    if r == 0.0:
        return 0.0, MU0 * I * R ** 2 / (2 * (R ** 2 + z ** 2) ** 1.5)
    m = 4 * R * r / ((R + r) ** 2 + z ** 2)
    K, E = ellipk(m), ellipe(m)
    pre = MU0 * I / (2 * np.pi * np.sqrt((R + r) ** 2 + z ** 2))
    bz = pre * (K + (R ** 2 - r ** 2 - z ** 2) / ((R - r) ** 2 + z ** 2) * E)
    br = pre * z / r * (-K + (R ** 2 + r ** 2 + z ** 2) / ((R - r) ** 2 + z ** 2) * E)
    return br, bz


def loop_A(R, I, r, z):
    m = 4 * R * r / ((R + r) ** 2 + z ** 2)
    k = np.sqrt(m)
    return MU0 * I / (np.pi * k) * np.sqrt(R / r) * ((1 - m / 2) * ellipk(m) - ellipe(m))


def cyl_to_cart(p, br, bz):
    r = np.hypot(p[0], p[1])
    if r == 0:
        return np.array([0.0, 0.0, bz])
    return np.array([br * p[0] / r, br * p[1] / r, bz])


HI = np.array([4, 1, 1, 100, 1, 1], dtype=np.int32)   # 400 angular nodes
PTS = np.array([[0.2, 0.1, 0.15], [0.5, -0.3, 0.4], [-0.05, 0.02, -0.3], [1.0, 0.5, -0.7], [0.0, 0.0, 0.25]])


def test_filament_B_and_A_closed_form():
    R, N, I = 0.3, 7, 2.5
    c = O.make_coil(R, 0.0, 0.0, N, I)
    a, b, _ = O.coil_fields(c, HI, PTS, True, True, False)
    for p, av, bv in zip(PTS, a, b):
        r, z = np.hypot(p[0], p[1]), p[2]
        ref = cyl_to_cart(p, *loop_B(R, N * I, r, z))
        assert np.linalg.norm(bv - ref) <= 1e-12 * np.linalg.norm(ref)
        if r > 0:
            aphi = loop_A(R, N * I, r, z)
            aref = np.array([-aphi * p[1] / r, aphi * p[0] / r, 0.0])
            assert np.linalg.norm(av - aref) <= 1e-12 * np.linalg.norm(aref)


def test_on_axis_closed_forms():
    I, N = 1.5, 300
    zs = np.array([-0.4, -0.05, 0.0, 0.03, 0.2, 0.9])
    pts = np.stack([np.zeros_like(zs), np.zeros_like(zs), zs], axis=1)
    args = np.array([1, 2, 2, 40, 80, 80], dtype=np.int32)
    # thin solenoid R, length b
    R, bl = 0.2, 0.3
    _, B, _ = O.coil_fields(O.make_coil(R, 0.0, bl, N, I), args, pts, False, True, False)
    n = N / bl
    ref = MU0 * n * I / 2 * ((zs + bl / 2) / np.sqrt(R ** 2 + (zs + bl / 2) ** 2)
                             - (zs - bl / 2) / np.sqrt(R ** 2 + (zs - bl / 2) ** 2))
    assert np.max(np.abs(B[:, 2] - ref) / np.abs(ref)) < 1e-12
    # pancake R1..R2 at z: mu0 N I/(2a) [ln((R2+s2)/(R1+s1)) - R2/s2 + R1/s1]
    R1, a = 0.1, 0.15
    R2 = R1 + a
    _, B, _ = O.coil_fields(O.make_coil(R1, a, 0.0, N, I), args, pts, False, True, False)
    s1, s2 = np.sqrt(R1 ** 2 + zs ** 2), np.sqrt(R2 ** 2 + zs ** 2)
    ref = MU0 * N * I / (2 * a) * (np.log((R2 + s2) / (R1 + s1)) - R2 / s2 + R1 / s1)
    assert np.max(np.abs(B[:, 2] - ref) / np.abs(ref)) < 1e-10
    # thick coil: mu0 J/2 [t ln((R2+sqrt(R2^2+t^2))/(R1+sqrt(R1^2+t^2)))]_{t=z-b/2}^{z+b/2}
    R1, a, bl = 0.1, 0.1, 0.2
    R2 = R1 + a
    _, B, _ = O.coil_fields(O.make_coil(R1, a, bl, N, I), args, pts, False, True, False)
    J = N * I / (a * bl)
    f = lambda t: t * np.log((R2 + np.sqrt(R2 ** 2 + t ** 2)) / (R1 + np.sqrt(R1 ** 2 + t ** 2)))
    ref = MU0 * J / 2 * (f(zs + bl / 2) - f(zs - bl / 2))
    assert np.max(np.abs(B[:, 2] - ref) / np.abs(ref)) < 1e-10


def test_off_axis_brute_force():
    """Thin solenoid and rectangular coil off axis: adaptive quadrature of the loop field over the winding."""
    I, N = 1.0, 100
    pts = np.array([[0.35, 0.0, 0.12], [0.05, 0.0, -0.25], [0.6, 0.0, 0.6]])
    args = np.array([2, 2, 2, 80, 60, 60], dtype=np.int32)
    R, bl = 0.2, 0.3
    _, B, _ = O.coil_fields(O.make_coil(R, 0.0, bl, N, I), args, pts, False, True, False)
    for p, bv in zip(pts, B):
        r = p[0]
        br = integrate.quad(lambda zp: loop_B(R, N * I / bl, r, p[2] - zp)[0], -bl / 2, bl / 2, epsabs=0,
                            epsrel=1e-13)[0]
        bz = integrate.quad(lambda zp: loop_B(R, N * I / bl, r, p[2] - zp)[1], -bl / 2, bl / 2, epsabs=0,
                            epsrel=1e-13)[0]
        ref = np.array([br, 0.0, bz])
        assert np.linalg.norm(bv - ref) <= 1e-10 * np.linalg.norm(ref)
    R1, a, bl = 0.1, 0.08, 0.12
    _, B, _ = O.coil_fields(O.make_coil(R1, a, bl, N, I), args, pts, False, True, False)
    J = N * I / (a * bl)
    for p, bv in zip(pts, B):
        r = p[0]
        comp = []
        for k in (0, 1):
            comp.append(integrate.dblquad(lambda zp, rho: loop_B(rho, J, r, p[2] - zp)[k], R1, R1 + a,
                                          -bl / 2, bl / 2, epsabs=0, epsrel=1e-11)[0])
        ref = np.array([comp[0], 0.0, comp[1]])
        assert np.linalg.norm(bv - ref) <= 1e-9 * np.linalg.norm(ref)


TYPES = [(0.15, 0.05, 0.08), (0.15, 0.0, 0.08), (0.15, 0.05, 0.0), (0.15, 0.0, 0.0)]


@pytest.mark.parametrize("geom", TYPES, ids=["rect", "thin", "flat", "filament"])
def test_gradient_and_curl_by_finite_differences(geom):
    R, a, b = geom
    c = O.make_coil(R, a, b, 50, 1.3, pos=(0.05, -0.1, 0.2), alpha=0.7, beta=1.9)
    args = np.array([1, 1, 1, 80, 40, 40], dtype=np.int32)
    pts = np.array([[0.45, 0.2, 0.5], [-0.3, 0.4, -0.1], [0.1, -0.5, 0.8]])
    _, _, G = O.coil_fields(c, args, pts, False, False, True)
    T = np.array(c.T).reshape(3, 3)
    h = 1e-5
    for p, g in zip(pts, G):
        jac = np.zeros((3, 3))
        curlA = np.zeros(3)
        dA = np.zeros((3, 3))
        for j in range(3):
            e = np.zeros(3)
            e[j] = h
            ap, bp, _ = O.coil_fields(c, args, np.array([p + e]), True, True, False)
            am, bm, _ = O.coil_fields(c, args, np.array([p - e]), True, True, False)
            jac[:, j] = (bp[0] - bm[0]) / (2 * h)
            dA[:, j] = (ap[0] - am[0]) / (2 * h)
        # reference convention: G = T G_local; the true Jacobian is T G_local T^T = G T^T
        assert np.linalg.norm(g @ T.T - jac) <= 1e-6 * np.linalg.norm(jac)
        curlA = np.array([dA[2, 1] - dA[1, 2], dA[0, 2] - dA[2, 0], dA[1, 0] - dA[0, 1]])
        _, b0, _ = O.coil_fields(c, args, np.array([p]), False, True, False)
        assert np.linalg.norm(curlA - b0[0]) <= 1e-6 * np.linalg.norm(b0[0])
        assert abs(np.trace(jac)) <= 1e-6 * np.linalg.norm(jac)


def test_rotation_covariance_and_linearity():
    args = np.array([1, 1, 1, 50, 20, 20], dtype=np.int32)
    c0 = O.make_coil(0.2, 0.05, 0.1, 80, 2.0)
    pos, al, be = np.array([0.3, -0.2, 0.1]), 1.1, 0.4
    c1 = O.make_coil(0.2, 0.05, 0.1, 80, 2.0, pos=tuple(pos), alpha=al, beta=be)
    c2 = O.make_coil(0.2, 0.05, 0.1, 80, -6.0, pos=tuple(pos), alpha=al, beta=be)
    T = np.array(c1.T).reshape(3, 3)
    pts = np.array([[0.9, 0.4, -0.5], [-0.6, 0.1, 0.7]])
    local = (pts - pos) @ np.array(c1.Tinv).reshape(3, 3).T
    _, B0, _ = O.coil_fields(c0, args, local, False, True, False)
    _, B1, _ = O.coil_fields(c1, args, pts, False, True, False)
    _, B2, _ = O.coil_fields(c2, args, pts, False, True, False)
    assert np.allclose(B1, B0 @ T.T, rtol=1e-12, atol=0)
    assert np.allclose(B2, -3.0 * B1, rtol=1e-11, atol=0)
    assert np.allclose(T @ T.T, np.eye(3), atol=1e-15)
    assert abs(np.linalg.det(T) - 1) < 1e-15
    # T e_z is the coil axis (sin a cos b, sin a sin b, cos a)
    assert np.allclose(T[:, 2], [np.sin(al) * np.cos(be), np.sin(al) * np.sin(be), np.cos(al)], atol=1e-15)


def test_dipole_far_field():
    R, a, b, N, I = 0.1, 0.04, 0.05, 120, 3.0
    c = O.make_coil(R, a, b, N, I)
    args = np.array([1, 1, 1, 40, 20, 20], dtype=np.int32)
    m = np.pi * N * I * (R * R + R * a + a * a / 3)   # CalculateAttributes.cxx:12-16
    for d in (30.0, 100.0):
        p = d * np.array([0.48, -0.6, 0.64])
        _, B, _ = O.coil_fields(c, args, np.array([p]), False, True, False)
        rhat = p / d
        mv = np.array([0, 0, m])
        ref = MU0 / (4 * np.pi) * (3 * np.dot(mv, rhat) * rhat - mv) / d ** 3
        assert np.linalg.norm(B[0] - ref) <= 5 * (R / d) ** 2 * np.linalg.norm(ref)


def test_reference_fp64_gradient_conditioning():
    """The reference's own double gradient integrands (CalculateFieldsMethods.cxx, restated by the oracle in
    double) differ from the long double evaluation of the same rule by well over 1e-7 near the winding of a
    rectangular coil; the accuracy claims of the device path are therefore made against the long double truth."""
    import helpers as H
    import workloads as W
    params, pts = W.single_rect_workload(65536, seed=0)
    pts = pts[:2048]
    _, args = H.gpu_args_rows(params)
    c = H.oracle_coil(params[0])
    _, _, g = O.coil_fields(c, args[0], pts, False, False, True, axis_ratio=O.AXIS_RATIO_GPU, threads=8)
    _, _, gt = O.coil_fields(c, args[0], pts, False, False, True, axis_ratio=O.AXIS_RATIO_GPU, threads=8,
                             extended=True)
    num = np.linalg.norm((g - gt).reshape(len(pts), 9), axis=1)
    e = num / np.linalg.norm(gt.reshape(len(pts), 9), axis=1)
    assert 1e-7 < np.max(e) < 1e-3
    assert np.median(e) < 1e-12
