"""GPU parity of the drop-in C++ / Python API (c_coil with ComputeMethod GPU) against the oracle.

Covers the reference's public GPU entry points:
  - Coil.compute_all_{A_potential,B_field,E_field}_vectors and compute_all_B_gradient_matrices (Coil.h:410-492);
  - CoilGroup.compute_all_B_field_vectors / compute_all_B_gradient_matrices (CoilGroup.h:96-133);
  - Coil.compute_all_{mutual_inductance,force_torque}_arrangements (Coil.h:638-675), the pure-GPU rule with the
    reference's force scale 2 pi N2 / I1 (CalculateForceArrangements.cxx:125);
  - CoilGroup.compute_all_mutual_inductance_arrangements (CoilGroup.h:194-223) with the group precision remapped
    to (pf - 1) / 2 + 1 (CalculateMInductanceGPU.cxx:18).
All in CCOIL_FP64 against the oracle's long double (fields) or double (pairs) evaluation of the same rule:
relative error <= 1e-11 for A, B, E, M, F, T and <= 1e-7 for the gradient (near-winding conditioning).
"""
import numpy as np
import pytest

import helpers as H
import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cc():
    import c_coil
    c_coil.coil.Coil.set_gpu_precision(c_coil.coil.GPU_FP64)
    yield c_coil
    c_coil.coil.Coil.set_gpu_precision(c_coil.coil.GPU_FP32C)


def _va(cc, pts):
    va = cc.tensor.Vector3Array()
    for p in pts:
        va.append(float(p[0]), float(p[1]), float(p[2]))
    return va


def _np3(va):
    return np.stack([np.array(va.x()), np.array(va.y()), np.array(va.z())], axis=1)


def _np9(ma):
    return np.stack([np.array(getattr(ma, k)()) for k in ("xx", "xy", "xz", "yx", "yy", "yz", "zx", "zy", "zz")],
                    axis=1)


def _args6(a):
    return np.array([1, 1, 1, a.angular_increment_count, a.thickness_increment_count, a.length_increment_count],
                    dtype=np.int32)


def test_coil_fields(cc):
    C, V = cc.coil, cc.tensor.Vector3
    p = dict(R=0.1, a=0.05, b=0.08, N=100, I=2.0, freq=60.0, pos=np.array([0.1, -0.05, 0.02]), alpha=0.4,
             beta=1.1)
    coil = C.Coil(p["R"], p["a"], p["b"], p["N"], p["I"], p["freq"], C.PrecisionFactor(5.0), 8, V(*p["pos"]),
                  p["alpha"], p["beta"])
    rng = np.random.default_rng(41)
    pts = rng.uniform(-0.5, 0.5, size=(512, 3))
    va = _va(cc, pts)
    A = _np3(coil.compute_all_A_potential_vectors(va, C.GPU))
    B = _np3(coil.compute_all_B_field_vectors(va, C.GPU))
    E = _np3(coil.compute_all_E_field_vectors(va, C.GPU))
    G = _np9(coil.compute_all_B_gradient_matrices(va, C.GPU))
    oc = H.oracle_coil(p)
    args = _args6(coil.get_precision_settings_GPU())
    a, b, g = O.coil_fields(oc, args, pts, True, True, True, axis_ratio=O.AXIS_RATIO_GPU, threads=8, extended=True)
    assert np.nanmax(H.rel_err_vec(A, a)) < 1e-11
    assert np.nanmax(H.rel_err_vec(B, b)) < 1e-11
    assert np.nanmax(H.rel_err_vec(E, 2 * np.pi * p["freq"] * a)) < 1e-11
    assert np.nanmax(H.rel_err_vec(G, g.reshape(-1, 9))) < 1e-7


def test_group_fields(cc):
    import workloads as W
    C, V = cc.coil, cc.tensor.Vector3
    params, pts = W.group_workload(10, 512, seed=42)
    group = cc.coil_group.CoilGroup()
    for p in params:
        group.add_coil(p["R"], p["a"], p["b"], int(p["N"]), p["I"], C.PrecisionFactor(5.0), 8, V(*p["pos"]),
                       p["alpha"], p["beta"])
    va = _va(cc, pts)
    B = _np3(group.compute_all_B_field_vectors(va, C.GPU))
    G = _np9(group.compute_all_B_gradient_matrices(va, C.GPU))
    coils = [H.oracle_coil(p) for p in params]
    args = [_args6(C.PrecisionArguments.get_coil_precision_arguments_GPU(m, C.PrecisionFactor(5.0)))
            for m in group.get_member_coils()]
    _, b, g = O.group_fields(coils, args, pts, False, True, True, axis_ratio=O.AXIS_RATIO_GPU, threads=8,
                             extended=True)
    assert np.nanmax(H.rel_err_vec(B, b)) < 1e-11
    assert np.nanmax(H.rel_err_vec(G, g.reshape(-1, 9))) < 1e-7


def _posed(p, pos, y, z):
    q = dict(p)
    q["pos"], q["alpha"], q["beta"] = np.asarray(pos, dtype=float), y, z
    return H.oracle_coil(q)


def test_single_pair_arrangements(cc):
    C, V = cc.coil, cc.tensor.Vector3
    pp_ = dict(R=0.1, a=0.03, b=0.06, N=200, I=3.0, pos=np.zeros(3), alpha=0.0, beta=0.0)
    ps_ = dict(R=0.06, a=0.02, b=0.04, N=80, I=1.5, pos=np.zeros(3), alpha=0.0, beta=0.0)
    prim = C.Coil(pp_["R"], pp_["a"], pp_["b"], pp_["N"], pp_["I"])
    sec = C.Coil(ps_["R"], ps_["a"], ps_["b"], ps_["N"], ps_["I"])
    rng = np.random.default_rng(43)
    n = 12
    ppos = rng.uniform(-0.05, 0.05, (n, 3))
    spos = ppos + np.column_stack([rng.uniform(-0.1, 0.1, n), rng.uniform(-0.1, 0.1, n), rng.uniform(0.25, 0.4, n)])
    pY, pZ = list(rng.uniform(0, 0.5, n)), list(rng.uniform(0, 6.28, n))
    sY, sZ = list(rng.uniform(0, 0.5, n)), list(rng.uniform(0, 6.28, n))
    M = C.Coil.compute_all_mutual_inductance_arrangements(prim, sec, _va(cc, ppos), _va(cc, spos), pY, pZ, sY, sZ,
                                                         C.PrecisionFactor(5.0), C.GPU)
    FT = C.Coil.compute_all_force_torque_arrangements(prim, sec, _va(cc, ppos), _va(cc, spos), pY, pZ, sY, sZ,
                                                      C.PrecisionFactor(5.0), C.GPU)
    a12 = O.pair_precision_args(H.oracle_coil(pp_), H.oracle_coil(ps_), 5.0, True, False, True)
    for k in range(n):
        op, os_ = _posed(pp_, ppos[k], pY[k], pZ[k]), _posed(ps_, spos[k], sY[k], sZ[k])
        m = O.mutual_inductance_gpu_rule(op, os_, a12[3:6], a12[9:12])
        F, T = O.force_torque_gpu_rule(op, os_, a12[3:6], a12[9:12], 2 * np.pi * ps_["N"] / pp_["I"])
        assert abs(M[k] - m) <= 1e-11 * abs(m), k
        f, t = FT[k]
        assert np.linalg.norm(np.array([f.x, f.y, f.z]) - F) <= 1e-11 * np.linalg.norm(F), k
        assert np.linalg.norm(np.array([t.x, t.y, t.z]) - T) <= 1e-11 * np.linalg.norm(T), k


def test_group_arrangements_precision_remap(cc):
    import workloads as W
    C, V = cc.coil, cc.tensor.Vector3
    params, _ = W.group_workload(4, 8, seed=44)
    group = cc.coil_group.CoilGroup()
    for p in params:
        group.add_coil(p["R"], p["a"], p["b"], int(p["N"]), p["I"], C.PrecisionFactor(5.0), 8, V(*p["pos"]),
                       p["alpha"], p["beta"])
    ps_ = dict(R=0.08, a=0.02, b=0.05, N=60, I=1.0, pos=np.zeros(3), alpha=0.0, beta=0.0)
    sec = C.Coil(ps_["R"], ps_["a"], ps_["b"], ps_["N"], ps_["I"])
    rng = np.random.default_rng(45)
    n = 6
    spos = rng.uniform(1.5, 2.0, (n, 3))                  # outside the group's cube
    sY, sZ = list(rng.uniform(0, 3.0, n)), list(rng.uniform(0, 6.28, n))
    pf = 7.0
    M = group.compute_all_mutual_inductance_arrangements(sec, _va(cc, spos), sY, sZ, C.PrecisionFactor(pf), C.GPU)
    gpf = (pf - 1.0) / 2.0 + 1.0
    sec_args = O.secondary_precision_args_gpu(H.oracle_coil(ps_), gpf)
    for k in range(n):
        os_ = _posed(ps_, spos[k], sY[k], sZ[k])
        ref = 0.0
        for p in params:
            oc = H.oracle_coil(p)
            pa = O.coil_precision_args(oc, gpf, True)
            ref += O.mutual_inductance_gpu_rule(oc, os_, pa[3:6], sec_args[3:6])
        assert abs(M[k] - ref) <= 1e-11 * abs(ref), (k, M[k], ref)


def test_dipole_force_and_torque(cc):
    """CoilGroup.compute_all_force_on_dipole_moments: F = grad(m . B) (true Jacobian), T = m x B, against the oracle's
    long double field and central differences of m . B."""
    import workloads as W
    C, V = cc.coil, cc.tensor.Vector3
    params, _ = W.group_workload(5, 8, seed=46)
    group = cc.coil_group.CoilGroup()
    for p in params:
        group.add_coil(p["R"], p["a"], p["b"], int(p["N"]), p["I"], C.PrecisionFactor(5.0), 8, V(*p["pos"]),
                       p["alpha"], p["beta"])
    rng = np.random.default_rng(47)
    pts = rng.uniform(1.6, 2.2, size=(16, 3)) * rng.choice([-1.0, 1.0], size=(16, 3))
    mom = rng.standard_normal((16, 3))
    out = group.compute_all_force_on_dipole_moments(_va(cc, pts), _va(cc, mom), C.GPU)
    coils = [H.oracle_coil(p) for p in params]
    args = [_args6(C.PrecisionArguments.get_coil_precision_arguments_GPU(m, C.PrecisionFactor(5.0)))
            for m in group.get_member_coils()]
    J = np.zeros((len(pts), 3, 3))
    for c, a in zip(coils, args):
        _, _, g = O.coil_fields(c, a, pts, False, False, True, axis_ratio=O.AXIS_RATIO_GPU, threads=4, extended=True)
        J += g @ np.array(c.T).reshape(3, 3).T
    _, b, _ = O.group_fields(coils, args, pts, False, True, False, axis_ratio=O.AXIS_RATIO_GPU, threads=4,
                             extended=True)
    h = 1e-6
    for k, (f, t) in enumerate(out):
        F = np.array([f.x, f.y, f.z])
        Tq = np.array([t.x, t.y, t.z])
        assert np.linalg.norm(F - J[k].T @ mom[k]) <= 1e-11 * np.linalg.norm(F)
        assert np.linalg.norm(Tq - np.cross(mom[k], b[k])) <= 1e-11 * np.linalg.norm(Tq)
        fd = np.zeros(3)
        for ax in range(3):
            e = np.zeros(3)
            e[ax] = h
            _, bp, _ = O.group_fields(coils, args, (pts[k] + e)[None], False, True, False,
                                      axis_ratio=O.AXIS_RATIO_GPU, extended=True)
            _, bm, _ = O.group_fields(coils, args, (pts[k] - e)[None], False, True, False,
                                      axis_ratio=O.AXIS_RATIO_GPU, extended=True)
            fd[ax] = mom[k] @ (bp[0] - bm[0]) / (2 * h)
        assert np.linalg.norm(F - fd) <= 1e-7 * np.linalg.norm(fd), (k, F, fd)
