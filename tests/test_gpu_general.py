"""GPU parity of the general pair path and of the group force arrangements, through the drop-in API.

  - The reference's 38 general geometries (data/values_MInductance_general.txt, inputs only: two unrotated coils, the
    secondary displaced by (dr, 0, dz)) through Coil.compute_mutual_inductance and Coil.compute_ampere_force with
    ComputeMethod GPU. Every row has coplanar axes, so the mutual inductance takes Coil::improvedPrecisionCase's
    half-ring rule (SupportFunctions.cxx:124-161, CalculateMInductanceGeneral.cxx:33-36) and the force the full ring
    (CalculateForceGeneral.cxx:37). Compared with the oracle's general path (orc_mutual_inductance_general,
    orc_force_torque_general) on the product's planned CoilPairArguments: relative error <= 1e-10; the drop-in M is
    also checked to be the ccoil_pairs call with the half ring and its offset, not the full ring.
  - CoilGroup.compute_all_force_torque_arrangements (CalculateForceGPU.cxx:11-87, ForceGroup.cu): every member
    against one secondary in several poses, group precision remapped to (pf - 1) / 2 + 1, force scaled by
    2 pi N2 I2, and a member passed as the secondary left out of its own sum (CalculateForceGPU.cxx:41-48). Compared
    with the oracle's GPU pair rule summed over the members: relative error <= 1e-11.
All in CCOIL_FP64.
"""
import os

import numpy as np
import pytest

import ccoil_abi as abi
import helpers as H
import oracle as O
import workloads as W

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def cc():
    import c_coil
    c_coil.coil.Coil.set_gpu_precision(c_coil.coil.GPU_FP64)
    yield c_coil
    c_coil.coil.Coil.set_gpu_precision(c_coil.coil.GPU_FP32C)


def _a12(args):
    pp, sp = args.primary_precision, args.secondary_precision
    return [pp.angular_block_count, pp.thickness_block_count, pp.length_block_count, pp.angular_increment_count,
            pp.thickness_increment_count, pp.length_increment_count, sp.angular_block_count,
            sp.thickness_block_count, sp.length_block_count, sp.angular_increment_count,
            sp.thickness_increment_count, sp.length_increment_count]


def test_reference_general_geometries(cc, ctx):
    C = cc.coil
    rows = np.loadtxt(os.path.join(GOLDEN, "values_MInductance_general.txt"))
    assert rows.shape == (38, 10)
    half_vs_full = []
    for k, (R1, a1, b1, N1, R2, a2, b2, N2, dz, dr) in enumerate(rows):
        p = dict(R=R1, a=a1, b=b1, N=N1, I=1.0, pos=np.zeros(3), alpha=0.0, beta=0.0)
        s = dict(R=R2, a=a2, b=b2, N=N2, I=1.0, pos=np.array([dr, 0.0, dz]), alpha=0.0, beta=0.0)
        cp, cs = W.make_coil(p), W.make_coil(s)
        args = C.CoilPairArguments.get_appropriate_coil_pair_arguments(cp, cs, C.PrecisionFactor(5.0), C.GPU,
                                                                      False, False)
        a12 = _a12(args)
        op, os_ = H.oracle_coil(p), H.oracle_coil(s)
        m_ref = O.mutual_inductance_general(op, os_, a12)
        m = C.Coil.compute_mutual_inductance(cp, cs, args, C.GPU)
        assert abs(m - m_ref) <= 1e-10 * abs(m_ref), (k, m, m_ref)
        F_ref, T_ref = O.force_torque_general(op, os_, a12)
        F, T = C.Coil.compute_ampere_force(cp, cs, args, C.GPU)
        F, T = np.array([F.x, F.y, F.z]), np.array([T.x, T.y, T.z])
        # rows with dz = 0 have no axial force by symmetry: the scale is the pair's largest force component
        fscale = max(np.linalg.norm(F_ref), 1e-300)
        assert np.linalg.norm(F - F_ref) <= 1e-10 * fscale, (k, F, F_ref)
        assert np.linalg.norm(T - T_ref) <= 1e-10 * max(np.linalg.norm(T_ref), fscale * (R1 + R2)), (k, T, T_ref)
        # the same pair through the C ABI with and without the coplanar half ring (offset 0: the secondary is
        # displaced along x, SupportFunctions.cxx:150-157)
        P = abi.CoilSet.from_rows([W.coil_row(cp, args.primary_precision)])
        S = abi.CoilSet.from_rows([W.coil_row(cs, args.secondary_precision)])
        half, _ = ctx.pairs(P, S, 1, abi.PAIR_M, abi.FP64, ring_offset=[0.0], half_ring=[1])
        full, _ = ctx.pairs(P, S, 1, abi.PAIR_M, abi.FP64, ring_offset=[0.0], half_ring=[0])
        assert half["M"][0] == m, k
        half_vs_full.append(abs(full["M"][0] - m) / abs(m))
    print(f"\nhalf ring vs full ring, relative difference: max {max(half_vs_full):.3g}")
    assert max(half_vs_full) > 1e-9   # the two rules are different quadratures of the same integral


def test_group_force_torque_arrangements(cc):
    C, V = cc.coil, cc.tensor.Vector3
    params, _ = W.group_workload(5, 8, seed=48)
    group = cc.coil_group.CoilGroup()
    for p in params:
        group.add_coil(p["R"], p["a"], p["b"], int(p["N"]), p["I"], C.PrecisionFactor(5.0), 8, V(*p["pos"]),
                       p["alpha"], p["beta"])
    ps_ = dict(R=0.07, a=0.02, b=0.04, N=90, I=2.5, pos=np.zeros(3), alpha=0.0, beta=0.0)
    sec = C.Coil(ps_["R"], ps_["a"], ps_["b"], ps_["N"], ps_["I"])
    rng = np.random.default_rng(49)
    n = 7
    spos = rng.uniform(1.5, 2.0, (n, 3)) * rng.choice([-1.0, 1.0], size=(n, 3))   # outside the group's cube
    sY, sZ = list(rng.uniform(0, 3.0, n)), list(rng.uniform(0, 6.28, n))
    va = cc.tensor.Vector3Array()
    for q in spos:
        va.append(float(q[0]), float(q[1]), float(q[2]))
    pf = 7.0
    gpf = (pf - 1.0) / 2.0 + 1.0
    out = group.compute_all_force_torque_arrangements(sec, va, sY, sZ, C.PrecisionFactor(pf), C.GPU)
    sec_args = O.secondary_precision_args_gpu(H.oracle_coil(ps_), gpf)

    def reference(members, sp, s_params, k):
        q = dict(s_params)
        q["pos"], q["alpha"], q["beta"] = sp[k], sY[k], sZ[k]
        os_ = H.oracle_coil(q)
        F, T = np.zeros(3), np.zeros(3)
        for p in members:
            oc = H.oracle_coil(p)
            pa = O.coil_precision_args(oc, gpf, True)
            f, t = O.force_torque_gpu_rule(oc, os_, pa[3:6], sec_args[3:6], 2 * np.pi * s_params["N"] * s_params["I"])
            F += f
            T += t
        return F, T

    for k, (f, t) in enumerate(out):
        F_ref, T_ref = reference(params, spos, ps_, k)
        assert np.linalg.norm(np.array([f.x, f.y, f.z]) - F_ref) <= 1e-11 * np.linalg.norm(F_ref), k
        assert np.linalg.norm(np.array([t.x, t.y, t.z]) - T_ref) <= 1e-11 * np.linalg.norm(T_ref), k

    # a member as the secondary: its own pair is left out (CalculateForceGPU.cxx:41-48)
    member = group[0]
    p0 = params[0]
    mpos = spos + 0.25
    va2 = cc.tensor.Vector3Array()
    for q in mpos:
        va2.append(float(q[0]), float(q[1]), float(q[2]))
    out2 = group.compute_all_force_torque_arrangements(member, va2, sY, sZ, C.PrecisionFactor(pf), C.GPU)
    sec_args = O.secondary_precision_args_gpu(H.oracle_coil(p0), gpf)
    for k, (f, t) in enumerate(out2):
        F_ref, T_ref = reference(params[1:], mpos, p0, k)
        assert np.linalg.norm(np.array([f.x, f.y, f.z]) - F_ref) <= 1e-11 * np.linalg.norm(F_ref), k
        assert np.linalg.norm(np.array([t.x, t.y, t.z]) - T_ref) <= 1e-11 * np.linalg.norm(T_ref), k
