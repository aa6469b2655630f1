"""GPU parity of the coaxial (z-axis) rules against the oracle.

  ccoil_pairs_zaxis      Coil::calculateMutualInductanceZAxisFast / calculateAmpereForceZAxisFast
  ccoil_self_inductance  Coil::calculateSelfInductance
  c_coil drop-in         Coil.compute_mutual_inductance / compute_force_torque on coaxial pairs (fast and slow
                         rules), Coil.compute_and_set_self_inductance, CoilGroup.compute_mutual_inductance /
                         compute_force_torque with coaxial and non-coaxial members.

The device evaluates the reference's per-node expressions in double. Its distance to the oracle's double evaluation
(the reference's CPU arithmetic) is bounded by 1e-12 of the value plus four times the double evaluation's own
distance to the long double one: the four length terms of the fast rules cancel, and a last-bit difference of log
or cos is amplified by the same factor in both. Slow rules go through the field kernel: relative error <= 1e-11
against the oracle's long double field.
"""
import numpy as np
import pytest

import ccoil_abi as abi
import oracle as O

pytestmark = pytest.mark.gpu


def _row(c, args6):
    return dict(use_fast=int(c.use_fast), n_ang=int(args6[3]), n_thk=int(args6[4]), n_len=int(args6[5]),
                blocks=tuple(int(v) for v in args6[0:3]), R=c.R, a=c.a, b=c.b, J=c.J, current=c.I, turns=float(c.N),
                frequency=c.freq, pos=tuple(c.pos), alpha=c.alpha, beta=c.beta)


def _close(gpu, d, ld):
    return abs(gpu - d) <= 1e-12 * abs(ld) + 4 * abs(d - ld)


def _coaxial_case(rng):
    def dims():
        thin = rng.random() < 0.3
        return (rng.uniform(0.02, 0.2), 0.0 if thin else rng.uniform(0.005, 0.1), rng.uniform(0.01, 0.3),
                int(rng.integers(1, 500)), rng.uniform(0.5, 3.0) * rng.choice([-1, 1]))
    p = O.make_coil(*dims(), pos=(0, 0, rng.uniform(-0.5, 0.5)))
    s = O.make_coil(*dims(), pos=(0, 0, rng.uniform(-0.5, 0.5)))
    return p, s, O.pair_precision_args(p, s, float(rng.uniform(2.0, 8.0)), False, True, False)


def test_pairs_zaxis_match_oracle(ctx):
    rng = np.random.default_rng(71)
    n = 24
    cases = [_coaxial_case(rng) for _ in range(n)]
    assert any(a[0] > 1 or a[1] > 1 or a[7] > 1 for _, _, a in cases)   # composite (blocked) rules are covered
    P = abi.CoilSet.from_rows([_row(p, a[:6]) for p, _, a in cases])
    S = abi.CoilSet.from_rows([_row(s, a[6:]) for _, s, a in cases])
    out, rep = ctx.pairs_zaxis(P, S, n, abi.PAIR_M | abi.PAIR_FT)
    evals = sum(int(a[0] * a[3] * a[1] * a[4] * a[7] * a[10]) for _, _, a in cases)
    assert rep.node_evaluations == 2 * evals
    for k, (p, s, a) in enumerate(cases):
        m_d, m_ld = O.mutual_inductance_zaxis_fast(p, s, a), O.mutual_inductance_zaxis_fast(p, s, a, True)
        f_d, f_ld = O.force_zaxis_fast(p, s, a), O.force_zaxis_fast(p, s, a, True)
        assert _close(out["M"][k], m_d, m_ld), (k, out["M"][k], m_d, m_ld)
        assert _close(out["Fz"][k], f_d, f_ld), (k, out["Fz"][k], f_d, f_ld)
    # one quantity at a time and indexed pairs give the same values
    idx = np.array([3, 0, 3, 7], dtype=np.int64)
    only_f, _ = ctx.pairs_zaxis(P, S, 4, abi.PAIR_FT, prim_index=idx, sec_index=idx)
    assert np.array_equal(only_f["Fz"], out["Fz"][idx])
    only_m, _ = ctx.pairs_zaxis(P, S, 4, abi.PAIR_M, prim_index=idx, sec_index=idx)
    assert np.array_equal(only_m["M"], out["M"][idx])


def test_pairs_zaxis_device_pointers(ctx):
    import torch
    rng = np.random.default_rng(72)
    cases = [_coaxial_case(rng) for _ in range(5)]
    P = abi.CoilSet.from_rows([_row(p, a[:6]) for p, _, a in cases])
    S = abi.CoilSet.from_rows([_row(s, a[6:]) for _, s, a in cases])
    host, _ = ctx.pairs_zaxis(P, S, 5, abi.PAIR_M | abi.PAIR_FT)
    dev = {"M": torch.zeros(5, dtype=torch.float64, device="cuda"),
           "Fz": torch.zeros(5, dtype=torch.float64, device="cuda")}
    ctx.pairs_zaxis(P, S, 5, abi.PAIR_M | abi.PAIR_FT, flags=abi.DEVICE_POINTERS, out=dev)
    assert np.array_equal(dev["M"].cpu().numpy(), host["M"])
    assert np.array_equal(dev["Fz"].cpu().numpy(), host["Fz"])


def test_self_inductance_matches_oracle(ctx):
    rng = np.random.default_rng(73)
    coils, radial, args = [], [], []
    for k in range(10):
        thin = k % 3 == 0
        c = O.make_coil(rng.uniform(0.02, 0.2), 0.0 if thin else rng.uniform(0.005, 0.1), rng.uniform(0.01, 0.3),
                        int(rng.integers(1, 500)), rng.uniform(0.5, 3.0))
        a = O.pair_precision_args(c, c, float(rng.uniform(2.0, 9.0)), False, True, False)
        coils.append(c)
        args.append(a)
        radial.append((a[7], a[10]))
    C = abi.CoilSet.from_rows([_row(c, a[:6]) for c, a in zip(coils, args)])
    L, rep = ctx.self_inductance(C, np.array(radial))
    for k, (c, a) in enumerate(zip(coils, args)):
        d, ld = O.self_inductance(c, a), O.self_inductance(c, a, True)
        assert _close(L[k], d, ld), (k, L[k], d, ld)


def test_zaxis_argument_errors(ctx):
    rng = np.random.default_rng(74)
    p, s, a = _coaxial_case(rng)
    P, S = abi.CoilSet.from_rows([_row(p, a[:6])]), abi.CoilSet.from_rows([_row(s, a[6:])])
    with pytest.raises(abi.CCoilError):
        ctx.pairs_zaxis(P, S, 1, abi.PAIR_DM, out={"M": np.zeros(1)})
    with pytest.raises(abi.CCoilError):
        ctx.pairs_zaxis(P, S, 2, abi.PAIR_M)            # index 1 out of range
    r = _row(s, a[6:])
    r["current"] = 0.0
    with pytest.raises(abi.CCoilError):
        ctx.pairs_zaxis(P, abi.CoilSet.from_rows([r]), 1, abi.PAIR_M)
    # the force does not divide by the currents
    out, _ = ctx.pairs_zaxis(P, abi.CoilSet.from_rows([r]), 1, abi.PAIR_FT)
    assert np.isfinite(out["Fz"][0])
    with pytest.raises(abi.CCoilError):
        ctx.self_inductance(P, np.array([[1, 101]]))


# ---------------------------------------------------------------------------------------------------------------
# Drop-in API
# ---------------------------------------------------------------------------------------------------------------

@pytest.fixture(scope="module")
def cc():
    import c_coil
    c_coil.coil.Coil.set_gpu_precision(c_coil.coil.GPU_FP64)
    yield c_coil
    c_coil.coil.Coil.set_gpu_precision(c_coil.coil.GPU_FP32C)


def _both(cc, dims, z=0.0, alpha=0.0, beta=0.0, x=0.0):
    C, V = cc.coil, cc.tensor.Vector3
    R, a, b, N, I = dims
    coil = C.Coil(R, a, b, N, I, 50.0, C.PrecisionFactor(5.0), 8, V(x, 0.0, z), alpha, beta)
    return coil, O.make_coil(R, a, b, N, I, 50.0, pos=(x, 0.0, z), alpha=alpha, beta=beta)


def test_dropin_coaxial_fast_pairs(cc):
    C = cc.coil
    for pdims, sdims, z in [((0.1, 0.05, 0.1, 100, 2.0), (0.08, 0.04, 0.06, 50, 1.5), 0.3),
                            ((0.05, 0.0, 0.2, 300, 1.0), (0.12, 0.03, 0.02, 20, -2.0), -0.4)]:
        prim, op = _both(cc, pdims, beta=0.7)       # a rotation about z keeps the pair coaxial
        sec, os_ = _both(cc, sdims, z=z)
        for pf in (3.0, 6.0):
            a = O.pair_precision_args(op, os_, pf, True, True, False)
            M = C.Coil.compute_mutual_inductance(prim, sec, C.PrecisionFactor(pf), C.GPU)
            assert _close(M, O.mutual_inductance_zaxis_fast(op, os_, a), O.mutual_inductance_zaxis_fast(op, os_, a, True))
            F, T = C.Coil.compute_ampere_force(prim, sec, C.PrecisionFactor(pf), C.GPU)
            fz = O.force_zaxis_fast(op, os_, a)
            assert F.x == 0.0 and F.y == 0.0 and T.x == 0.0 and T.y == 0.0 and T.z == 0.0
            assert _close(F.z, fz, O.force_zaxis_fast(op, os_, a, True))


def test_dropin_coaxial_slow_pairs(cc):
    """A FLAT or FILAMENT member makes the pair take the slow rule: the primary's A (or B) at the secondary's r-z
    grid from the device field rule (the primary's increments, one block), as the reference's GPU field path
    evaluates it."""
    C = cc.coil
    for pdims, sdims, z in [((0.1, 0.05, 0.1, 100, 2.0), (0.08, 0.04, 0.0, 50, 1.5), 0.3),     # FLAT secondary
                            ((0.1, 0.0, 0.0, 1, 2.0), (0.07, 0.0, 0.0, 1, 1.0), 0.05)]:        # two loops
        prim, op = _both(cc, pdims)
        sec, os_ = _both(cc, sdims, z=z)
        a = O.pair_precision_args(op, os_, 5.0, True, True, False)
        fa = np.array([1, 1, 1, a[3], a[4], a[5]], dtype=np.int32)
        M = C.Coil.compute_mutual_inductance(prim, sec, C.PrecisionFactor(5.0), C.GPU)
        m = O.mutual_inductance_zaxis_slow(op, os_, a, fa, extended=True)
        assert abs(M - m) <= 1e-11 * abs(m), (M, m)
        F, T = C.Coil.compute_ampere_force(prim, sec, C.PrecisionFactor(5.0), C.GPU)
        f = O.force_zaxis_slow(op, os_, a, fa, extended=True)
        assert abs(F.z - f) <= 1e-11 * abs(f), (F.z, f)
        assert F.x == 0.0 and F.y == 0.0 and T.z == 0.0


def test_dropin_self_inductance(cc):
    C, V = cc.coil, cc.tensor.Vector3
    for dims in [(0.1, 0.05, 0.1, 100, 2.0), (0.05, 0.0, 0.2, 300, 1.0)]:      # RECTANGULAR, THIN
        coil, oc = _both(cc, dims, z=0.2, alpha=0.3, beta=1.2, x=0.05)
        L = coil.compute_and_set_self_inductance(C.PrecisionFactor(6.0))
        centred = O.make_coil(*dims, 50.0)
        a = O.pair_precision_args(centred, centred, 6.0, False, True, False)
        assert _close(L, O.self_inductance(centred, a), O.self_inductance(centred, a, True))
        assert coil.get_self_inductance() == L
        P = coil.get_position_vector()
        assert (P.x, P.y, P.z) == (0.05, 0.0, 0.2) and tuple(coil.get_rotation_angles()) == (0.3, 1.2)
    # FLAT: the slow coaxial mutual inductance of the coil with itself
    dims = (0.1, 0.05, 0.0, 30, 1.0)
    coil, _ = _both(cc, dims)
    L = coil.compute_and_set_self_inductance(C.PrecisionFactor(4.0))
    oc = O.make_coil(*dims, 50.0)
    a = O.pair_precision_args(oc, oc, 4.0, False, True, False)
    fa = np.array([1, 1, 1, a[3], a[4], a[5]], dtype=np.int32)
    ref = O.mutual_inductance_zaxis_slow(oc, oc, a, fa, extended=True)
    assert abs(L - ref) <= 1e-10 * abs(ref), (L, ref)
    # a filament loop has no self-inductance
    loop, _ = _both(cc, (0.1, 0.0, 0.0, 1, 1.0))
    with pytest.raises(Exception, match="divergent"):
        loop.compute_and_set_self_inductance(C.PrecisionFactor(4.0))


def test_dropin_group_mixed_members(cc):
    """CoilGroup sums per member with each member's own rule: coaxial members the z-axis rules, the others the
    general rule (CoilGroup ComputeMInductance.cxx / ComputeForce.cxx call Coil's pair methods per member)."""
    C, V = cc.coil, cc.tensor.Vector3
    members = [((0.1, 0.05, 0.1, 100, 2.0), (0, 0, 0.0), 0.0), ((0.06, 0.02, 0.05, 80, 1.0), (0, 0, -0.3), 0.0),
               ((0.05, 0.02, 0.04, 60, 1.5), (0.4, 0.1, 0.2), 0.5)]
    sdims = (0.08, 0.04, 0.06, 50, 1.5)
    group = cc.coil_group.CoilGroup()
    for dims, pos, alpha in members:
        R, a, b, N, I = dims
        group.add_coil(R, a, b, N, I, C.PrecisionFactor(5.0), 8, V(*pos), alpha, 0.0)
    sec = C.Coil(*sdims, 50.0, C.PrecisionFactor(5.0), 8, V(0.0, 0.0, 0.35), 0.0, 0.0)
    os_ = O.make_coil(*sdims, 50.0, pos=(0, 0, 0.35))
    M = group.compute_mutual_inductance(sec, C.PrecisionFactor(5.0), C.GPU)
    F, T = group.compute_force_torque(sec, C.PrecisionFactor(5.0), C.GPU)
    m_ref, f_ref, t_ref = 0.0, np.zeros(3), np.zeros(3)
    for dims, pos, alpha in members:
        op = O.make_coil(*dims, 50.0, pos=pos, alpha=alpha)
        if pos[0] == 0 and pos[1] == 0 and alpha == 0:
            a = O.pair_precision_args(op, os_, 5.0, True, True, False)
            m_ref += O.mutual_inductance_zaxis_fast(op, os_, a, True)
            f_ref[2] += O.force_zaxis_fast(op, os_, a, True)
        else:
            a = O.pair_precision_args(op, os_, 5.0, True, False, False)
            m_ref += O.mutual_inductance_general(op, os_, a)
            f, t = O.force_torque_general(op, os_, a)
            f_ref += f
            t_ref += t
    assert abs(M - m_ref) <= 1e-9 * abs(m_ref), (M, m_ref)
    assert np.linalg.norm(np.array([F.x, F.y, F.z]) - f_ref) <= 1e-9 * np.linalg.norm(f_ref)
    assert np.linalg.norm(np.array([T.x, T.y, T.z]) - t_ref) <= 1e-9 * np.linalg.norm(t_ref)


def test_dropin_reference_zaxis_comparison_workload(cc):
    """The reference's own z-axis comparison (Compare::mutualInductanceZAxis, CompMInductanceZAxis.cxx:16-53): the
    19 coaxial geometries of data/values_MInductance_zAxis.txt at PrecisionFactor 1..8 with the CPU_ST arguments the
    reference plans, here computed on the GPU, against the oracle's restatement of the same rule."""
    import os
    C, V = cc.coil, cc.tensor.Vector3
    rows = np.loadtxt(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                                   "values_MInductance_zAxis.txt"))
    for R1, a1, b1, N1, R2, a2, b2, N2, d in rows:
        prim = C.Coil(R1, a1, b1, int(N1))
        sec = C.Coil(R2, a2, b2, int(N2))
        sec.set_position_and_orientation(V(0.0, 0.0, d), 0.0, 0.0)
        op, os_ = O.make_coil(R1, a1, b1, int(N1)), O.make_coil(R2, a2, b2, int(N2), pos=(0.0, 0.0, d))
        fast = op.use_fast and os_.use_fast
        for pf in range(1, 9):
            args = C.CoilPairArguments.get_appropriate_coil_pair_arguments(prim, sec, C.PrecisionFactor(pf), C.CPU_ST,
                                                                          True, False)
            M = C.Coil.compute_mutual_inductance(prim, sec, args, C.GPU)
            a = O.pair_precision_args(op, os_, float(pf), False, True, False)
            if fast:
                assert _close(M, O.mutual_inductance_zaxis_fast(op, os_, a),
                              O.mutual_inductance_zaxis_fast(op, os_, a, True)), (R1, R2, d, pf)
            else:
                fa = np.array([1, 1, 1, a[3], a[4], a[5]], dtype=np.int32)
                m = O.mutual_inductance_zaxis_slow(op, os_, a, fa, extended=True)
                assert abs(M - m) <= 1e-11 * abs(m), (R1, R2, d, pf)


def test_dropin_reference_self_inductance_workload(cc):
    """The reference's self-inductance comparison (Compare::selfInductance, CompMInductanceZAxis.cxx:56-140): a
    RECTANGULAR, a THIN and a FLAT coil and Grover's (alpha, beta) thick coils at several precision factors."""
    C = cc.coil
    r1 = 0.1
    dims = [(0.03, 0.03, 0.12, 3600), (0.03, 0.0, 0.12, 120), (0.03, 0.03, 0.0, 30)]
    dims += [(r1, r1 * (al - 1), 2 * r1 * be, 1) for al, be in ((1.5, 0.25), (3.0, 1.0), (4.0, 3.0), (7.0, 6.0), (9.0, 4.0))]
    for R, a, b, N in dims:
        coil = C.Coil(R, a, b, N)
        oc = O.make_coil(R, a, b, N)
        for pf in (2.0, 5.0, 8.0):
            L = coil.compute_and_set_self_inductance(C.PrecisionFactor(pf))
            args = O.pair_precision_args(oc, oc, pf, False, True, False)
            if oc.type == O.FLAT:
                fa = np.array([1, 1, 1, args[3], args[4], args[5]], dtype=np.int32)
                ref = O.mutual_inductance_zaxis_slow(oc, oc, args, fa, extended=True)
                assert abs(L - ref) <= 1e-10 * abs(ref), (R, a, b, pf)
            else:
                assert _close(L, O.self_inductance(oc, args), O.self_inductance(oc, args, True)), (R, a, b, pf)


def test_dropin_reference_zaxis_force_workloads(cc):
    """The reference's coaxial force comparisons (CompForce.cxx: forceTorqueFilamentsZAxis :8-20,
    forceTorqueThinCoilsZAxis :102-150, forceTorqueZAxis :325-337 on 25 of its 500 positions) with the arguments the
    reference plans for CPU_ST / CPU_MT, computed on the GPU, against the oracle."""
    C, V = cc.coil, cc.tensor.Vector3
    cases = [((0.5, 0.0, 0.0, 1, 100.0), (0.3, 0.0, 0.0, 1, 200.0), [0.4], range(1, 9)),
             ((0.5, 0.0, 0.3, 200, 1.0), (0.5, 0.0, 0.2, 100, 1.0), [0.4], range(1, 11)),
             ((0.1, 0.0, 0.15, 225, 1.0), (0.09, 0.0, 0.15, 300, 1.0), [0.05], range(1, 13)),
             ((0.25, 0.0, 0.06, 40, 1.0), (0.2, 0.0, 0.04, 120, 1.0), [0.3], range(1, 11)),
             ((0.0762, 0.0832, 0.0, 516, 1.42), (0.0762, 0.0832, 0.0, 516, 1.42), [0.0468], range(1, 9)),
             ((0.03, 0.03, 0.12, 3600, 1.0), (0.02, 0.025, 0.04, 1000, 1.0),
              [0.08 + i * 0.001 for i in range(0, 500, 20)], [8])]
    for pdims, sdims, zs, pfs in cases:
        prim = C.Coil(*pdims)
        op = O.make_coil(*pdims)
        for z in zs:
            sec = C.Coil(*sdims)
            sec.set_position_and_orientation(V(0.0, 0.0, z), 0.0, 0.0)
            os_ = O.make_coil(*sdims, pos=(0.0, 0.0, z))
            for pf in pfs:
                args = C.CoilPairArguments.get_appropriate_coil_pair_arguments(prim, sec, C.PrecisionFactor(pf),
                                                                              C.CPU_ST, True, False)
                F, T = C.Coil.compute_ampere_force(prim, sec, args, C.GPU)
                assert F.x == 0.0 and F.y == 0.0 and (T.x, T.y, T.z) == (0.0, 0.0, 0.0)
                a = O.pair_precision_args(op, os_, float(pf), False, True, False)
                if op.use_fast and os_.use_fast:
                    assert _close(F.z, O.force_zaxis_fast(op, os_, a), O.force_zaxis_fast(op, os_, a, True)), \
                        (pdims, z, pf)
                else:
                    fa = np.array([1, 1, 1, a[3], a[4], a[5]], dtype=np.int32)
                    f = O.force_zaxis_slow(op, os_, a, fa, extended=True)
                    assert abs(F.z - f) <= 1e-11 * abs(f), (pdims, z, pf, F.z, f)
