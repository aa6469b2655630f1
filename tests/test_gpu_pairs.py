"""GPU parity of the pair path (ccoil_pairs through the C ABI, and the c_coil drop-in API) against the FP64
oracle. Tolerances: FP64 relative error <= 1e-10; FP32C <= 1e-6 for M and |F|, |T| (pairs whose secondary
nodes fall in the primary's near-winding shell are evaluated in FP64 node by node)."""
import numpy as np
import pytest

import ccoil_abi as abi
import helpers as H
import oracle as O
import workloads as W

pytestmark = pytest.mark.gpu


def _rows(params, incs):
    rows = []
    for p, inc in zip(params, incs):
        r = W.coil_row(W.make_coil(p))
        r.update(n_ang=inc[0], n_thk=inc[1], n_len=inc[2], blocks=(1, 1, 1))
        rows.append(r)
    return rows


def _small_incs(rng, params, sec):
    out = []
    for p in params:
        ang = int(rng.integers(6, 17))
        thk = int(rng.integers(3, 7)) if p["type"] in ("FLAT", "RECTANGULAR") else 1
        ln = (int(rng.integers(3, 7)) if p["type"] in ("THIN", "RECTANGULAR") else 1) if sec else 1
        out.append((ang, thk, ln))
    return out


@pytest.mark.parametrize("prec,tol", [(abi.FP64, 1e-10), (abi.FP64_IEEE, 1e-10), (abi.FP32C, 1e-6)])
def test_gpu_rule_pairs_match_oracle(ctx, prec, tol):
    rng = np.random.default_rng(7)
    n = 24
    prim = W.random_coil_params(rng, n)
    sec = W.random_coil_params(rng, n)
    pi, si = _small_incs(rng, prim, False), _small_incs(rng, sec, True)
    P, S = abi.CoilSet.from_rows(_rows(prim, pi)), abi.CoilSet.from_rows(_rows(sec, si))
    out, rep = ctx.pairs(P, S, n, abi.PAIR_M | abi.PAIR_FT, prec)
    assert rep.node_evaluations == sum(a[0] * a[1] * b[0] * b[1] * b[2] for a, b in zip(pi, si))
    for k in range(n):
        op, os_ = H.oracle_coil(prim[k]), H.oracle_coil(sec[k])
        m = O.mutual_inductance_gpu_rule(op, os_, pi[k], si[k])
        F, T = O.force_torque_gpu_rule(op, os_, pi[k], si[k], 2 * np.pi * sec[k]["N"] * sec[k]["I"])
        assert abs(out["M"][k] - m) <= tol * abs(m), (k, out["M"][k], m)
        assert np.linalg.norm(out["F"][k] - F) <= tol * np.linalg.norm(F), k
        assert np.linalg.norm(out["T"][k] - T) <= tol * max(np.linalg.norm(T), 1e-300), k


def test_force_reference_scale_flag(ctx):
    rng = np.random.default_rng(8)
    prim = W.random_coil_params(rng, 3)
    sec = W.random_coil_params(rng, 3)
    pi, si = _small_incs(rng, prim, False), _small_incs(rng, sec, True)
    P, S = abi.CoilSet.from_rows(_rows(prim, pi)), abi.CoilSet.from_rows(_rows(sec, si))
    phys, _ = ctx.pairs(P, S, 3, abi.PAIR_FT, abi.FP64)
    refs, _ = ctx.pairs(P, S, 3, abi.PAIR_FT, abi.FP64, flags=abi.FORCE_REFERENCE_SCALE)
    for k in range(3):
        ratio = (1.0 / prim[k]["I"]) / sec[k]["I"]   # (2 pi N2 / I1) / (2 pi N2 I2)
        assert np.allclose(refs["F"][k], phys["F"][k] * ratio, rtol=1e-12, atol=0)


def test_drop_in_general_path_matches_oracle():
    import c_coil
    C = c_coil.coil
    rng = np.random.default_rng(9)
    C.Coil.set_gpu_precision(C.GPU_FP64)
    try:
        for _ in range(6):
            p, s = W.random_coil_params(rng, 2)
            cp, cs = W.make_coil(p), W.make_coil(s)
            args = C.CoilPairArguments.get_appropriate_coil_pair_arguments(cp, cs, C.PrecisionFactor(2.0), C.GPU,
                                                                          False, False)
            a12 = [args.primary_precision.angular_block_count, args.primary_precision.thickness_block_count,
                   args.primary_precision.length_block_count, args.primary_precision.angular_increment_count,
                   args.primary_precision.thickness_increment_count, args.primary_precision.length_increment_count,
                   args.secondary_precision.angular_block_count, args.secondary_precision.thickness_block_count,
                   args.secondary_precision.length_block_count, args.secondary_precision.angular_increment_count,
                   args.secondary_precision.thickness_increment_count,
                   args.secondary_precision.length_increment_count]
            op, os_ = H.oracle_coil(p), H.oracle_coil(s)
            m_ref = O.mutual_inductance_general(op, os_, a12)
            m = C.Coil.compute_mutual_inductance(cp, cs, args, C.GPU)
            assert abs(m - m_ref) <= 1e-10 * abs(m_ref)
            F_ref, T_ref = O.force_torque_general(op, os_, a12)
            F, T = C.Coil.compute_ampere_force(cp, cs, args, C.GPU)
            assert np.linalg.norm(np.array([F.x, F.y, F.z]) - F_ref) <= 1e-10 * np.linalg.norm(F_ref)
            assert np.linalg.norm(np.array([T.x, T.y, T.z]) - T_ref) <= 1e-10 * np.linalg.norm(T_ref)
    finally:
        C.Coil.set_gpu_precision(C.GPU_FP32C)
