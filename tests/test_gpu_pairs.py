"""GPU parity of the pair path (ccoil_pairs through the C ABI, and the c_coil drop-in API) against the FP64
oracle. Tolerances: CCOIL_FP64 and CCOIL_FP64_IEEE relative error <= 1e-10; CCOIL_FP32C <= 3e-5 for M and |F|, |T|
(secondary nodes in the primary's near-winding shell are evaluated with the FP64 arithmetic; far from the primary the
FP32 node error is amplified by the angular sums, as in the field path)."""
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


@pytest.mark.parametrize("prec,tol", [(abi.FP64, 1e-10), (abi.FP64_IEEE, 1e-10), (abi.FP32C, 3e-5)])
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


def _shifted(rows, k, axis, h):
    out = [dict(r) for r in rows]
    pos = list(out[k]["pos"])
    pos[axis] += h
    out[k]["pos"] = tuple(pos)
    return out


@pytest.mark.parametrize("prec", [abi.FP64, abi.FP32C])
def test_mutual_inductance_gradient(ctx, prec):
    """CCOIL_PAIR_DM is the exact derivative of the discrete M with respect to the secondary's position: it
    matches central differences of M itself (FP64 to 1e-7), and F / (I1 I2) up to the quadrature error (1e-5)."""
    rng = np.random.default_rng(31)
    prim = W.random_coil_params(rng, 6)
    sec = W.random_coil_params(rng, 6)
    for s in sec:
        s["pos"] = s["pos"] * 2.0                      # keep the pairs apart
    incs_p = [(60, 12 if p["type"] in ("FLAT", "RECTANGULAR") else 1, 1) for p in prim]
    incs_s = [(48, 8 if s["type"] in ("FLAT", "RECTANGULAR") else 1, 8 if s["type"] in ("THIN", "RECTANGULAR") else 1)
              for s in sec]
    prow, srow = _rows(prim, incs_p), _rows(sec, incs_s)
    P = abi.CoilSet.from_rows(prow)
    out, _ = ctx.pairs(P, abi.CoilSet.from_rows(srow), 6, abi.PAIR_M | abi.PAIR_FT | abi.PAIR_DM, prec)
    ref, _ = ctx.pairs(P, abi.CoilSet.from_rows(srow), 6, abi.PAIR_M | abi.PAIR_FT | abi.PAIR_DM, abi.FP64)
    for k in range(6):
        dm = out["dM"][k]
        I1, I2 = prim[k]["I"], sec[k]["I"]
        force = out["F"][k] / (I1 * I2)
        print(f"pair {k}: dM {dm}, F/(I1 I2) {force}")
        assert np.linalg.norm(dm - force) <= 1e-5 * np.linalg.norm(force)
        if prec == abi.FP64:
            h = 1e-6
            fd = np.zeros(3)
            for ax in range(3):
                mp, _ = ctx.pairs(P, abi.CoilSet.from_rows(_shifted(srow, k, ax, h)), 6, abi.PAIR_M, abi.FP64)
                mm, _ = ctx.pairs(P, abi.CoilSet.from_rows(_shifted(srow, k, ax, -h)), 6, abi.PAIR_M, abi.FP64)
                fd[ax] = (mp["M"][k] - mm["M"][k]) / (2 * h)
            assert np.linalg.norm(dm - fd) <= 1e-7 * np.linalg.norm(fd), (k, dm, fd)
        else:
            assert np.linalg.norm(dm - ref["dM"][k]) <= 3e-5 * np.linalg.norm(ref["dM"][k])


def test_mutual_inductance_matrix_and_gradient(ctx):
    """The off-diagonal M matrix of a 48-coil group through index arrays, with its position gradient: dM_ij
    against F_ij / (I_i I_j), reciprocity M_ij = M_ji, and translation invariance of the gradient."""
    params, _ = W.group_workload(48, 8, seed=9)
    for p in params:
        p["pos"] = p["pos"] * 4.0                      # spread the coils so that no two overlap
    prim, sec = abi.CoilSet.from_rows(W.group_rows(params)), abi.CoilSet.from_rows(W.secondary_rows(params))
    pi, si = W.matrix_pairs(len(params))
    out, rep = ctx.pairs(prim, sec, len(pi), abi.PAIR_M | abi.PAIR_FT | abi.PAIR_DM, abi.FP64,
                         prim_index=pi, sec_index=si)
    n = len(params)
    M = np.zeros((n, n))
    M[pi, si] = out["M"]
    I = np.array([p["I"] for p in params])
    force = out["F"] / (I[pi] * I[si])[:, None]
    e = np.linalg.norm(out["dM"] - force, axis=1) / np.linalg.norm(force, axis=1)
    print(f"\nM matrix {n}x{n}: dM vs F/(I I) max {e.max():.3g} median {np.median(e):.3g}; "
          f"reciprocity max {np.max(np.abs(M - M.T)) / np.max(np.abs(M)):.3g}")
    assert np.median(e) < 1e-6 and np.percentile(e, 99) < 1e-4
    assert np.max(np.abs(M - M.T)) <= 1e-3 * np.max(np.abs(M))


def test_row_split_matches_oracle_and_large_batch(ctx):
    """A small batch splits large primaries' thickness rows over several work items (report.split > 1); a large
    batch does not. Both match the oracle's GPU pair rule, and each other to rounding."""
    rng = np.random.default_rng(41)
    prim = W.random_coil_params(rng, 3)
    sec = W.random_coil_params(rng, 3)
    for p in prim:
        p.update(type="RECTANGULAR", a=max(p["R"] * 0.3, 0.01), b=max(p["R"] * 0.4, 0.01))
    for s in sec:
        s["pos"] = s["pos"] * 2.0
    pi = [(40, 20, 1)] * 3
    si = [(24, 4 if s["type"] in ("FLAT", "RECTANGULAR") else 1, 4 if s["type"] in ("THIN", "RECTANGULAR") else 1)
          for s in sec]
    prow, srow = _rows(prim, pi), _rows(sec, si)
    small, rep = ctx.pairs(abi.CoilSet.from_rows(prow), abi.CoilSet.from_rows(srow), 3, abi.PAIR_M | abi.PAIR_FT,
                           abi.FP64)
    assert rep.split > 1
    for k in range(3):
        op, os_ = H.oracle_coil(prim[k]), H.oracle_coil(sec[k])
        m = O.mutual_inductance_gpu_rule(op, os_, pi[k], si[k])
        F, T = O.force_torque_gpu_rule(op, os_, pi[k], si[k], 2 * np.pi * sec[k]["N"] * sec[k]["I"])
        assert abs(small["M"][k] - m) <= 1e-10 * abs(m), (k, small["M"][k], m)
        assert np.linalg.norm(small["F"][k] - F) <= 1e-10 * np.linalg.norm(F), k
        assert np.linalg.norm(small["T"][k] - T) <= 1e-10 * max(np.linalg.norm(T), 1e-300), k
    bp, bs = W.pair_workload(8192, seed=42)
    big, rep_big = ctx.pairs(abi.CoilSet.from_rows(bp + prow), abi.CoilSet.from_rows(bs + srow), 8195,
                             abi.PAIR_M | abi.PAIR_FT, abi.FP64)
    assert rep_big.split == 1
    assert np.allclose(big["M"][-3:], small["M"], rtol=1e-12, atol=0)
    assert np.allclose(big["F"][-3:], small["F"], rtol=1e-12, atol=1e-12 * np.abs(small["F"]).max())
    assert np.allclose(big["T"][-3:], small["T"], rtol=1e-12, atol=1e-12 * np.abs(small["T"]).max())


def test_fp32c_near_nodes_run_in_fp64(ctx):
    """FP32C evaluates secondary nodes inside the primary's near-winding shell in FP64 (queued and run in full
    waves): a secondary wound inside the primary's winding has only such nodes, so FP32C equals FP64 up to the
    order of the sums."""
    prim = [dict(type="RECTANGULAR", R=0.1, a=0.02, b=0.1, N=200, I=2.0, pos=np.zeros(3), alpha=0.0, beta=0.0)]
    sec = [dict(type="THIN", R=0.105, a=0.0, b=0.05, N=50, I=1.5, pos=np.array([0.001, -0.002, 0.003]), alpha=0.02,
                beta=0.4)]
    prow, srow = _rows(prim, [(48, 16, 1)]), _rows(sec, [(64, 1, 24)])
    P, S = abi.CoilSet.from_rows(prow), abi.CoilSet.from_rows(srow)
    o64, _ = ctx.pairs(P, S, 1, abi.PAIR_M | abi.PAIR_FT, abi.FP64)
    o32, _ = ctx.pairs(P, S, 1, abi.PAIR_M | abi.PAIR_FT, abi.FP32C)
    assert abs(o32["M"][0] - o64["M"][0]) <= 1e-12 * abs(o64["M"][0])
    assert np.linalg.norm(o32["F"][0] - o64["F"][0]) <= 1e-12 * np.linalg.norm(o64["F"][0])
    # the torque of a nearly coaxial pair is small: its error is measured against the force times the radius too
    scale_t = np.linalg.norm(o64["T"][0]) + 0.1 * np.linalg.norm(o64["F"][0])
    assert np.linalg.norm(o32["T"][0] - o64["T"][0]) <= 1e-12 * scale_t
