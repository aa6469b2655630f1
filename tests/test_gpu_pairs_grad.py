"""ccoil_pairs_grad (the adjoint of the mutual inductance M with respect to both coils' parameters) against the oracle.

  - Fourth-order central finite differences (five points) of the oracle's GPU pair rule
    (orc_mutual_inductance_gpu_rule, the same quadrature with the same orders) in every free parameter of both coils:
    position, alpha, beta, R, and the thickness a (FLAT, RECTANGULAR) and length b (THIN, RECTANGULAR) where the coil
    type has them (a thin dimension is clamped to 1e-25 R by the Coil constructor and is not a parameter). Step
    1e-3 of the coil size (1e-3 rad for angles): M's own rounding (about 5e-14 relative for a sum of 10^6 node
    terms) makes steps of 1e-6 noisy at 1e-6, while the stencil's truncation at 1e-3 stays near 1e-12; relative
    error <= 1e-6 of max(|FD|, 1e-3 of the pair's largest derivative).
  - The reference's force and torque (ccoil_pairs, the B rule of Force.cu / CalculateForceGeneral.cxx:114-128) are
    I1 I2 dM/dp and I1 I2 dM/dtheta: F = I1 I2 dM/d(secondary position), T . v = I1 I2 dM/dalpha (alpha turns the
    secondary about v = (-sin beta, cos beta, 0)) and T_z = I1 I2 dM/dbeta (beta turns it about z), on separated
    pairs at high orders, where the two quadratures agree: <= 1e-7.
  - CCOIL_FP32C against CCOIL_FP64: <= 1e-4 of the pair's largest derivative.
"""
import numpy as np
import pytest

import ccoil_abi as abi
import oracle as O
import workloads as W

pytestmark = pytest.mark.gpu

TERMS = abi.PAIR_GRAD_TERMS


def _pairs(n, seed, incs=None, sep=2.5, max_inc=80):
    """n pairs of the mixed-coil distributions, the secondary at sep times the radii from the primary; increments
    incs, or drawn from {8, 12, ..., max_inc} per active dimension."""
    rng = np.random.default_rng(seed)
    prim = W.random_coil_params(rng, n)
    sec = W.random_coil_params(rng, n)
    prow, srow = [], []
    for p, s in zip(prim, sec):
        d = rng.standard_normal(3)
        s["pos"] = np.asarray(p["pos"]) + sep * (p["R"] + p["a"] + s["R"] + s["a"] + 0.5 * (p["b"] + s["b"])) * d / np.linalg.norm(d)
        cp, cs = W.make_coil(p), W.make_coil(s)
        pick = (lambda: int(incs)) if incs else (lambda: int(4 * rng.integers(2, max_inc // 4 + 1)))
        pa = (pick(), pick() if p["type"] in (W.FLAT, W.RECTANGULAR) else 1, 1)
        sa = (pick(), pick() if s["type"] in (W.FLAT, W.RECTANGULAR) else 1,
              pick() if s["type"] in (W.THIN, W.RECTANGULAR) else 1)
        rp = W.coil_row(cp)
        rp.update(n_ang=pa[0], n_thk=pa[1], n_len=pa[2], blocks=(1, 1, 1))
        rs = W.coil_row(cs)
        rs.update(n_ang=sa[0], n_thk=sa[1], n_len=sa[2], blocks=(1, 1, 1))
        prow.append(rp)
        srow.append(rs)
    return prim, sec, prow, srow


def _oracle_coil(p, **change):
    q = dict(p)
    pos = np.array(q["pos"], dtype=float)
    for k, v in change.items():
        if k in ("x", "y", "z"):
            pos["xyz".index(k)] += v
        else:
            q[k] = q[k] + v
    return O.make_coil(q["R"], q["a"], q["b"], int(q["N"]), q["I"], 50.0, tuple(pos), q["alpha"], q["beta"])


def _free(p, name):
    if name == "a":
        return p["type"] in (W.FLAT, W.RECTANGULAR)
    if name == "b":
        return p["type"] in (W.THIN, W.RECTANGULAR)
    return True


def test_gradient_matches_oracle_finite_differences(ctx):
    # increments up to 16: the oracle's rule costs (primary nodes) x (secondary nodes) per M, 64 M per pair here
    prim, sec, prow, srow = _pairs(8, seed=61, max_inc=16)
    P, S = abi.CoilSet.from_rows(prow), abi.CoilSet.from_rows(srow)
    M, grad, _ = ctx.pairs_grad(P, S, len(prow), abi.FP64)
    names = ("x", "y", "z", "alpha", "beta", "R", "a", "b")
    worst = 0.0
    for k, (p, s, rp, rs) in enumerate(zip(prim, sec, prow, srow)):
        pinc, sinc = (rp["n_ang"], rp["n_thk"], rp["n_len"]), (rs["n_ang"], rs["n_thk"], rs["n_len"])
        m0 = O.mutual_inductance_gpu_rule(_oracle_coil(p), _oracle_coil(s), pinc, sinc)
        assert abs(M[k] - m0) <= 1e-10 * abs(m0), k
        scale = np.max(np.abs(grad[k]))
        size = p["R"] + s["R"]
        for side, coil in (("sec", s), ("prim", p)):
            for name in names:
                if not _free(coil, name):
                    continue
                h = 1e-3 if name in ("alpha", "beta") else 1e-3 * (size if name in "xyz" else coil[name])
                vals = []
                for step in (2.0, 1.0, -1.0, -2.0):
                    ch = {name: step * h}
                    pc = _oracle_coil(p, **ch) if side == "prim" else _oracle_coil(p)
                    sc = _oracle_coil(s, **ch) if side == "sec" else _oracle_coil(s)
                    vals.append(O.mutual_inductance_gpu_rule(pc, sc, pinc, sinc))
                fd = (-vals[0] + 8.0 * vals[1] - 8.0 * vals[2] + vals[3]) / (12.0 * h)
                got = grad[k, TERMS.index(f"{side}_{name}")]
                err = abs(got - fd) / max(abs(fd), 1e-3 * scale)
                worst = max(worst, err)
                assert err <= 1e-6, (k, side, name, got, fd, p["type"], s["type"])
    print(f"\nworst relative difference from the oracle's finite differences: {worst:.3g}")


def test_force_and_torque_are_current_times_gradient(ctx):
    prim, sec, prow, srow = _pairs(12, seed=62, incs=60, sep=3.0)
    P, S = abi.CoilSet.from_rows(prow), abi.CoilSet.from_rows(srow)
    n = len(prow)
    _, grad, _ = ctx.pairs_grad(P, S, n, abi.FP64)
    ft, _ = ctx.pairs(P, S, n, abi.PAIR_FT, abi.FP64)
    for k, (p, s) in enumerate(zip(prim, sec)):
        II = p["I"] * s["I"]
        F = II * grad[k, 0:3]
        assert np.linalg.norm(ft["F"][k] - F) <= 1e-7 * np.linalg.norm(F), k
        v = np.array([-np.sin(s["beta"]), np.cos(s["beta"]), 0.0])
        tq = ft["T"][k]
        tscale = np.linalg.norm(tq)
        assert abs(tq @ v - II * grad[k, TERMS.index("sec_alpha")]) <= 1e-7 * tscale, k
        assert abs(tq[2] - II * grad[k, TERMS.index("sec_beta")]) <= 1e-7 * tscale, k


def test_fp32c_gradient_against_fp64(ctx):
    _, _, prow, srow = _pairs(64, seed=63)
    P, S = abi.CoilSet.from_rows(prow), abi.CoilSet.from_rows(srow)
    _, g64, _ = ctx.pairs_grad(P, S, len(prow), abi.FP64)
    _, g32, _ = ctx.pairs_grad(P, S, len(prow), abi.FP32C)
    scale = np.max(np.abs(g64), axis=1, keepdims=True)
    err = np.max(np.abs(g32 - g64) / scale)
    print(f"\nFP32C gradient against FP64: max {err:.3g} of the pair's largest derivative")
    assert err <= 1e-4
    with pytest.raises(abi.CCoilError):
        ctx.pairs_grad(P, S, len(prow), abi.FP32)
