"""GPU checks at the benchmark's full sizes, through properties that hold at any size (the oracle covers the small
sizes in the other test files):

  - the 128-coil mixed group at 2^20 points (B and the gradient):
      * CCOIL_FP64 (hardware reciprocals refined by one Newton step) against CCOIL_FP64_IEEE (IEEE division and
        square root, the same integrands): two arithmetic paths of the same rule;
      * superposition: the group equals the sum of its two halves, to rounding of the sum of the coils' magnitudes;
      * CCOIL_FP32C against CCOIL_FP64 relative to the sum of the coils' magnitudes at each point (the
        superposition's own conditioning), B and the gradient within 3e-6 at every point, and bitwise reproducibility of FP32C (its near-winding pass has no
        floating-point atomics);
  - 65,536 random pairs of mixed type with increments ~ U{4, ..., 80} (M, F, T):
      * CCOIL_FP64 against CCOIL_FP64_IEEE;
      * permutation invariance: a pair's result does not depend on its position in the batch or on the other
        pairs (the launch orders work items longest first), bitwise.
"""
import numpy as np
import pytest

import ccoil_abi as abi
import workloads as W

pytestmark = pytest.mark.gpu


def _norm(x):
    return np.linalg.norm(np.asarray(x).reshape(len(x), -1), axis=1)


def test_group_full_size(ctx):
    import torch
    dev = torch.device("cuda:0")
    params, pts = W.group_workload(128, 1 << 20, seed=0)
    rows = W.group_rows(params)
    n = len(pts)
    q = abi.B | abi.G
    d_pts = torch.from_numpy(np.ascontiguousarray(pts)).to(dev)

    def run(rs, prec):
        out = {"B": torch.empty((n, 3), dtype=torch.float64, device=dev),
               "G": torch.empty((n, 9), dtype=torch.float64, device=dev)}
        ctx.fields(abi.CoilSet.from_rows(rs), d_pts, q, prec, abi.DEVICE_POINTERS, out=out)
        return out

    full = run(rows, abi.FP64)
    # sum over coils of |B_k| and |G_k| at every point
    mb = torch.zeros(n, dtype=torch.float64, device=dev)
    mg = torch.zeros(n, dtype=torch.float64, device=dev)
    for r in rows:
        one = run([r], abi.FP64)
        mb += torch.linalg.vector_norm(one["B"], dim=1)
        mg += torch.linalg.vector_norm(one["G"], dim=1)
    b64, g64 = full["B"].cpu().numpy(), full["G"].cpu().numpy()
    mb, mg = mb.cpu().numpy(), mg.cpu().numpy()
    assert np.all(mb > 0) and np.all(mg > 0)

    ieee = run(rows, abi.FP64_IEEE)
    eb = _norm(ieee["B"].cpu().numpy() - b64) / _norm(b64)
    eg = _norm(ieee["G"].cpu().numpy() - g64) / _norm(g64)
    print(f"\nFP64 vs FP64_IEEE: B max {eb.max():.3g}, G max {eg.max():.3g}")
    assert eb.max() < 1e-12 and eg.max() < 1e-11

    h1, h2 = run(rows[:64], abi.FP64), run(rows[64:], abi.FP64)
    sb = _norm(h1["B"].cpu().numpy() + h2["B"].cpu().numpy() - b64) / mb
    sg = _norm(h1["G"].cpu().numpy() + h2["G"].cpu().numpy() - g64) / mg
    print(f"superposition of two halves: B max {sb.max():.3g}, G max {sg.max():.3g} (relative to sum_k |.|)")
    assert sb.max() < 1e-14 and sg.max() < 1e-14

    f1 = run(rows, abi.FP32C)
    f2 = run(rows, abi.FP32C)
    assert torch.equal(f1["B"], f2["B"]) and torch.equal(f1["G"], f2["G"])
    cb = _norm(f1["B"].cpu().numpy() - b64) / mb
    cg = _norm(f1["G"].cpu().numpy() - g64) / mg
    print(f"FP32C vs FP64 relative to sum_k |.|: B max {cb.max():.3g} p99 {np.percentile(cb, 99):.3g}; "
          f"G max {cg.max():.3g} p99 {np.percentile(cg, 99):.3g} p99.9 {np.percentile(cg, 99.9):.3g}")
    assert cb.max() < 2e-6
    # the hoop term B_rho/r of pairs near a coil's axis goes to the FP64 pass, so the gradient has no 1/r blow-up
    assert cg.max() < 3e-6 and np.percentile(cg, 99) < 1e-6 and np.percentile(cg, 99.9) < 2e-6


def test_pairs_full_size(ctx):
    prow, srow = W.pair_workload(65536, seed=0)
    P, S = abi.CoilSet.from_rows(prow), abi.CoilSet.from_rows(srow)
    n = len(prow)
    what = abi.PAIR_M | abi.PAIR_FT
    o64, _ = ctx.pairs(P, S, n, what, abi.FP64)
    oie, _ = ctx.pairs(P, S, n, what, abi.FP64_IEEE)
    em = np.abs(oie["M"] - o64["M"]) / np.abs(o64["M"])
    ef = _norm(oie["F"] - o64["F"]) / _norm(o64["F"])
    et = _norm(oie["T"] - o64["T"]) / _norm(o64["T"])
    print(f"\npairs FP64 vs FP64_IEEE: M max {em.max():.3g} p99 {np.percentile(em, 99):.3g}; "
          f"F max {ef.max():.3g} p99 {np.percentile(ef, 99):.3g}; T max {et.max():.3g} p99 {np.percentile(et, 99):.3g}")
    assert np.percentile(em, 99) < 1e-12 and np.percentile(ef, 99) < 1e-12 and np.percentile(et, 99) < 1e-12
    assert em.max() < 1e-8 and ef.max() < 1e-8 and et.max() < 1e-8

    perm = np.random.default_rng(5).permutation(n)
    op, _ = ctx.pairs(P, S, n, what, abi.FP64, prim_index=perm, sec_index=perm)
    assert np.array_equal(op["M"], o64["M"][perm])
    assert np.array_equal(op["F"], o64["F"][perm]) and np.array_equal(op["T"], o64["T"][perm])
