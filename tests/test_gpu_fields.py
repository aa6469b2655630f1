"""GPU parity of the field path (ccoil_fields through the C ABI) against the oracle.

The oracle integrates the same quadrature rule (the coil's GPU PrecisionArguments, one block per layer).
Accuracy is measured against the oracle's long double evaluation of that rule (the truth), because the
reference's own FP64 gradient integrands lose up to ~3e-5 to cancellation near the winding
(test_oracle_fields.py::test_reference_fp64_gradient_conditioning pins that). Metric: per-point vector-norm relative error
|x - ref| / |ref|. Tolerances:
  CCOIL_FP64           A, B <= 1e-11 at every point; gradient <= 1e-9 away from the near-winding shell, <= 1e-6
                       within it (the node sum's own conditioning there; see test_fp64_matches_truth)
  CCOIL_FP64_REFERENCE A, B <= 1e-10 against the oracle's double restatement of the same integrands
  CCOIL_FP32C          single coil: B and gradient <= 1e-6 away from the near-winding shell (all of them are
                       re-evaluated in FP64); mixed groups: see test_mixed_group_fp32c_accuracy
  CCOIL_FP32           the reference's float integrands: reported, B far <= 1e-4
"""
import numpy as np
import pytest

import ccoil_abi as abi
import helpers as H
import oracle as O
import workloads as W

pytestmark = pytest.mark.gpu


def _run(ctx, rows, pts, q, prec, flags=0):
    cs = abi.CoilSet.from_rows(rows)
    return ctx.fields(cs, pts, q, prec, flags)


def _truth(params, args, pts, A=True, B=True, G=True):
    coils = [H.oracle_coil(p) for p in params]
    return O.group_fields(coils, args, pts, A, B, G, axis_ratio=O.AXIS_RATIO_GPU, threads=8, extended=True)


def _double(params, args, pts, A=True, B=True, G=True):
    coils = [H.oracle_coil(p) for p in params]
    return O.group_fields(coils, args, pts, A, B, G, axis_ratio=O.AXIS_RATIO_GPU, threads=8)


def test_device_gauss_legendre_matches_leggauss(ctx):
    for n in range(1, 101):
        x, w = ctx.gauss_legendre(n)
        xr, wr = np.polynomial.legendre.leggauss(n)
        assert np.max(np.abs(x - xr)) < 1e-14, n
        assert np.max(np.abs(w - wr)) < 1e-14, n


@pytest.mark.parametrize("prec", [abi.FP64, abi.FP64_IEEE])
@pytest.mark.parametrize("case", ["rect", "mixed"])
def test_fp64_matches_truth(ctx, case, prec):
    if case == "rect":
        params, pts = W.single_rect_workload(65536, seed=0)   # the plain-method workload at its full size
    else:
        params, pts = W.group_workload(12, 2048, seed=1)
    rows, args = H.gpu_args_rows(params)
    if case == "rect":
        assert list(args[0][3:]) == [59, 27, 27]
    out, rep = _run(ctx, rows, pts, abi.A | abi.B | abi.G, prec)
    assert rep.near_pairs == 0
    a, b, g = _truth(params, args, pts)
    near = H.near_mask(params, pts)
    ea, eb = H.rel_err_vec(out["A"], a), H.rel_err_vec(out["B"], b)
    eg = H.rel_err_vec(out["G"], g.reshape(len(pts), 9))
    inside = H.inside_mask(params, pts)
    print(f"\nprecision {prec} {case}: A max {np.nanmax(ea):.3g} B max {np.nanmax(eb):.3g} G max {np.nanmax(eg):.3g} "
          f"(inside the winding {np.nanmax(eg[inside], initial=0):.3g}, {inside.sum()} points; outside "
          f"{np.nanmax(eg[~inside]):.3g}) G far max {np.nanmax(eg[~near]):.3g}")
    assert np.nanmax(ea) < 1e-11
    assert np.nanmax(eb) < 1e-11
    # Within the near-winding shell the gradient's quadrature sum is ill-conditioned: close to the winding's surface
    # (1 mm above the end face of the 0.1 m coil of this workload) single nodes contribute up to ~1e9 times the sum,
    # so a sum of node values each correct to a few ulp carries ~1e-7 relative error (FP64 and FP64_IEEE agree with
    # each other there to 1e-15; a host evaluation of the same double integrands reproduces the same figure).
    assert np.nanmax(eg[near]) < 1e-6
    assert np.nanmax(eg[~near]) < 1e-9


def test_fp64_reference_matches_oracle_double(ctx):
    params, pts = W.group_workload(12, 1024, seed=1)
    rows, args = H.gpu_args_rows(params)
    out, _ = _run(ctx, rows, pts, abi.A | abi.B | abi.E | abi.G, abi.FP64_REFERENCE)
    a, b, g = _double(params, args, pts)
    e_ref = np.zeros_like(a)
    for p, ar in zip(params, args):
        c = H.oracle_coil(p)
        ac, _, _ = O.coil_fields(c, ar, pts, True, False, False, axis_ratio=O.AXIS_RATIO_GPU)
        e_ref += 2 * np.pi * c.freq * ac
    for key, ref in (("A", a), ("B", b), ("E", e_ref)):
        e = H.rel_err_vec(out[key], ref)
        assert np.nanmax(e) < 1e-10, (key, np.nanmax(e))
    # the gradient of the reference integrands agrees with the truth only to their own conditioning
    _, _, gt = _truth(params, args, pts, A=False, B=False)
    eg = H.rel_err_vec(out["G"], gt.reshape(len(pts), 9))
    assert np.nanmax(eg) < 1e-4


def test_single_rect_fp32c_and_fp32(ctx):
    params, pts = W.single_rect_workload(65536, seed=0)
    pts = pts[:4096]
    rows, args = H.gpu_args_rows(params)
    _, b, g = _truth(params, args, pts, A=False)
    near = H.near_mask(params, pts)
    out, rep = _run(ctx, rows, pts, abi.B | abi.G, abi.FP32C)
    eb = H.rel_err_vec(out["B"], b)
    eg = H.rel_err_vec(out["G"], g.reshape(len(pts), 9))
    print(f"\nFP32C rect: B far max {np.nanmax(eb[~near]):.3g} p99 {np.nanpercentile(eb[~near], 99):.3g}; "
          f"G far max {np.nanmax(eg[~near]):.3g}; all-point B max {np.nanmax(eb):.3g}; near pairs {rep.near_pairs}")
    assert rep.near_pairs == H.fp64_pair_count(params, pts, rows=rows)
    _, rep_b = _run(ctx, rows, pts, abi.B, abi.FP32C)   # without the gradient only the winding shell goes to FP64
    assert rep_b.near_pairs == int(near.sum())
    _, rep_d = _run(ctx, rows, pts, abi.B | abi.G, abi.FP32C, abi.DIRECT)   # no far-field series: the axis cone too
    assert rep_d.near_pairs == H.fp64_pair_count(params, pts)
    assert np.nanmax(eb[~near]) < 1e-6
    assert np.nanmax(eg[~near]) < 1e-6
    assert np.nanmax(eb) < 1e-9 or np.nanmax(eb[near]) < 1e-9
    out, _ = _run(ctx, rows, pts, abi.B, abi.FP32)
    eb = H.rel_err_vec(out["B"], b)
    print(f"FP32 rect: B far max {np.nanmax(eb[~near]):.3g}")
    assert np.nanmax(eb[~near]) < 1e-4


@pytest.mark.parametrize("seed", [2, 3])
def test_mixed_group_fp32c_accuracy(ctx, seed):
    """Mixed groups: the summed field can be much smaller than single coils' contributions, so the error is
    also reported relative to sum_k |B_k| (the conditioning of the superposition)."""
    params, pts = W.group_workload(12, 2048, seed=seed)
    rows, args = H.gpu_args_rows(params)
    out, rep = _run(ctx, rows, pts, abi.B | abi.G, abi.FP32C)
    _, b, g = _truth(params, args, pts, A=False)
    near = H.near_mask(params, pts)
    eb = H.rel_err_vec(out["B"], b)
    eg = H.rel_err_vec(out["G"], g.reshape(len(pts), 9))
    scale = np.zeros(len(pts))
    for p, a in zip(params, args):
        _, bk, _ = O.coil_fields(H.oracle_coil(p), a, pts, False, True, False, axis_ratio=O.AXIS_RATIO_GPU)
        scale += np.linalg.norm(bk, axis=1)
    ebn = np.linalg.norm(out["B"] - b, axis=1) / scale
    far = ~near
    print(f"\nFP32C seed={seed}: B far max {np.nanmax(eb[far]):.3g} p99 {np.nanpercentile(eb[far], 99):.3g} "
          f"(vs sum|B_k|: max {np.max(ebn[far]):.3g}) | G far max {np.nanmax(eg[far]):.3g} "
          f"p99 {np.nanpercentile(eg[far], 99):.3g} | near pairs {rep.near_pairs}")
    assert np.max(ebn[far]) < 5e-6
    assert np.nanpercentile(eb[far], 99) < 3e-6
    assert np.nanmax(eb[near]) < 1e-5


def test_gradient_jacobian_convention(ctx):
    """CCOIL_G_JACOBIAN = T G T^T summed over coils; its trace (div B) vanishes."""
    params, pts = W.group_workload(3, 256, seed=4)
    rows, args = H.gpu_args_rows(params)
    o2, _ = _run(ctx, rows, pts, abi.G_JACOBIAN, abi.FP64)
    ref = np.zeros((len(pts), 3, 3))
    for p, a in zip(params, args):
        c = H.oracle_coil(p)
        T = np.array(c.T).reshape(3, 3)
        _, _, g = O.coil_fields(c, a, pts, False, False, True, axis_ratio=O.AXIS_RATIO_GPU, threads=4, extended=True)
        ref += g @ T.T
    e = H.rel_err_vec(o2["G"], ref.reshape(len(pts), 9))
    assert np.nanmax(e) < 1e-7
    Jm = o2["G"].reshape(-1, 3, 3)
    tr = np.abs(np.trace(Jm, axis1=1, axis2=2)) / np.linalg.norm(Jm.reshape(-1, 9), axis=1)
    assert np.median(tr) < 1e-6


@pytest.mark.parametrize("prec", [abi.FP64, abi.FP32C, abi.FP32, abi.FP64_REFERENCE, abi.FP64_IEEE])
def test_quantity_subsets_agree(ctx, prec):
    """Each quantity is the same whether requested alone or together with the others (the gradient's
    rho-phi term reuses the B_rho sum, which must be formed also when B is not requested)."""
    params, pts = W.group_workload(8, 512, seed=12)
    rows, _ = H.gpu_args_rows(params)
    full, _ = _run(ctx, rows, pts, abi.A | abi.B | abi.G, prec)
    for key, bit in (("A", abi.A), ("B", abi.B), ("G", abi.G)):
        alone, _ = _run(ctx, rows, pts, bit, prec)
        e = H.rel_err_vec(alone[key], full[key])
        assert np.nanmax(e) < (1e-12 if prec in (abi.FP64, abi.FP64_REFERENCE, abi.FP64_IEEE) else 1e-5), \
            (key, np.nanmax(e))


def test_device_pointers_with_torch(ctx):
    torch = pytest.importorskip("torch")
    params, pts = W.group_workload(4, 512, seed=5)
    rows, args = H.gpu_args_rows(params)
    host, _ = _run(ctx, rows, pts, abi.B | abi.G, abi.FP32C)
    dp = torch.tensor(pts, device="cuda:0", dtype=torch.float64)
    out = {"B": torch.zeros((len(pts), 3), device="cuda:0", dtype=torch.float64),
           "G": torch.zeros((len(pts), 9), device="cuda:0", dtype=torch.float64)}
    torch.cuda.synchronize()
    ctx.fields(abi.CoilSet.from_rows(rows), dp, abi.B | abi.G, abi.FP32C, abi.DEVICE_POINTERS, out=out)
    assert np.array_equal(out["B"].cpu().numpy(), host["B"])
    assert np.array_equal(out["G"].cpu().numpy(), host["G"])


def test_argument_errors(ctx):
    params, pts = W.group_workload(2, 8, seed=6)
    rows, _ = H.gpu_args_rows(params)
    cs = abi.CoilSet.from_rows(rows)
    with pytest.raises(abi.CCoilError) as e:
        ctx.fields(cs, pts, abi.G | abi.G_JACOBIAN, abi.FP64)
    assert e.value.status == abi.ERR_ARGUMENT
    with pytest.raises(abi.CCoilError):
        ctx.fields(cs, pts, abi.B, 7)
    bad = dict(rows[0])
    bad["n_ang"] = 101
    with pytest.raises(abi.CCoilError):
        ctx.fields(abi.CoilSet.from_rows([bad]), pts, abi.B, abi.FP64)
    out, _ = ctx.fields(cs, pts[:0], abi.B, abi.FP64)
    assert out["B"].shape == (0, 3)


@pytest.mark.parametrize("prec,flags,splits", [(abi.FP64, 0, (1, 4, 16)), (abi.FP32C, abi.DIRECT, (1, 4, 16)),
                                                (abi.FP32C, 0, (1, 1, 16))])
def test_coil_chunking_matches_single_pass(ctx, prec, flags, splits):
    """Smaller calls split a point's thickness rows over 4 or 16 lanes and the coils over grid rows (summed in
    order by field_reduce_kernel); a large call runs one lane and one pass per point. FP32C with the series runs the
    series-only kernel (one lane per point) from 16,384 points on, where the record pass takes the rule's pairs. All
    give the same field to rounding, except that FP32C with the series takes it from the coils' cost floor only with
    one lane per point: a 256-point call keeps the float rule out to 2 r_max (agreement to FP32C's accuracy there)."""
    params, pts = W.group_workload(16, 600000, seed=13)
    rows, _ = H.gpu_args_rows(params)
    big, rep_big = _run(ctx, rows, pts, abi.A | abi.B | abi.G, prec, flags)
    mid, rep_mid = _run(ctx, rows, pts[:40000], abi.A | abi.B | abi.G, prec, flags)
    small, rep_small = _run(ctx, rows, pts[:256], abi.A | abi.B | abi.G, prec, flags)
    assert (rep_big.split, rep_mid.split, rep_small.split) == splits
    small_tol = 1e-5 if (prec == abi.FP32C and not flags & abi.DIRECT) else 1e-12
    for key in ("A", "B", "G"):
        for part, tol in ((small, small_tol), (mid, 1e-12)):
            e = H.rel_err_vec(part[key][:256], big[key][:256])
            assert np.nanmax(e) < tol, (key, np.nanmax(e))


def test_fp32c_near_pass_is_deterministic(ctx):
    """FP32C re-evaluates near-winding (coil, point) pairs in FP64 and adds them per point in coil order (no
    floating-point atomics): repeated calls give bitwise equal results, also where a point is near several coils;
    with CCOIL_NO_NEAR_FP64 nothing is re-evaluated."""
    params, pts = W.group_workload(48, 8192, seed=11)
    pts = pts * 0.4                                   # dense: many points near several coils
    rows, _ = H.gpu_args_rows(params)
    cs = abi.CoilSet.from_rows(rows)
    q = abi.A | abi.B | abi.G
    first, rep = ctx.fields(cs, pts, q, abi.FP32C)
    near = H.near_mask(params, pts)
    assert rep.near_pairs > len(pts) // 10
    for _ in range(3):
        again, rep2 = ctx.fields(cs, pts, q, abi.FP32C)
        assert rep2.near_pairs == rep.near_pairs
        for k in ("A", "B", "G"):
            assert np.array_equal(again[k], first[k]), k
    ref, _ = ctx.fields(cs, pts, q, abi.FP64)
    far_only, rep3 = ctx.fields(cs, pts, q, abi.FP32C, abi.NO_NEAR_FP64)
    assert rep3.near_pairs == 0
    # the FP64 near pass brings the near points to FP32C's far-point accuracy
    e_near = H.rel_err_vec(first["B"], ref["B"])[near]
    e_skip = H.rel_err_vec(far_only["B"], ref["B"])[near]
    assert np.nanmedian(e_near) < np.nanmedian(e_skip)


def test_fp32c_small_call_near_and_far_points(ctx):
    """A small FP32C call (16 lanes per point): near-winding points are re-evaluated with the CCOIL_FP64 arithmetic
    and match CCOIL_FP64; far points keep the FP32 error."""
    rng = np.random.default_rng(61)
    for typ, R, a, b in (("RECTANGULAR", 0.1, 0.05, 0.08), ("FLAT", 0.12, 0.06, 0.0), ("THIN", 0.09, 0.0, 0.1)):
        params = [dict(type=typ, R=R, a=a, b=b, N=300, I=2.5, pos=np.array([0.05, 0.02, -0.03]), alpha=0.6, beta=1.2)]
        rows, _ = H.gpu_args_rows(params)
        pts = np.array(params[0]["pos"]) + rng.uniform(-0.3, 0.3, size=(200, 3))
        near = H.near_mask(params, pts)
        assert 0 < near.sum() < len(pts)
        o64, _ = _run(ctx, rows, pts, abi.B | abi.G, abi.FP64)
        o32, rep = _run(ctx, rows, pts, abi.B | abi.G, abi.FP32C)
        assert rep.split == 16 and rep.near_pairs == H.fp64_pair_count(params, pts, rows=rows)
        eb = H.rel_err_vec(o32["B"], o64["B"])
        eg = H.rel_err_vec(o32["G"], o64["G"])
        assert np.nanmax(eb[near]) < 1e-13 and np.nanmax(eg[near]) < 1e-12, typ
        assert np.nanmax(eb[~near]) < 1e-5 and np.nanmax(eg[~near]) < 1e-4, typ
