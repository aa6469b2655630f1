"""ccoil_fields_noalloc (the field entry point that allocates nothing) against ccoil_fields through the C ABI.

Both run the same kernels with the same launch shape, so the outputs must be bitwise equal. Two calls in a row, the
second with fewer points and fewer coils, check that the shared scratch keeps nothing that a later call depends on
(counts, cursors and the near pass's state are cleared by every call). Precisions: CCOIL_FP64 (one and several coil
chunks) and CCOIL_FP32C (the near pass in the scratch). A call too large for the 64 MiB scratch is rejected with
CCOIL_ERR_ARGUMENT before anything is launched into it.
"""
import numpy as np
import pytest

import ccoil_abi as abi
import workloads as W

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _outs(n, q):
    out = {}
    for key, bit, w in (("A", abi.A, 3), ("B", abi.B, 3), ("E", abi.E, 3), ("G", abi.G | abi.G_JACOBIAN, 9)):
        if q & bit:
            out[key] = torch.full((n, w), float("nan"), dtype=torch.float64, device="cuda")
    return out


@pytest.mark.parametrize("prec,q", [(abi.FP64, abi.A | abi.B | abi.E), (abi.FP64, abi.B | abi.G),
                                    (abi.FP32C, abi.B | abi.G), (abi.FP32C, abi.A | abi.B)])
def test_noalloc_matches_fields(ctx, prec, q):
    params, pts = W.group_workload(24, 6000, seed=21)
    rows = W.group_rows(params)
    for n_coils, n_pts in ((24, 6000), (9, 1500)):          # the second call smaller in both
        cs = abi.CoilSet.from_rows(rows[:n_coils])
        x = torch.from_numpy(np.ascontiguousarray(pts[:n_pts])).to("cuda")
        ref, rrep = ctx.fields(cs, x, q, prec, abi.DEVICE_POINTERS, out=_outs(n_pts, q))
        got, grep = ctx.fields_noalloc(cs, x, q, prec, _outs(n_pts, q))
        torch.cuda.synchronize()
        print(f"\nprecision {prec} q {q}: {n_coils} coils x {n_pts} points, near pairs {grep.near_pairs}, "
              f"split {grep.split}")
        assert grep.near_pairs == rrep.near_pairs and grep.split == rrep.split
        for k in ref:
            assert torch.isfinite(got[k]).all(), k
            assert torch.equal(got[k], ref[k]), (k, float((got[k] - ref[k]).abs().max()))
        if prec == abi.FP32C and n_coils == 24:
            assert grep.near_pairs > 0


def test_noalloc_rejects_oversized_calls(ctx):
    params, _ = W.group_workload(4, 8, seed=22)
    cs = abi.CoilSet.from_rows(W.group_rows(params))
    n = 1 << 22   # FP32C needs 16 bytes of counts and offsets per point alone: 64 MiB
    x = torch.zeros((n, 3), dtype=torch.float64, device="cuda")
    with pytest.raises(abi.CCoilError) as err:
        ctx.fields_noalloc(cs, x, abi.B, abi.FP32C, _outs(n, abi.B))
    assert err.value.status == abi.ERR_ARGUMENT and "scratch" in str(err.value)
    # host pointers are refused
    with pytest.raises(abi.CCoilError):
        abi.lib()
        ctx._check(abi.lib().ccoil_fields_noalloc(ctx.h, None, None, 0, abi.B, abi.FP64, 0, None, None, None, None,
                                                  None))
