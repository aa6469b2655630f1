"""Plans (a field or pair call captured into a HIP graph and replayed) and the autotuner, through the C ABI.

  - A field plan (CCOIL_FP64 and CCOIL_FP32C with its near pass, B and the gradient) and a pair plan (M, F, T) are
    replayed several times with eager library calls, torch work and synchronisations between the replays; every
    replay must equal the eager call bitwise (same kernels, same launch shape). Writing new point coordinates into
    the plan's input buffer and replaying gives the eager result for those points.
  - An FP32C plan made on points with no near pairs and replayed on points that have many more than it holds
    reports CCOIL_ERR_CAPACITY.
  - CCOIL_AUTOTUNE: the first call of a size times the candidate shapes and caches the fastest, in memory and in the
    tuning file; later calls, and a new context that loads the file, replay that shape: bitwise equal results, and
    within 1e-12 of the rule's shape (a different order of the same sums).
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


def _busy_work(ctx, cs, x, q, prec):
    """Eager work between replays: torch kernels and an eager library call into other buffers."""
    y = torch.randn(1 << 20, device="cuda", dtype=torch.float64)
    (y * 2.0 + 1.0).sum().item()
    ctx.fields(cs, x, q, prec, abi.DEVICE_POINTERS, out=_outs(x.shape[0], q))
    torch.cuda.synchronize()


@pytest.mark.parametrize("prec", [abi.FP64, abi.FP32C])
def test_field_plan_replays_match_eager(ctx, prec):
    params, pts = W.group_workload(16, 5000, seed=31)
    cs = abi.CoilSet.from_rows(W.group_rows(params))
    q = abi.B | abi.G
    x = torch.from_numpy(pts).to("cuda")
    ref, rrep = ctx.fields(cs, x, q, prec, abi.DEVICE_POINTERS, out=_outs(len(pts), q))
    out = _outs(len(pts), q)
    plan = ctx.fields_plan(cs, x, q, prec, out)
    for k in range(4):
        for t in out.values():
            t.fill_(float("nan"))
        plan.launch()
        rep = plan.synchronize()
        assert rep.near_pairs == rrep.near_pairs and rep.node_evaluations == rrep.node_evaluations
        for key in ref:
            assert torch.equal(out[key], ref[key]), (k, key)
        _busy_work(ctx, cs, x, q, prec)
    # new point coordinates in the same buffer
    pts2 = W.group_workload(16, 5000, seed=32)[1]
    x2 = torch.from_numpy(pts2).to("cuda")
    ref2, _ = ctx.fields(cs, x2, q, prec, abi.DEVICE_POINTERS, out=_outs(len(pts2), q))
    x.copy_(x2)
    torch.cuda.synchronize()
    plan.launch()
    plan.synchronize()
    for key in ref2:
        assert torch.equal(out[key], ref2[key]), key
    plan.close()


def test_pair_plan_replays_match_eager(ctx):
    prow, srow = W.pair_workload(300, seed=33)
    P, S = abi.CoilSet.from_rows(prow), abi.CoilSet.from_rows(srow)
    what = abi.PAIR_M | abi.PAIR_FT
    n = len(prow)
    ref, _ = ctx.pairs(P, S, n, what, abi.FP64)
    out = {"M": torch.full((n,), float("nan"), dtype=torch.float64, device="cuda"),
           "F": torch.full((n, 3), float("nan"), dtype=torch.float64, device="cuda"),
           "T": torch.full((n, 3), float("nan"), dtype=torch.float64, device="cuda")}
    plan = ctx.pairs_plan(P, S, n, what, abi.FP64, out)
    params, pts = W.group_workload(4, 512, seed=34)
    cs = abi.CoilSet.from_rows(W.group_rows(params))
    x = torch.from_numpy(pts).to("cuda")
    for k in range(3):
        for t in out.values():
            t.fill_(float("nan"))
        plan.launch()
        rep = plan.synchronize()
        assert rep.node_evaluations > 0
        for key in ("M", "F", "T"):
            assert np.array_equal(out[key].cpu().numpy(), ref[key]), (k, key)
        _busy_work(ctx, cs, x, abi.B, abi.FP64)
    plan.close()


def test_fp32c_plan_reports_capacity(ctx):
    params, _ = W.group_workload(6, 8, seed=35)
    rows = W.group_rows(params)
    cs = abi.CoilSet.from_rows(rows)
    n = 6000
    far = np.random.default_rng(36).uniform(20.0, 30.0, (n, 3))     # far from every coil: no near pairs
    x = torch.from_numpy(far).to("cuda")
    out = _outs(n, abi.B)
    plan = ctx.fields_plan(cs, x, abi.B, abi.FP32C, out)
    plan.launch()
    assert plan.synchronize().near_pairs == 0
    # every point on the first coil's winding (mid-radius, mid-length): each is near that coil
    c = rows[0]
    T = W.rotation(c["alpha"], c["beta"])
    phi = np.linspace(0.0, 2 * np.pi, n, endpoint=False)
    rho = c["R"] + 0.5 * c["a"]
    loc = np.stack([rho * np.cos(phi), rho * np.sin(phi), np.zeros(n)], axis=1)
    x.copy_(torch.from_numpy(loc @ T.T + np.array(c["pos"])))
    torch.cuda.synchronize()
    plan.launch()
    with pytest.raises(abi.CCoilError) as err:
        plan.synchronize()
    assert err.value.status == abi.ERR_CAPACITY
    plan.close()


def test_autotune_cache_replays_the_same_shape(ctx, tmp_path):
    params, pts = W.group_workload(20, 20000, seed=37)
    cs = abi.CoilSet.from_rows(W.group_rows(params))
    q = abi.B | abi.G
    x = torch.from_numpy(pts).to("cuda")
    path = tmp_path / "tuning.txt"
    ctx.set_tuning_file(str(path))
    try:
        rule, _ = ctx.fields(cs, x, q, abi.FP64, abi.DEVICE_POINTERS, out=_outs(len(pts), q))
        t1, _ = ctx.fields(cs, x, q, abi.FP64, abi.DEVICE_POINTERS | abi.AUTOTUNE, out=_outs(len(pts), q))
        lines = path.read_text().splitlines()
        assert len(lines) == 1 and lines[0].startswith("fields ") and "split=" in lines[0], lines
        t2, _ = ctx.fields(cs, x, q, abi.FP64, abi.DEVICE_POINTERS | abi.AUTOTUNE, out=_outs(len(pts), q))
        assert len(path.read_text().splitlines()) == 1           # the second call replays the cached shape
        for k in t1:
            assert torch.equal(t1[k], t2[k]), k
            scale = rule[k].abs().max()
            assert float((t1[k] - rule[k]).abs().max()) <= 1e-12 * float(scale), k
        # a new context loading the file takes the same shape without timing
        ctx2 = abi.Context(0)
        ctx2.set_tuning_file(str(path))
        t3, _ = ctx2.fields(cs, x, q, abi.FP64, abi.DEVICE_POINTERS | abi.AUTOTUNE, out=_outs(len(pts), q))
        for k in t1:
            assert torch.equal(t1[k], t3[k]), k
        # pairs: the lane target
        prow, srow = W.pair_workload(200, seed=38)
        P, S = abi.CoilSet.from_rows(prow), abi.CoilSet.from_rows(srow)
        p_rule, _ = ctx2.pairs(P, S, len(prow), abi.PAIR_M, abi.FP64)
        p1, _ = ctx2.pairs(P, S, len(prow), abi.PAIR_M, abi.FP64, flags=abi.AUTOTUNE)
        p2, _ = ctx2.pairs(P, S, len(prow), abi.PAIR_M, abi.FP64, flags=abi.AUTOTUNE)
        assert any(l.startswith("pairs ") for l in path.read_text().splitlines())
        assert np.array_equal(p1["M"], p2["M"])
        assert np.allclose(p1["M"], p_rule["M"], rtol=1e-12, atol=0)
        ctx2.close()
    finally:
        ctx.set_tuning_file(None)
