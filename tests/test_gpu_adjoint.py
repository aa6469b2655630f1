"""GPU tests of the adjoint (ccoil_field_gradient) and its torch.autograd.Function.

Checks:
  - every one of the 9 per-coil derivatives (current, R, a, b, position, alpha, beta) of L = sum_x g(x) . B(x)
    for a fixed random g against central finite differences (h = 1e-6 relative) of the oracle's long double field
    with frozen quadrature orders, for all four coil types, rotated; points outside the near-winding shell;
    tolerance 1e-6 relative to the largest derivative of the same kind;
  - FP32C adjoint against the FP64 adjoint (1e-5);
  - the design loss of the coil-design workload through torch autograd equals the adjoint called directly, and
    matches central differences of the GPU forward (FP64) to 1e-5.
"""
import numpy as np
import pytest

import ccoil_abi as abi
import helpers as H
import oracle as O
import workloads as W

pytestmark = pytest.mark.gpu

KEYS = abi.GRAD_TERMS


def _oracle_B(p, args, pts):
    c = H.oracle_coil(p)
    _, b, _ = O.coil_fields(c, args, pts, False, True, False, axis_ratio=O.AXIS_RATIO_GPU, threads=8,
                            extended=True)
    return b


def _perturbed(p, key, h):
    q = dict(p)
    q["pos"] = np.array(p["pos"], dtype=float)
    if key == "current":
        q["I"] = p["I"] + h
    elif key in ("px", "py", "pz"):
        q["pos"][("px", "py", "pz").index(key)] += h
    else:
        q[key] = p[key] + h
    return q


def _value(p, key):
    if key == "current":
        return p["I"]
    if key in ("px", "py", "pz"):
        return p["pos"][("px", "py", "pz").index(key)]
    return p[key]


@pytest.mark.parametrize("typ,R,a,b", [("RECTANGULAR", 0.1, 0.05, 0.08), ("THIN", 0.12, 0.0, 0.1),
                                       ("FLAT", 0.1, 0.06, 0.0), ("FILAMENT", 0.09, 0.0, 0.0)])
def test_adjoint_matches_finite_differences(ctx, typ, R, a, b):
    rng = np.random.default_rng(21)
    p = dict(type=typ, R=R, a=a, b=b, N=100, I=2.0, pos=np.array([0.1, -0.05, 0.02]), alpha=0.4, beta=1.1)
    rows, args = H.gpu_args_rows([p])
    pts = W.cube_points(rng, 512, 0.5)
    pts = pts[~H.near_mask([p], pts)]
    g = rng.standard_normal(pts.shape)
    grad, rep = ctx.field_gradient(abi.CoilSet.from_rows(rows), pts, g, abi.FP64)
    fd = {}
    for key in KEYS:
        if key in ("a", "b") and p[key] == 0.0:
            continue                       # clamped thin dimension: not a free parameter
        h = 1e-6 * max(abs(_value(p, key)), 0.1 if key in ("px", "py", "pz") else 1.0)
        bp = _oracle_B(_perturbed(p, key, h), args[0], pts)
        bm = _oracle_B(_perturbed(p, key, -h), args[0], pts)
        fd[key] = float(np.sum(g * (bp - bm)) / (2 * h))
    pos = ("px", "py", "pz")
    scale = {k: max(abs(fd[x]) for x in (pos if k in pos else (k,))) for k in fd}
    for k, v in fd.items():
        got = grad[0, KEYS.index(k)]
        print(f"{typ} d/d{k}: adjoint {got:.12e} fd {v:.12e}")
        assert abs(got - v) <= 1e-6 * max(abs(v), 1e-3 * scale[k]), (k, got, v)


@pytest.mark.parametrize("n_coils,n_points,seed", [(16, 1024, 3), (64, 4096, 0)])
def test_adjoint_fp32c_against_fp64(ctx, n_coils, n_points, seed):
    """On a small set and on the design workload itself (64 coils, 4,096 points, the bench's adjoint leg)."""
    params, pts = W.adjoint_workload(n_coils, n_points, seed=seed)
    rows = W.group_rows(params)
    rng = np.random.default_rng(4)
    g = rng.standard_normal(pts.shape)
    cs = abi.CoilSet.from_rows(rows)
    g64, _ = ctx.field_gradient(cs, pts, g, abi.FP64)
    gie, _ = ctx.field_gradient(cs, pts, g, abi.FP64_IEEE)
    g32, _ = ctx.field_gradient(cs, pts, g, abi.FP32C)
    for k in range(len(KEYS)):
        s = np.max(np.abs(g64[:, k]))
        assert np.max(np.abs(gie[:, k] - g64[:, k])) <= 1e-12 * s, KEYS[k]
        assert np.max(np.abs(g32[:, k] - g64[:, k])) <= 1e-5 * s, KEYS[k]


def test_zero_current_and_empty_inputs(ctx):
    params, pts = W.adjoint_workload(3, 64, seed=5)
    rows = W.group_rows(params)
    g = np.ones_like(pts)
    ref, _ = ctx.field_gradient(abi.CoilSet.from_rows(rows), pts, g, abi.FP64)
    rows0 = [dict(r) for r in rows]
    I = rows0[1]["current"]
    rows0[1]["current"], rows0[1]["J"] = 0.0, 0.0
    got, _ = ctx.field_gradient(abi.CoilSet.from_rows(rows0), pts, g, abi.FP64)
    assert np.allclose(got[0], ref[0], rtol=1e-13) and np.allclose(got[2], ref[2], rtol=1e-13)
    assert got[1, 0] == pytest.approx(ref[1, 0], rel=1e-12)      # dL/dI does not depend on I
    assert np.all(got[1, 1:] == 0.0)
    empty, _ = ctx.field_gradient(abi.CoilSet.from_rows(rows), pts[:0], g[:0], abi.FP64)
    assert np.all(empty == 0.0)
    with pytest.raises(abi.CCoilError):
        ctx.field_gradient(abi.CoilSet.from_rows(rows), pts, g, abi.FP32)


def test_torch_design_loss_gradient(ctx):
    torch = pytest.importorskip("torch")
    import ccoil_torch as CT
    params, pts = W.adjoint_workload(64, 4096, seed=0)
    rows = W.group_rows(params)
    model = CT.CoilFieldModel(rows, ctx, abi.FP64)
    theta = model.initial_params("cuda").requires_grad_(True)
    x = torch.tensor(pts, dtype=torch.float64, device="cuda")
    loss = CT.design_loss(model.B(theta, x), x)
    loss.backward()
    got = theta.grad.cpu().numpy()
    # the same gradient from the adjoint called directly with dL/dB of the design loss
    with torch.no_grad():
        B = model.B(theta, x)
    Bq = B.clone().requires_grad_(True)
    CT.design_loss(Bq, x).backward()
    direct, _ = ctx.field_gradient(model.coilset(theta), pts, Bq.grad.cpu().numpy(), abi.FP64)
    assert np.allclose(got, direct, rtol=1e-12, atol=0)
    # central differences of the GPU forward for a few coils and every parameter
    rng = np.random.default_rng(6)
    for ci in rng.choice(len(rows), 4, replace=False):
        for k in range(len(KEYS)):
            base = theta.detach().clone()
            h = 1e-6 * max(abs(float(base[ci, k])), 0.05 if KEYS[k] in ("px", "py", "pz") else 1.0)
            if KEYS[k] == "a" and params[ci]["type"] == "THIN":
                continue                   # clamped thin dimension: not a free parameter
            vals = []
            for sgn in (1, -1):
                th = base.clone()
                th[ci, k] += sgn * h
                with torch.no_grad():
                    vals.append(float(CT.design_loss(model.B(th, x), x)))
            fd = (vals[0] - vals[1]) / (2 * h)
            scale = np.max(np.abs(got[:, k]))
            assert abs(got[ci, k] - fd) <= 1e-5 * max(abs(fd), 1e-2 * scale), (ci, KEYS[k], got[ci, k], fd)
