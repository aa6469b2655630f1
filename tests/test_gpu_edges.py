"""GPU edge cases of the field and pair paths against the oracle's long double evaluation of the same rule.

  - points on the coil axis (the GPU gradient's on-axis branch r/R <= 1e-5, FieldGGroup.cu:134), at the centre, on the
    end planes z = +-b/2 and mirrored in z (the fast rule is evaluated at |z| with odd terms flipped);
  - points far away (10^3 R), where the angular sums extract a small first harmonic;
  - empty inputs: no points, no pairs, no coils (a CoilGroup without members has zero field).
Tolerances as in test_gpu_fields.py: CCOIL_FP64 A, B <= 1e-11; gradient <= 1e-7. CCOIL_FP32C on the same points
(its FP64 pass and series where they apply, the float rule elsewhere): 5e-6, and 1e-11 / 1e-9 on the axis and far
points.
"""
import numpy as np
import pytest

import ccoil_abi as abi
import helpers as H
import oracle as O
import workloads as W

pytestmark = pytest.mark.gpu


def _coils():
    base = dict(N=100, I=1.5, freq=50.0, pos=np.array([0.1, -0.2, 0.05]), alpha=0.7, beta=1.9)
    return [dict(base, type="RECTANGULAR", R=0.1, a=0.05, b=0.08), dict(base, type="THIN", R=0.12, a=0.0, b=0.1),
            dict(base, type="FLAT", R=0.09, a=0.06, b=0.0), dict(base, type="FILAMENT", R=0.11, a=0.0, b=0.0)]


def _local_points(p, rng):
    """Points given in the coil frame (z along the axis), mapped to global coordinates."""
    c = H.oracle_coil(p)
    T = np.array(c.T).reshape(3, 3)
    R, a, b = c.R, c.a, c.b
    loc = [(0, 0, 0), (0, 0, 0.3 * R), (0, 0, -0.3 * R), (0, 0, 2 * R), (0, 0, -5 * R),
           (1e-7 * R, 0, 0.4 * R), (0, 2e-6 * R, -0.4 * R)]                      # on and next to the axis
    for sgn in (1, -1):                                                         # end planes and mirror images
        loc += [(0.5 * R, 0, sgn * 0.5 * b), (1.5 * (R + a), 0.1 * R, sgn * 0.5 * b), (0.3 * R, 0.2 * R, sgn * 0.7 * R)]
    far = rng.standard_normal((8, 3))
    loc += list(1e3 * R * far / np.linalg.norm(far, axis=1, keepdims=True))
    loc = np.array(loc, dtype=float)
    return loc @ T.T + np.array(p["pos"])


@pytest.mark.parametrize("prec", [abi.FP64, abi.FP64_IEEE])
def test_axis_end_planes_mirror_and_far_points(ctx, prec):
    rng = np.random.default_rng(91)
    for p in _coils():
        rows, args = H.gpu_args_rows([p])
        pts = _local_points(p, rng)
        out, _ = ctx.fields(abi.CoilSet.from_rows(rows), pts, abi.A | abi.B | abi.G, prec)
        a, b, g = O.coil_fields(H.oracle_coil(p), args[0], pts, True, True, True, axis_ratio=O.AXIS_RATIO_GPU,
                                extended=True)
        # the vector potential vanishes on the axis, where both sides hold only rounding residue (~1e-16 of the
        # off-axis scale): below 1e-4 of that scale the error is taken relative to 1e-4 of the scale
        scale_a = np.max(np.linalg.norm(a, axis=1))
        ea = np.linalg.norm(out["A"] - a, axis=1) / np.maximum(np.linalg.norm(a, axis=1), 1e-4 * scale_a)
        assert np.max(ea) < 1e-11, (p["type"], ea)
        assert np.nanmax(H.rel_err_vec(out["B"], b)) < 1e-11, p["type"]
        assert np.nanmax(H.rel_err_vec(out["G"], g.reshape(-1, 9))) < 1e-7, p["type"]


def test_fp32c_axis_end_planes_mirror_and_far_points(ctx):
    """CCOIL_FP32C on the same points: the axis points go to the FP64 pass (the gradient's axis cone) or take the
    FP64 series, the points at 10^3 R take the series, the end-plane points beside the winding are in its near shell;
    the rest are float-rule points. A, B, gradient within 5e-6 everywhere; the axis points and the far points within
    1e-11 (A, B) and 1e-9 (gradient)."""
    rng = np.random.default_rng(91)
    for p in _coils():
        rows, args = H.gpu_args_rows([p])
        pts = _local_points(p, rng)
        out, rep = ctx.fields(abi.CoilSet.from_rows(rows), pts, abi.A | abi.B | abi.G, abi.FP32C)
        a, b, g = O.coil_fields(H.oracle_coil(p), args[0], pts, True, True, True, axis_ratio=O.AXIS_RATIO_GPU,
                                extended=True)
        scale_a = np.max(np.linalg.norm(a, axis=1))
        ea = np.linalg.norm(out["A"] - a, axis=1) / np.maximum(np.linalg.norm(a, axis=1), 1e-4 * scale_a)
        eb = H.rel_err_vec(out["B"], b)
        eg = H.rel_err_vec(out["G"], g.reshape(-1, 9))
        exact = np.r_[np.arange(7), np.arange(len(pts) - 8, len(pts))]   # axis points, points at 10^3 R
        print(f"\nFP32C {p['type']}: A {ea.max():.3g} B {np.nanmax(eb):.3g} G {np.nanmax(eg):.3g}; axis and far "
              f"points B {np.nanmax(eb[exact]):.3g} G {np.nanmax(eg[exact]):.3g}; near pairs {rep.near_pairs}")
        assert ea.max() < 5e-6 and np.nanmax(eb) < 5e-6 and np.nanmax(eg) < 5e-6, p["type"]
        assert ea[exact].max() < 1e-11 and np.nanmax(eb[exact]) < 1e-11 and np.nanmax(eg[exact]) < 1e-9, p["type"]


def test_empty_inputs(ctx):
    rows, _ = H.gpu_args_rows(_coils())
    cs = abi.CoilSet.from_rows(rows)
    pts = np.zeros((0, 3))
    out, rep = ctx.fields(cs, pts, abi.A | abi.B | abi.G, abi.FP64)
    assert out["B"].shape == (0, 3) and out["G"].shape == (0, 9) and rep.node_evaluations == 0
    none = abi.CoilSet.from_rows([])
    pts = np.random.default_rng(92).uniform(-1, 1, (17, 3))
    out = {"A": np.full((17, 3), np.nan), "B": np.full((17, 3), np.nan), "G": np.full((17, 9), np.nan)}
    out, _ = ctx.fields(none, pts, abi.A | abi.B | abi.G, abi.FP64, out=out)
    assert np.all(out["A"] == 0.0) and np.all(out["B"] == 0.0) and np.all(out["G"] == 0.0)
    prow, srow = W.pair_workload(2, seed=3)
    P, S = abi.CoilSet.from_rows(prow), abi.CoilSet.from_rows(srow)
    res, rep = ctx.pairs(P, S, 0, abi.PAIR_M | abi.PAIR_FT, abi.FP64)
    assert res["M"].shape == (0,) and rep.node_evaluations == 0
    res, _ = ctx.pairs_zaxis(P, S, 0, abi.PAIR_M | abi.PAIR_FT)
    assert res["M"].shape == (0,) and res["Fz"].shape == (0,)
    L, _ = ctx.self_inductance(none, np.zeros((0, 2)))
    assert L.shape == (0,)
