"""Points placed on the classification boundaries of the two-pass calls: the near-winding shell (distance max(a, b,
1e-3 R) from the cross-section rectangle), the axis cone (r^2 = 0.01 (r^2 + z^2 + R^2)) and the far-field distance.

field_kernel and the record pass (near_count / near_fill) classify each (coil, point) pair separately; a pair that
the two classified differently would leave a slot unwritten or fill it twice, so the library checks the fill
against the field kernel's counts and fails the call with CCOIL_ERR_DEVICE. Here the points lie within a few ulps
of each boundary, on both sides, for an unrotated coil at the origin (local and global coordinates agree, so some
points lie exactly on a boundary) and for rotated coils of every type:

  - every call succeeds (no classification mismatch), at sizes below and above the size that turns the visit order
    and the record pass on;
  - one coil at a time, CCOIL_FP32C within 1e-4 (B) and 1e-3 (gradient) of that coil's CCOIL_FP64 field at every
    point: a dropped or doubled contribution would be an error of order 1;
  - CCOIL_FP64 with CCOIL_FAR_FIELD within 1e-9 of the plain rule, and the group calls within 1e-3 of CCOIL_FP64.
"""
import numpy as np
import pytest

import ccoil_abi as abi
import helpers as H

pytestmark = pytest.mark.gpu

COILS = [
    dict(type="RECTANGULAR", R=0.1, a=0.05, b=0.08, pos=(0.0, 0.0, 0.0), alpha=0.0, beta=0.0),
    dict(type="RECTANGULAR", R=0.08, a=0.03, b=0.12, pos=(0.9, -0.4, 0.3), alpha=0.7, beta=1.9),
    dict(type="THIN", R=0.12, a=0.0, b=0.2, pos=(-0.8, 0.5, -0.2), alpha=2.1, beta=-0.6),
    dict(type="FLAT", R=0.05, a=0.1, b=0.0, pos=(0.3, 0.9, -0.7), alpha=1.2, beta=0.4),
    dict(type="FILAMENT", R=0.1, a=0.0, b=0.0, pos=(-0.4, -0.9, 0.6), alpha=-0.9, beta=2.7),
]


def _params(c):
    return dict(R=c["R"], a=c["a"], b=c["b"], N=250, I=3.0, pos=np.array(c["pos"]), alpha=c["alpha"],
                beta=c["beta"])


def _local_boundary_points(rng, c, row, n):
    """(n, 3) local (x, y, z) on the three boundaries, n // 3 each."""
    R, a, b = c["R"], c["a"], c["b"]
    w = max(a, b, 1e-3 * R)
    m = n // 3
    # near shell: the rectangle [R, R + a] x [-b/2, b/2] grown by w (sides, ends, rounded corners)
    t = rng.uniform(0.0, 1.0, m)
    part = rng.integers(0, 4, m)
    ang = rng.uniform(0.0, 0.5 * np.pi, m)
    r_sh = np.where(part == 0, R + a + w, np.where(part == 1, R + a * t, R + a + w * np.cos(ang)))
    z_sh = np.where(part == 0, (t - 0.5) * b, np.where(part == 1, 0.5 * b + w, 0.5 * b + w * np.sin(ang)))
    inner = (part == 3) & (R - w > 0)
    r_sh = np.where(inner, R - w, r_sh)
    z_sh = np.where(inner, (t - 0.5) * b, z_sh)
    z_sh *= rng.choice([-1.0, 1.0], m)
    # axis cone: r^2 = 0.01 (z^2 + R^2) / 0.99
    z_ax = rng.uniform(-3.0 * R, 3.0 * R, m)
    r_ax = np.sqrt(0.01 * (z_ax * z_ax + R * R) / 0.99)
    # far-field distance (of a call of n points)
    _, dfar = H.far_radius(row, n)
    th = rng.uniform(0.0, np.pi, n - 2 * m)
    r_far, z_far = dfar * np.sin(th), dfar * np.cos(th)
    r = np.concatenate([r_sh, r_ax, r_far])
    z = np.concatenate([z_sh, z_ax, z_far])
    phi = rng.uniform(0.0, 2.0 * np.pi, n)
    if c["alpha"] == 0.0 and c["beta"] == 0.0:
        phi[: n // 4] = 0.0   # on the x axis: r and z are exactly the boundary's values in the kernels too
    return np.stack([r * np.cos(phi), r * np.sin(phi), z], axis=1)


def _boundary_points(rng, c, row, n):
    loc = _local_boundary_points(rng, c, row, n)
    Tinv = np.array(H.oracle_coil(_params(c)).Tinv).reshape(3, 3)
    pts = np.asarray(c["pos"]) + loc @ Tinv
    # a few ulps either side (the exact points of the unrotated coil are kept)
    jit = rng.integers(-3, 4, pts.shape) * np.spacing(np.abs(pts) + 1e-300)
    if c["alpha"] == 0.0 and c["beta"] == 0.0:
        jit[: n // 8] = 0.0
    return pts + jit


def _rows():
    params = [_params(c) for c in COILS]
    rows, _ = H.gpu_args_rows(params)
    return params, rows


def _err(got, ref):
    n = len(ref)
    num = np.linalg.norm((got - ref).reshape(n, -1), axis=1)
    den = np.linalg.norm(ref.reshape(n, -1), axis=1)
    return num / den


@pytest.mark.parametrize("n_per_coil", [3000, 24000])
def test_points_on_the_classification_boundaries(ctx, n_per_coil):
    rng = np.random.default_rng(21)
    params, rows = _rows()
    q = abi.B | abi.G
    per_coil = []
    for k, (c, row) in enumerate(zip(COILS, rows)):
        assert H.far_radius(row)[1] > 0, c["type"]
        pts = _boundary_points(rng, c, row, n_per_coil)
        per_coil.append(pts)
        cs = abi.CoilSet.from_rows([row])
        print(f"\n[boundaries] coil {k} ({c['type']}), {len(pts)} points", flush=True)
        ref, _ = ctx.fields(cs, pts, q, abi.FP64)
        got, rep = ctx.fields(cs, pts, q, abi.FP32C)
        assert rep.near_pairs > 0
        eb, eg = _err(got["B"], ref["B"]), _err(got["G"], ref["G"])
        print(f"  FP32C: B max {eb.max():.3g}, gradient max {eg.max():.3g}, {rep.near_pairs} FP64 pairs")
        assert eb.max() <= 1e-4 and eg.max() <= 1e-3, (k, eb.max(), eg.max())
        ser, _ = ctx.fields(cs, pts, q, abi.FP64, abi.FAR_FIELD)
        es = max(_err(ser["B"], ref["B"]).max(), _err(ser["G"], ref["G"]).max())
        print(f"  FP64 with the series: max {es:.3g}")
        assert es <= 1e-9, (k, es)
    # all coils over all boundary points at once
    cs = abi.CoilSet.from_rows(rows)
    pts = np.concatenate(per_coil)
    print(f"[boundaries] group of {len(rows)}, {len(pts)} points", flush=True)
    ref, _ = ctx.fields(cs, pts, q, abi.FP64)
    for prec, flags, tol in ((abi.FP32C, 0, 1e-3), (abi.FP64, abi.FAR_FIELD, 1e-9)):
        got, _ = ctx.fields(cs, pts, q, prec, flags)
        e = max(_err(got["B"], ref["B"]).max(), _err(got["G"], ref["G"]).max())
        print(f"  precision {prec} flags {flags}: max {e:.3g}")
        assert e <= tol, (prec, flags, e)
