"""CCOIL_FP32C's far-field series (beyond a coil's far-field distance the exterior zonal-harmonic series in FP64
replaces the float rule) against the rule.

  - The series against CCOIL_FP64 (the rule in double) at points beyond the threshold, for every coil type, thin and
    thick windings, thickness orders 1-20 (the series sums the rule's own thickness nodes) and angular orders from
    the threshold's bound: A, B and the gradient within 1e-11 of the coil's field, and within 1e-6 even for an
    angular order (12) whose threshold lies far out. No near pairs there.
  - Coils of fewer than 5 angular increments never take the series: bitwise the CCOIL_DIRECT results.
  - Reproducibility: repeated calls are bitwise equal, and so are permuted points (the visit order, a spatial sort
    with atomics, decides only which wave holds a point).
  - A mixed group at 65,536 points (the visit order on) against the FP64 rule away from the windings: the FP64
    series (CCOIL_FAR_FIELD) within 1e-9 per point; FP32C with the series at p99 <= 1e-6, max <= 1e-5.
  - CCOIL_FP64 with CCOIL_FAR_FIELD against the plain rule: within 1e-12 beyond the threshold, bitwise inside it.
"""
import numpy as np
import pytest

import ccoil_abi as abi
import helpers as H
import workloads as W

pytestmark = pytest.mark.gpu


def _coil(typ, R, a, b, n_ang, n_thk, pos=(0.03, -0.02, 0.05), alpha=0.7, beta=1.9):
    p = dict(type=typ, R=R, a=a, b=b, N=250, I=3.0, pos=np.array(pos), alpha=alpha, beta=beta)
    row = W.coil_row(W.make_coil(p))
    row.update(n_ang=n_ang, n_thk=n_thk if typ in ("RECTANGULAR", "FLAT") else 1, blocks=(1, 1, 1))
    return p, row


def _shell_points(rng, row, y, n):
    """n points at distance y r_max from the coil's centre, directions uniform."""
    rmax, _ = H.far_radius(row)
    d = rng.standard_normal((n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return np.asarray(row["pos"]) + y * rmax * d


@pytest.mark.parametrize("typ,R,a,b", [("RECTANGULAR", 0.1, 0.05, 0.08), ("RECTANGULAR", 0.04, 0.2, 0.15),
                                       ("THIN", 0.1, 0.0, 0.3), ("FLAT", 0.02, 0.2, 0.0), ("FILAMENT", 0.1, 0.0, 0.0)])
def test_series_matches_the_fp64_rule(ctx, typ, R, a, b):
    rng = np.random.default_rng(7)
    worst = 0.0
    for n_ang, n_thk in ((60, 1), (48, 7), (32, 20), (12, 9)):
        p, row = _coil(typ, R, a, b, n_ang, n_thk)
        rmax, dfar = H.far_radius(row)
        assert dfar > 0, (typ, n_ang)
        cs = abi.CoilSet.from_rows([row])
        for y in (dfar / rmax * 1.02, dfar / rmax * 1.5, 10.0 * dfar / rmax):
            pts = _shell_points(rng, row, y, 512)
            q = abi.A | abi.B | abi.G
            ref, _ = ctx.fields(cs, pts, q, abi.FP64)
            got, rep = ctx.fields(cs, pts, q, abi.FP32C)
            assert rep.near_pairs == 0
            for k in ("A", "B", "G"):
                scale = np.max(np.linalg.norm(ref[k].reshape(len(pts), -1), axis=1))
                err = np.max(np.linalg.norm((got[k] - ref[k]).reshape(len(pts), -1), axis=1)) / scale
                worst = max(worst, err)
                assert err <= (1e-11 if n_ang >= 32 else 1e-9), (typ, n_ang, n_thk, y, k, err)
    print(f"\n{typ}: series against the FP64 rule, worst {worst:.3g} of the coil's field")


def test_low_angular_orders_keep_the_rule(ctx):
    rng = np.random.default_rng(8)
    p, row = _coil("RECTANGULAR", 0.1, 0.05, 0.08, 4, 6)
    assert H.far_radius(row)[1] == 0.0
    pts = _shell_points(rng, row, 20.0, 256)
    cs = abi.CoilSet.from_rows([row])
    q = abi.A | abi.B | abi.G
    a, _ = ctx.fields(cs, pts, q, abi.FP32C)
    d, _ = ctx.fields(cs, pts, q, abi.FP32C, abi.DIRECT)
    for k in ("A", "B", "G"):
        assert np.array_equal(a[k], d[k]), k


def test_far_field_results_do_not_depend_on_the_visit_order(ctx):
    params, pts = W.group_workload(40, 1 << 15, seed=12)   # above the size that turns the visit order on
    rows, _ = H.gpu_args_rows(params)
    cs = abi.CoilSet.from_rows(rows)
    q = abi.B | abi.G
    first, _ = ctx.fields(cs, pts, q, abi.FP32C)
    again, _ = ctx.fields(cs, pts, q, abi.FP32C)
    perm = np.random.default_rng(3).permutation(len(pts))
    shuffled, _ = ctx.fields(cs, pts[perm], q, abi.FP32C)
    for k in ("B", "G"):
        assert np.array_equal(first[k], again[k]), k
        assert np.array_equal(first[k][perm], shuffled[k]), k


def test_mixed_group_against_fp64(ctx):
    """128 mixed coils at 65,536 points, away from the windings, per point against the FP64 rule: the FP64 series
    (CCOIL_FAR_FIELD) within 1e-9; FP32C with its series at p99 <= 1e-6 and max <= 1e-5 (the float rule's own
    rounding, several float ulps per pair, on the pairs inside twice r_max), no worse than without the series."""
    params, pts = W.group_workload(128, 1 << 16, seed=5)
    rows, _ = H.gpu_args_rows(params)
    cs = abi.CoilSet.from_rows(rows)
    q = abi.B | abi.G
    ref, _ = ctx.fields(cs, pts, q, abi.FP64)
    far = ~H.near_mask(params, pts)
    res = {}
    for name, prec, flags in (("FP64 series", abi.FP64, abi.FAR_FIELD), ("FP32C series", abi.FP32C, 0),
                              ("FP32C direct", abi.FP32C, abi.DIRECT)):
        out, rep = ctx.fields(cs, pts, q, prec, flags)
        eb = H.rel_err_vec(out["B"], ref["B"])[far]
        eg = H.rel_err_vec(out["G"], ref["G"])[far]
        res[name] = (np.nanmax(eb), np.nanpercentile(eb, 99), np.nanmax(eg), np.nanpercentile(eg, 99))
        print(f"\n{name}: B far max {res[name][0]:.3g} p99 {res[name][1]:.3g} | G far max {res[name][2]:.3g} "
              f"p99 {res[name][3]:.3g} | near pairs {rep.near_pairs} | field kernel {rep.main_kernel_ms:.2f} ms, "
              f"all {rep.kernel_ms:.2f} ms")
    assert res["FP64 series"][0] <= 1e-9 and res["FP64 series"][2] <= 1e-9
    s, d = res["FP32C series"], res["FP32C direct"]
    assert s[1] <= 1e-6 and s[3] <= 1e-6 and s[0] <= 1e-5 and s[2] <= 1e-5
    assert s[0] <= 1.5 * d[0] and s[1] <= 1.5 * d[1]


@pytest.mark.parametrize("typ,R,a,b", [("RECTANGULAR", 0.1, 0.05, 0.08), ("FILAMENT", 0.1, 0.0, 0.0)])
def test_fp64_series_matches_the_rule(ctx, typ, R, a, b):
    """CCOIL_FP64 with CCOIL_FAR_FIELD: beyond the coil's (FP64) far-field distance the series, within 1e-12 of the
    coil's field of the rule; without the flag bitwise the plain rule."""
    rng = np.random.default_rng(9)
    p, row = _coil(typ, R, a, b, 48, 7)
    rmax, _ = H.far_radius(row)
    cs = abi.CoilSet.from_rows([row])
    pts = np.concatenate([_shell_points(rng, row, y, 256) for y in (1.5, 2.5, 6.0, 30.0)])
    q = abi.A | abi.B | abi.G
    ref, _ = ctx.fields(cs, pts, q, abi.FP64)
    got, _ = ctx.fields(cs, pts, q, abi.FP64, abi.FAR_FIELD)
    for k in ("A", "B", "G"):
        for part in range(4):
            sl = slice(256 * part, 256 * (part + 1))
            scale = np.max(np.linalg.norm(ref[k][sl].reshape(256, -1), axis=1))
            err = np.max(np.linalg.norm((got[k][sl] - ref[k][sl]).reshape(256, -1), axis=1)) / scale
            assert err <= 1e-12, (typ, k, part, err)
    inner = slice(0, 256)   # y = 1.5: inside twice r_max, the rule itself
    assert np.array_equal(got["B"][inner], ref["B"][inner])
