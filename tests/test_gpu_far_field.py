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
  - Calls of 16,384 points or more start the series at the coil's cost floor (down to 1.17 r_max, up to 250 terms):
    FP32C within 1e-11 and FP64 within 1e-13 of the coil's field of the FP64 rule just beyond it.
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


@pytest.mark.parametrize("typ,R,a,b", [("RECTANGULAR", 0.1, 0.05, 0.08), ("RECTANGULAR", 0.04, 0.2, 0.15),
                                       ("THIN", 0.1, 0.0, 0.3), ("FLAT", 0.02, 0.2, 0.0)])
def test_series_from_the_cost_floor(ctx, typ, R, a, b):
    """Calls of 16,384 points or more (one lane per point) take the series from the coil's cost floor (down to
    1.17 r_max) where its accuracy bound allows. Just beyond that distance and further out, for angular orders 80 to
    16: CCOIL_FP32C within 1e-11 of the coil's field of the FP64 rule, CCOIL_FP64 with CCOIL_FAR_FIELD within 1e-13."""
    rng = np.random.default_rng(23)
    worst = {}
    for n_ang, n_thk in ((80, 20), (48, 7), (32, 12), (16, 30)):
        p, row = _coil(typ, R, a, b, n_ang, n_thk)
        cs = abi.CoilSet.from_rows([row])
        for prec, flags, fp64, tol in ((abi.FP32C, 0, False, 1e-11), (abi.FP64, abi.FAR_FIELD, True, 1e-13)):
            rmax, dfar = H.far_distance(row, fp64=fp64, one_lane=True)
            assert dfar > 0, (typ, n_ang)
            ys = (dfar / rmax * 1.0001, dfar / rmax * 1.1, dfar / rmax * 1.4)
            pts = np.concatenate([_shell_points(rng, row, y, 6000) for y in ys])
            q = abi.A | abi.B | abi.G
            ref, _ = ctx.fields(cs, pts, q, abi.FP64)
            got, rep = ctx.fields(cs, pts, q, prec, flags)
            assert rep.near_pairs == 0
            for k in ("A", "B", "G"):
                for part in range(3):
                    sl = slice(6000 * part, 6000 * (part + 1))
                    scale = np.max(np.linalg.norm(ref[k][sl].reshape(6000, -1), axis=1))
                    err = np.max(np.linalg.norm((got[k][sl] - ref[k][sl]).reshape(6000, -1), axis=1)) / scale
                    worst[(prec, n_ang)] = max(worst.get((prec, n_ang), 0.0), err)
                    assert err <= tol, (typ, n_ang, n_thk, prec, round(dfar / rmax, 3), k, part, err)
    print(f"\n{typ} R={R} a={a} b={b}: worst " + ", ".join(f"{'FP64' if k[0] == abi.FP64 else 'FP32C'}/{k[1]}: {v:.2g}"
                                                      for k, v in worst.items()))


def test_module_far_field_switch():
    """c_coil.coil.Coil.set_gpu_far_field: per thread, -1 the precision's default, 0 every pair by the rule, 1 the
    series beyond each coil's distance. FP64 on: within 1e-12 of the rule at far points; FP32C off: bitwise the
    CCOIL_DIRECT call of the C ABI."""
    import c_coil
    C, V = c_coil.coil, c_coil.tensor.Vector3
    coil = C.Coil(0.1, 0.05, 0.08, 100, 2.0, 60.0, C.PrecisionFactor(5.0), 8, V(0.1, -0.05, 0.02), 0.4, 1.1)
    rng = np.random.default_rng(91)
    d = rng.standard_normal((500, 3))
    pts = np.array([0.1, -0.05, 0.02]) + 1.5 * d / np.linalg.norm(d, axis=1, keepdims=True)   # about 9 r_max away
    try:
        C.Coil.set_gpu_precision(C.GPU_FP64)
        assert C.Coil.get_gpu_far_field() == -1
        rule = coil.compute_all_B_field_vectors(pts, C.GPU)
        C.Coil.set_gpu_far_field(1)
        series = coil.compute_all_B_field_vectors(pts, C.GPU)
        err = np.linalg.norm(series - rule, axis=1) / np.linalg.norm(rule, axis=1)
        assert err.max() <= 1e-12 and not np.array_equal(series, rule)
        C.Coil.set_gpu_precision(C.GPU_FP32C)
        C.Coil.set_gpu_far_field(0)
        direct = coil.compute_all_B_field_vectors(pts, C.GPU)
        C.Coil.set_gpu_far_field(-1)
        default = coil.compute_all_B_field_vectors(pts, C.GPU)
        e32 = np.linalg.norm(default - rule, axis=1) / np.linalg.norm(rule, axis=1)
        assert e32.max() <= 1e-12 and not np.array_equal(direct, default)
        with pytest.raises(ValueError):
            C.Coil.set_gpu_far_field(2)
    finally:
        C.Coil.set_gpu_far_field(-1)
        C.Coil.set_gpu_precision(C.GPU_FP32C)


def test_pairs_series_against_the_rule(ctx):
    """Pairs: secondary nodes beyond the primary's far-field distance take the primary's series (CCOIL_FP64 with
    CCOIL_FAR_FIELD, CCOIL_FP32C by default). On 4,096 random pairs of the mixed distributions, M, F, T and dM of the
    FP64 series against the FP64 rule: M within 1e-10 of |M| + |F/(I1 I2)| d (d the centre distance: M's size over
    the pose's neighbourhood), F, T, dM within 1e-10 of their norms; FP32C with the series against FP64 no worse than
    without it; permuting the batch changes nothing, bitwise."""
    prow, srow = W.pair_workload(4096, seed=3)
    P, S = abi.CoilSet.from_rows(prow), abi.CoilSet.from_rows(srow)
    n = len(prow)
    what = abi.PAIR_M | abi.PAIR_FT | abi.PAIR_DM
    ref, r64 = ctx.pairs(P, S, n, what, abi.FP64)
    ser, s64 = ctx.pairs(P, S, n, what, abi.FP64, abi.FAR_FIELD)
    cur = np.array([p["current"] * s["current"] for p, s in zip(prow, srow)])
    dist = np.linalg.norm(np.array([s["pos"] for s in srow]) - np.array([p["pos"] for p in prow]), axis=1)
    mscale = np.abs(ref["M"]) + np.linalg.norm(ref["F"], axis=1) / np.abs(cur) * dist
    em = np.abs(ser["M"] - ref["M"]) / mscale
    ef = np.linalg.norm(ser["F"] - ref["F"], axis=1) / np.linalg.norm(ref["F"], axis=1)
    et = np.linalg.norm(ser["T"] - ref["T"], axis=1) / np.linalg.norm(ref["T"], axis=1)
    ed = np.linalg.norm(ser["dM"] - ref["dM"], axis=1) / np.linalg.norm(ref["dM"], axis=1)
    changed = np.mean(ser["M"] != ref["M"])
    print(f"\npairs, FP64 series vs rule ({changed:.0%} of the pairs changed): M {em.max():.3g}, F {ef.max():.3g}, "
          f"T {et.max():.3g}, dM {ed.max():.3g}; kernel {s64.main_kernel_ms:.1f} ms against {r64.main_kernel_ms:.1f}")
    assert changed > 0.1
    assert em.max() <= 1e-10 and ef.max() <= 1e-10 and et.max() <= 1e-10 and ed.max() <= 1e-10
    f32, _ = ctx.pairs(P, S, n, what, abi.FP32C)
    d32, _ = ctx.pairs(P, S, n, what, abi.FP32C, abi.DIRECT)
    e32 = np.abs(f32["M"] - ref["M"]) / mscale
    ed32 = np.abs(d32["M"] - ref["M"]) / mscale
    f32f = np.linalg.norm(f32["F"] - ref["F"], axis=1) / np.linalg.norm(ref["F"], axis=1)
    print(f"FP32C M vs FP64 (relative to M's size): with the series max {e32.max():.3g} p99 "
          f"{np.percentile(e32, 99):.3g}; the rule only max {ed32.max():.3g} p99 {np.percentile(ed32, 99):.3g}; "
          f"F max {f32f.max():.3g}")
    assert e32.max() <= ed32.max() and np.percentile(e32, 99) <= np.percentile(ed32, 99)
    perm = np.random.default_rng(4).permutation(n)
    pp, _ = ctx.pairs(P, S, n, what, abi.FP64, abi.FAR_FIELD, prim_index=perm, sec_index=perm)
    for k in ("M", "F", "T", "dM"):
        assert np.array_equal(pp[k], ser[k][perm]), k
