"""ccoil_far_field_radius (host only) against the restatement of the far-field distance in helpers.py (the accuracy
bound searched from the cost floor on), over angular orders 1-100, thin and thick windings, fast and slow rules, both
precisions that use the series; the bound and the cost floor pinned on cases fixed by hand."""
import numpy as np
import pytest

import ccoil_abi as abi
import helpers as H
import workloads as W


def _rows(rng, n):
    rows = []
    for k in range(n):
        typ = ("RECTANGULAR", "THIN", "FLAT", "FILAMENT")[k % 4]
        p = dict(type=typ, R=rng.uniform(0.01, 0.5), a=rng.uniform(0.005, 0.3) if typ in ("RECTANGULAR", "FLAT") else 0.0,
                 b=rng.uniform(0.005, 0.3) if typ in ("RECTANGULAR", "THIN") else 0.0, N=100, I=1.0,
                 pos=rng.uniform(-1, 1, 3), alpha=rng.uniform(0, 3), beta=rng.uniform(0, 6))
        row = W.coil_row(W.make_coil(p))
        row.update(n_ang=int(1 + k % 100), n_thk=int(1 + k % 7), blocks=(1, 1, 1))
        rows.append(row)
    return rows


@pytest.mark.parametrize("prec", [abi.FP32C, abi.FP64])
def test_library_threshold_matches_the_bound(prec):
    """The distance is the accuracy bound's, searched from the cost floor on (a bisection over [floor, 100] whose
    last step is 9e-11 of the range: equal to that)."""
    rows = _rows(np.random.default_rng(1), 400)
    rmax, dist = abi.far_field_radius(abi.CoilSet.from_rows(rows), prec)
    for row, rm, d in zip(rows, rmax, dist):
        R, a, b = row["R"], row["a"], row["b"]
        assert rm == pytest.approx(np.hypot(R + a, 0.5 * b if row["use_fast"] else 0.0), rel=1e-15)
        _, dh = H.far_distance(row, fp64=prec == abi.FP64)
        assert d == pytest.approx(dh, rel=1e-9, abs=0.0), (row["n_ang"], row["n_thk"], dh / rm, d / rm)


def test_cost_floor_pinned_by_hand():
    # FP64 terms: log_tol = 35.2319 (float), L(y) = 2 log_tol / ln(y^2) + 2 at 15 per term + 200.
    # RECTANGULAR 80 x 20 nodes at 55: rule 88,000, y = exp(15 x 35.2319 / 87,770) = 1.0060: clamped to 1.17.
    assert H.far_cost_floor(80, 20, True, H.LOG_TOL["fp64"]) == 1.17
    # FILAMENT 80 x 1 at 12: rule 960, y = exp(528.48 / 730) = exp(0.72394) = 2.0626: clamped to 2.
    assert H.far_cost_floor(80, 1, False, H.LOG_TOL["fp64"]) == 2.0
    # slow 40 x 3: rule 1,440, y = exp(528.48 / 1,210) = exp(0.436759) = 1.54768
    assert H.far_cost_floor(40, 3, False, H.LOG_TOL["fp64"]) == pytest.approx(1.54768, abs=2e-5)
    # slow 12 x 1: rule 144 <= 230: 2
    assert H.far_cost_floor(12, 1, False, H.LOG_TOL["fp64"]) == 2.0
    # at the floor the FP64 series needs int(2 x 35.2319 / ln 1.3689) + 2 = 226 terms, within the 250 it carries
    assert int(2 * H.LOG_TOL["fp64"] / np.log(1.17 ** 2)) + 2 == 226


def test_other_precisions_never_use_the_series():
    rows = _rows(np.random.default_rng(2), 20)
    for prec in (abi.FP32, abi.FP64_REFERENCE, abi.FP64_IEEE):
        _, dist = abi.far_field_radius(abi.CoilSet.from_rows(rows), prec)
        assert np.all(dist == 0.0)


def test_bound_pinned_by_hand():
    # y = 2, q = 1: cosh(eta) = 1 + 1 / 4 = 1.25, eta = ln 2 (cosh ln 2 = (2 + 1/2) / 2); x_s = -1 + 2i ln2 / pi =
    # -1 + 0.441271i; rho_B = |x_s - sqrt(x_s^2 - 1)| = 1.98580 (by hand: x_s^2 - 1 = -0.194720 - 0.882542i, whose
    # root of positive real part is 0.595419 - 0.741110i, so x_s - root = -1.595419 + 1.182381i). With n = 40:
    # 1000 * 4 * rho^-80 = 5.8e-21 <= 1e-14, so y = 2 is the threshold; with n = 12: 4000 * rho^-24 = 2.8e-4 > 1e-11,
    # so it lies beyond 2.
    assert H.far_threshold(40, 1.0, 1e-14) == 2.0
    assert H.far_threshold(12, 1.0, 1e-11) > 2.0
    xs = complex(-1.0, 2.0 * np.log(2.0) / np.pi)
    sq = np.sqrt(xs * xs - 1)
    assert max(abs(xs + sq), abs(xs - sq)) == pytest.approx(1.98580, abs=1e-5)
    # fewer than 5 angular increments never use the series; a bound still above tol at y = 100 neither
    assert H.far_threshold(4, 1.0, 1.0) == 0.0
    assert H.far_threshold(8, 1.0, 1e-14) == 0.0
    # the threshold grows as the tolerance tightens and falls as the angular order grows
    assert H.far_threshold(16, 1.0, 1e-14) > H.far_threshold(16, 1.0, 1e-11) >= 2.0
    assert H.far_threshold(14, 1.0, 1e-11) > H.far_threshold(18, 1.0, 1e-11)


def test_restatements_follow_the_device_constants():
    """The host-side restatements (helpers.FAR_MIN_Y, the roofline model's term cap) carry the device's constants
    (kFarMinY in context.hpp, kMpTerms in multipole.hpp), read from the sources."""
    import os
    import re
    import roofline_model as RM
    src = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "c-coil_amd", "csrc")
    ctx = open(os.path.join(src, "context.hpp")).read()
    mp = open(os.path.join(src, "multipole.hpp")).read()
    assert float(re.search(r"kFarMinY = ([0-9.]+);", ctx).group(1)) == H.FAR_MIN_Y
    assert int(re.search(r"kMpTerms = (\d+);", mp).group(1)) == RM.SERIES_MAX_TERMS
