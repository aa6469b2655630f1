"""Shared test helpers: coil construction for product and oracle from one parameter dict, error metrics."""
import numpy as np

import oracle as O
import workloads as W

ORACLE_TYPE = {"RECTANGULAR": O.RECTANGULAR, "THIN": O.THIN, "FLAT": O.FLAT, "FILAMENT": O.FILAMENT}


def oracle_coil(p):
    return O.make_coil(p["R"], p["a"], p["b"], int(p["N"]), p.get("I", 1.0), p.get("freq", 50.0),
                       pos=tuple(float(x) for x in p["pos"]), alpha=p["alpha"], beta=p["beta"])


def gpu_args_rows(params, pf=5.0):
    """Product planning (c_coil) -> C ABI rows, and the same args as oracle int32[6]."""
    import c_coil
    rows, args = [], []
    for p in params:
        c = W.make_coil(p, pf)
        a = c_coil.coil.PrecisionArguments.get_coil_precision_arguments_GPU(c, c_coil.coil.PrecisionFactor(pf))
        rows.append(W.coil_row(c, a))
        # The GPU kernels use one block per layer; the oracle integrates the same rule.
        args.append(np.array([1, 1, 1, a.angular_increment_count, a.thickness_increment_count,
                              a.length_increment_count], dtype=np.int32))
    return rows, args


def rel_err_vec(x, ref, floor=0.0):
    """Per-point vector-norm relative error |x - ref| / |ref| (rows). Points with |ref| <= floor get NaN (excluded by
    nanmax); a row whose output x is not finite gets inf wherever it lies, so a NaN written by a kernel fails."""
    x = np.asarray(x).reshape(len(ref), -1)
    ref = np.asarray(ref).reshape(len(ref), -1)
    den = np.linalg.norm(ref, axis=1)
    num = np.linalg.norm(x - ref, axis=1)
    out = np.full(len(ref), np.nan)
    ok = den > floor
    out[ok] = num[ok] / den[ok]
    out[~np.isfinite(x).all(axis=1)] = np.inf
    return out


def far_threshold(n_ang, q, tol=1e-11, ylo=2.0):
    """Distance in units of r_max beyond which the far-field series may replace the rule on accuracy (0: never),
    restated from the bound the library documents: the least y in [ylo, 100] with 1000 y^2 rho_B^(-2 n_ang) <= tol,
    where rho_B is the Bernstein-ellipse parameter of x_s = -1 + 2i eta / pi and cosh eta = 1 + (y - 1)^2 q / (2 y),
    q = r_max / (R + a)."""
    def bound(y):
        d = y - 1.0
        eta = np.arccosh(1.0 + d * d * q / (2.0 * y))
        xs = complex(-1.0, 2.0 * eta / np.pi)
        sq = np.sqrt(xs * xs - 1.0)
        return 1000.0 * y * y * max(abs(xs + sq), abs(xs - sq)) ** (-2.0 * n_ang)
    if n_ang < 5:
        return 0.0
    if bound(ylo) <= tol:
        return ylo
    if bound(100.0) > tol:
        return 0.0
    lo, hi = ylo, 100.0
    for _ in range(40):
        mid = 0.5 * (lo + hi)
        if bound(mid) <= tol:
            hi = mid
        else:
            lo = mid
    return hi


FAR_MIN_Y = 1.17
LOG_TOL = {"fp64": float(np.float32(35.2319)), "fp32c": float(np.float32(30.6267))}   # ln(2 / tol) of the terms


def far_cost_floor(n_ang, n_thk, fast, log_tol):
    """The library's cost floor of the far-field distance (units of r_max), restated: the series of
    L(y) = 2 log_tol / ln(y^2) + 2 terms at 15 operations per term plus 200 against the rule's n_ang n_thk nodes at
    55 (fast method) or 12 (slow) operations, the nearest y in [1.17, 2] where the series is not dearer (2 when the
    rule costs 230 or less)."""
    rule = float(n_ang) * n_thk * (55.0 if fast else 12.0)
    if rule <= 230.0:
        return 2.0
    return min(2.0, max(FAR_MIN_Y, float(np.exp(15.0 * log_tol / (rule - 230.0)))))


def far_distance(row, fp64=False, one_lane=True):
    """(r_max, far-field distance) of a coil row as the library sets them (distance 0: the coil keeps its rule):
    the accuracy bound at the precision's tolerance from the cost floor on for launches with one lane per point
    (pairs, and field calls of 16,384 points or more in the series modes), from 2 otherwise."""
    R, a, b = row["R"], row["a"], row["b"]
    rmax = float(np.hypot(R + a, 0.5 * b if row["use_fast"] else 0.0))
    ylo = (far_cost_floor(row["n_ang"], row["n_thk"], bool(row["use_fast"]), LOG_TOL["fp64" if fp64 else "fp32c"])
           if one_lane else 2.0)
    y = far_threshold(row["n_ang"], rmax / (R + a), 1e-14 if fp64 else 1e-11, ylo)
    return rmax, y * rmax


def far_radius(row, n_points=1):
    """(r_max, far-field distance) of a coil row in a CCOIL_FP32C field call of n_points points (distance 0: the coil
    keeps its rule)."""
    return far_distance(row, fp64=False, one_lane=n_points >= 16384)


def fp64_pair_count(params, pts, width_floor=1e-3, rows=None):
    """(coil, point) pairs that CCOIL_FP32C evaluates in FP64 when the gradient is requested: the near-winding shell
    of near_mask, and the cone and cylinder around each coil's axis, r < 0.1 sqrt(r^2 + z^2 + R^2); with rows (the
    coils' GPU rows) less the pairs beyond the coil's far-field distance, which take the series instead."""
    n = 0
    for k, p in enumerate(params):
        c = oracle_coil(p)
        T = np.array(c.Tinv).reshape(3, 3)
        l = (pts - np.array(p["pos"])) @ T.T
        r = np.hypot(l[:, 0], l[:, 1])
        z = l[:, 2]
        dr = np.maximum(np.maximum(c.R - r, r - (c.R + c.a)), 0.0)
        dz = np.maximum(np.abs(z) - 0.5 * c.b, 0.0)
        w = max(c.a, c.b, width_floor * c.R)
        sel = (dr * dr + dz * dz < w * w) | (r * r < 0.01 * (r * r + z * z + c.R * c.R))
        if rows is not None:
            _, dfar = far_radius(rows[k], len(pts))
            if dfar > 0:
                sel &= ~(r * r + z * z > dfar * dfar)
        n += int(np.sum(sel))
    return n


def inside_mask(params, pts):
    """Points inside any coil's winding volume (R <= r <= R + a, |z| <= b/2 in the coil frame), where the reference
    warns that fields are not meaningful (Coil.h:150) and the quadrature's node values are near-singular."""
    inside = np.zeros(len(pts), dtype=bool)
    for p in params:
        c = oracle_coil(p)
        T = np.array(c.Tinv).reshape(3, 3)
        l = (pts - np.array(p["pos"])) @ T.T
        r = np.hypot(l[:, 0], l[:, 1])
        inside |= (r >= c.R) & (r <= c.R + c.a) & (np.abs(l[:, 2]) <= 0.5 * c.b)
    return inside


def near_mask(params, pts, width_floor=1e-3):
    """Points within one cross-section width (max(a, b), floored at 1e-3 R) of any coil's winding."""
    near = np.zeros(len(pts), dtype=bool)
    for p in params:
        c = oracle_coil(p)
        T = np.array(c.Tinv).reshape(3, 3)
        l = (pts - np.array(p["pos"])) @ T.T
        r = np.hypot(l[:, 0], l[:, 1])
        z = l[:, 2]
        dr = np.maximum(np.maximum(c.R - r, r - (c.R + c.a)), 0.0)
        dz = np.maximum(np.abs(z) - 0.5 * c.b, 0.0)
        w = max(c.a, c.b, width_floor * c.R)
        near |= dr * dr + dz * dz < w * w
    return near
