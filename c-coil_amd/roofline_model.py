"""Algorithmic work per quadrature node, frozen before the first measurement (used by bench.py's roofline).

The unit of the field path is one coil·point evaluation at the coil's GPU PrecisionArguments: n_thk * n_ang
nodes (the device kernels, like the reference's, have no length loop). The work per node is counted once from
the reference's float integrands after their obvious common subexpressions (K = J^-1/2, K/J = K^3, one shared
reciprocal), FMA counted as 2 flops, per-point work (the frame transform, cos phi = x / r) excluded:

  fast method (THIN, RECTANGULAR)  B: FieldBGroup.cu:47-77   gradient: FieldGGroup.cu:257-305
  slow method (FLAT, FILAMENT)     B: FieldBGroup.cu:78-101  gradient: FieldGGroup.cu:306-339
  vector potential                 FieldAGroup.cu

  variant                  special ops (rsqrt/rcp/sqrt/log)   flops
  fast B                   3                                  19
  fused fast B + gradient  3                                  63
  slow B                   1                                  13
  fused slow B + gradient  1                                  40
  fast A                   5                                  14
  slow A                   1                                   7

The device kernels evaluate rearranged, cancellation-free integrands that issue more instructions per node than
this; the extra instructions (and any compensation) are not algorithmic work and are not counted.
Pairs: one node-pair of M costs one A evaluation; one node-pair of F/T costs one B evaluation.

The far-field series (a field call's (coil, point) pairs beyond the coil's far-field distance, multipole.hpp) is
counted per pair from its FP64 loop: per odd term 13 flops for B (the next power of r_max / D, two sums with their
coefficient products, the Legendre recurrence of P and P') and 18 with the gradient (two more sums); per even term
7 (the power and the recurrences: a coil symmetric about its mid-plane, as every coil here is, has no even moments,
so those terms have no sums), and per pair
20 (B) or 30 (B + gradient) flops and 2 special ops (1 / sqrt(D^2), the log of the term count) for the set-up and
the outputs. The term count per pair is the kernel's: int(2 ln(2 / tol) / ln(D^2 / r_max^2)) + 2 in [4, SERIES_MAX_TERMS]
(multipole.hpp's kMpTerms).
"""

FAST_B = (3, 19)
FAST_BG = (3, 63)
SLOW_B = (1, 13)
SLOW_BG = (1, 40)
FAST_A = (5, 14)
SLOW_A = (1, 7)

# MI355X published peaks (dense, no sparsity): FP32 vector 157.3 TFLOP/s (packed FMA; 78.6 TFLOP/s unpacked),
# FP64 vector 78.6 TFLOP/s. Transcendentals issue at a quarter of the FP32 lane rate:
# 256 CUs * 16 lanes/clk * 2.4 GHz = 9.83e12 ops/s.
PEAK_FP32_TFLOPS = 157.3
PEAK_FP64_TFLOPS = 78.6
PEAK_TRANS_OPS = 256 * 16 * 2.4e9
PEAK_HBM_GBS = 8000.0


def field_work(rows, n_points, with_gradient=True):
    """(flops, special ops) of one pass of B (+ gradient) over n_points for coil rows (dicts with use_fast,
    n_ang, n_thk)."""
    flops = trans = 0
    for r in rows:
        nodes = int(r["n_ang"]) * int(r["n_thk"])
        t, f = (FAST_BG if with_gradient else FAST_B) if r["use_fast"] else (SLOW_BG if with_gradient else SLOW_B)
        flops += nodes * f
        trans += nodes * t
    return float(flops) * n_points, float(trans) * n_points


SERIES_B = (2, 20, 13)    # special ops and flops per pair, flops per odd term
SERIES_BG = (2, 30, 18)
SERIES_EVEN_TERM = 7      # flops per even term (no sums: the even moments are zero)
SERIES_MAX_TERMS = 250    # kMpTerms
LOG_TOL_FP64 = 35.2319    # ln(2 / 1e-15)
LOG_TOL_FP32C = 30.6267   # ln(2 / 1e-13)


def series_field_work(rows, pts, dist, rmax, log_tol, with_gradient=True, near_pass=False):
    """Work of one field_kernel launch that takes the far-field series beyond each coil's distance dist (0: never)
    and the rule elsewhere; near_pass: the FP32C near-winding (and, with the gradient, axis-cone) pairs are left to
    the near pass. Returns dict(rule_flops, rule_special, series_flops, series_special, rule_pairs, series_pairs,
    near_pairs)."""
    import numpy as np
    from workloads import rotation
    out = dict(rule_flops=0.0, rule_special=0.0, series_flops=0.0, series_special=0.0, rule_pairs=0,
               series_pairs=0, near_pairs=0, near_flops=0.0)
    sp, sf, st = SERIES_BG if with_gradient else SERIES_B
    for r, d, rm in zip(rows, dist, rmax):
        x = pts - np.asarray(r["pos"])
        D2 = np.einsum("ij,ij->i", x, x)
        far = (D2 > d * d) if d > 0 else np.zeros(len(pts), dtype=bool)
        near = np.zeros(len(pts), dtype=bool)
        if near_pass:
            l = x @ rotation(r["alpha"], r["beta"])
            rr, z = np.hypot(l[:, 0], l[:, 1]), l[:, 2]
            dr = np.maximum(np.maximum(r["R"] - rr, rr - (r["R"] + r["a"])), 0.0)
            dz = np.maximum(np.abs(z) - 0.5 * r["b"], 0.0)
            w = max(r["a"], r["b"], 1e-3 * r["R"])
            near = dr * dr + dz * dz < w * w
            if with_gradient:
                near |= rr * rr < 0.01 * (rr * rr + z * z + r["R"] ** 2)
            near &= ~far
        rule = ~far & ~near
        nodes = int(r["n_ang"]) * int(r["n_thk"])
        t, f = (FAST_BG if with_gradient else FAST_B) if r["use_fast"] else (SLOW_BG if with_gradient else SLOW_B)
        nr = int(rule.sum())
        out["rule_flops"] += float(nr) * nodes * f
        out["rule_special"] += float(nr) * nodes * t
        if far.any():
            L = np.clip((2.0 * log_tol / np.log(D2[far] / (rm * rm))).astype(np.int64) + 2, 4, SERIES_MAX_TERMS)
            odd, even = (L + 1) // 2, L // 2
            out["series_flops"] += float(far.sum()) * sf + float(odd.sum()) * st + float(even.sum()) * SERIES_EVEN_TERM
            out["series_special"] += float(far.sum()) * sp
        out["rule_pairs"] += nr
        out["series_pairs"] += int(far.sum())
        out["near_pairs"] += int(near.sum())
        out["near_flops"] += float(near.sum()) * nodes * f
    return out


def pair_work(prim_rows, sec_rows, with_m=True, with_ft=True):
    """(flops, special ops, node-pairs) of one pass of ccoil_pairs over the pairs (row k with row k)."""
    flops = trans = nodes = 0
    for p, s in zip(prim_rows, sec_rows):
        n = int(p["n_ang"]) * int(p["n_thk"]) * int(s["n_ang"]) * int(s["n_thk"]) * int(s["n_len"])
        nodes += n
        if with_m:
            t, f = FAST_A if p["use_fast"] else SLOW_A
            flops += n * f
            trans += n * t
        if with_ft:
            t, f = FAST_B if p["use_fast"] else SLOW_B
            flops += n * f
            trans += n * t
    return float(flops), float(trans), float(nodes)
