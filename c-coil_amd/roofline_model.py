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
