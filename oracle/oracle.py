"""ctypes front end of the FP64 oracle (oracle/ccoil_oracle.c).

TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this module, and only as the correctness reference or as the timed CPU baseline. The product (the HIP
library and the c_coil module) never imports it.

Each wrapper names the reference routine it restates; the C file cites file:line.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

RECTANGULAR, THIN, FLAT, FILAMENT = 0, 1, 2, 3
AXIS_RATIO_CPU = 1e-12   # CalculateFieldsMethods.cxx:60
AXIS_RATIO_GPU = 1e-5    # FieldGGroup.cu:134 (GPU kernels' on-axis branch)
GPU_INCREMENTS = 80      # reference CMakeLists.txt:16 default


class OrcCoil(ctypes.Structure):
    _fields_ = [
        ("R", ctypes.c_double), ("a", ctypes.c_double), ("b", ctypes.c_double),
        ("N", ctypes.c_int32), ("I", ctypes.c_double), ("J", ctypes.c_double),
        ("freq", ctypes.c_double), ("type", ctypes.c_int32), ("use_fast", ctypes.c_int32),
        ("pos", ctypes.c_double * 3), ("alpha", ctypes.c_double), ("beta", ctypes.c_double),
        ("T", ctypes.c_double * 9), ("Tinv", ctypes.c_double * 9),
    ]


def build():
    """Compile both oracle builds (make -C oracle)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = {}


def lib(fast=False):
    key = "fast" if fast else "ieee"
    if key not in _lib:
        name = "liboracle_fast.so" if fast else "liboracle.so"
        path = os.path.join(HERE, name)
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int32)
        cp = ctypes.POINTER(OrcCoil)
        L.orc_gauss_legendre.argtypes = [ctypes.c_int, dp, dp]
        L.orc_precision_factor.argtypes = [ctypes.c_double]
        L.orc_precision_factor.restype = ctypes.c_double
        L.orc_coil_make.argtypes = [cp, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int32,
                                    ctypes.c_double, ctypes.c_double, dp, ctypes.c_double, ctypes.c_double]
        L.orc_coil_precision_args.argtypes = [cp, ctypes.c_double, ctypes.c_int, ctypes.c_int, ip]
        L.orc_secondary_precision_args_gpu.argtypes = [cp, ctypes.c_double, ctypes.c_int, ip]
        L.orc_pair_precision_args.argtypes = [cp, cp, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                              ctypes.c_int, ip]
        L.orc_coil_fields.argtypes = [cp, ip, ctypes.c_int64, dp, dp, dp, dp, ctypes.c_double]
        L.orc_coil_fields_mt.argtypes = [cp, ip, ctypes.c_int64, dp, dp, dp, dp, ctypes.c_double, ctypes.c_int]
        L.orc_coil_fields_ext.argtypes = [cp, ip, ctypes.c_int64, dp, dp, dp, dp, ctypes.c_double, ctypes.c_int]
        L.orc_chunk_ends.argtypes = [ctypes.c_int64, ctypes.c_int, ctypes.POINTER(ctypes.c_int64)]
        L.orc_mutual_inductance_general.argtypes = [cp, cp, ip]
        L.orc_mutual_inductance_general.restype = ctypes.c_double
        L.orc_force_torque_general.argtypes = [cp, cp, ip, dp, dp]
        L.orc_mutual_inductance_gpu_rule.argtypes = [cp, cp, ip, ip]
        L.orc_mutual_inductance_gpu_rule.restype = ctypes.c_double
        L.orc_force_torque_gpu_rule.argtypes = [cp, cp, ip, ip, ctypes.c_double, dp, dp]
        L.orc_zaxis_rule.argtypes = [cp, cp, ip, ctypes.c_int, ctypes.c_int]
        L.orc_zaxis_rule.restype = ctypes.c_double
        _lib[key] = L
    return _lib[key]


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) if a is not None else None


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))


def gauss_legendre(n):
    x = np.zeros(n)
    w = np.zeros(n)
    lib().orc_gauss_legendre(n, _dp(x), _dp(w))
    return x, w


def precision_factor(pf):
    return lib().orc_precision_factor(pf)


def make_coil(R, a, b, N, I=1.0, freq=50.0, pos=(0.0, 0.0, 0.0), alpha=0.0, beta=0.0):
    """Coil constructor semantics (Coil.cxx:22-44)."""
    c = OrcCoil()
    p = np.ascontiguousarray(pos, dtype=np.float64)
    lib().orc_coil_make(ctypes.byref(c), R, a, b, int(N), I, freq, _dp(p), alpha, beta)
    return c


def coil_precision_args(coil, pf=5.0, gpu=False, gpu_increments=GPU_INCREMENTS):
    out = np.zeros(6, dtype=np.int32)
    lib().orc_coil_precision_args(ctypes.byref(coil), pf, int(gpu), gpu_increments, _ip(out))
    return out


def secondary_precision_args_gpu(coil, pf=5.0, gpu_increments=GPU_INCREMENTS):
    out = np.zeros(6, dtype=np.int32)
    lib().orc_secondary_precision_args_gpu(ctypes.byref(coil), pf, gpu_increments, _ip(out))
    return out


def pair_precision_args(prim, sec, pf=5.0, gpu=False, z_axis=False, pure_gpu=False,
                        gpu_increments=GPU_INCREMENTS):
    out = np.zeros(12, dtype=np.int32)
    lib().orc_pair_precision_args(ctypes.byref(prim), ctypes.byref(sec), pf, int(gpu), int(z_axis),
                                  int(pure_gpu), gpu_increments, _ip(out))
    return out


def coil_fields(coil, args, xyz, A=True, B=True, G=True, axis_ratio=AXIS_RATIO_CPU, threads=0, fast=False,
                extended=False):
    """Per-coil A/B/G (G = T*G_local) at points xyz (n,3). threads>0 uses the CPU_MT chunking;
    extended=True evaluates the same rule in long double (the accuracy truth)."""
    xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
    n = xyz.shape[0]
    a = np.zeros((n, 3)) if A else None
    b = np.zeros((n, 3)) if B else None
    g = np.zeros((n, 3, 3)) if G else None
    args = np.ascontiguousarray(args, dtype=np.int32)
    L = lib(fast)
    if extended:
        L.orc_coil_fields_ext(ctypes.byref(coil), _ip(args), n, _dp(xyz), _dp(a), _dp(b), _dp(g), axis_ratio,
                              max(threads, 1))
    elif threads > 0:
        L.orc_coil_fields_mt(ctypes.byref(coil), _ip(args), n, _dp(xyz), _dp(a), _dp(b), _dp(g), axis_ratio,
                             threads)
    else:
        L.orc_coil_fields(ctypes.byref(coil), _ip(args), n, _dp(xyz), _dp(a), _dp(b), _dp(g), axis_ratio)
    return a, b, g


def group_fields(coils, args_list, xyz, A=True, B=True, G=True, axis_ratio=AXIS_RATIO_CPU, threads=0, fast=False,
                 extended=False):
    """Sum over coils (CoilGroup ComputeFields.cxx:45-163 single-threaded branch semantics)."""
    xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
    n = xyz.shape[0]
    at = np.zeros((n, 3)) if A else None
    bt = np.zeros((n, 3)) if B else None
    gt = np.zeros((n, 3, 3)) if G else None
    for c, args in zip(coils, args_list):
        a, b, g = coil_fields(c, args, xyz, A, B, G, axis_ratio, threads, fast, extended)
        if A:
            at += a
        if B:
            bt += b
        if G:
            gt += g
    return at, bt, gt


def chunk_ends(op_count, threads):
    ends = (ctypes.c_int64 * (threads + 1))()
    lib().orc_chunk_ends(op_count, threads, ends)
    return list(ends)


def mutual_inductance_general(prim, sec, args12):
    args12 = np.ascontiguousarray(args12, dtype=np.int32)
    return lib().orc_mutual_inductance_general(ctypes.byref(prim), ctypes.byref(sec), _ip(args12))


def force_torque_general(prim, sec, args12):
    args12 = np.ascontiguousarray(args12, dtype=np.int32)
    F = np.zeros(3)
    T = np.zeros(3)
    lib().orc_force_torque_general(ctypes.byref(prim), ctypes.byref(sec), _ip(args12), _dp(F), _dp(T))
    return F, T


def mutual_inductance_gpu_rule(prim, sec, prim_inc, sec_inc):
    pi = np.ascontiguousarray(prim_inc, dtype=np.int32)
    si = np.ascontiguousarray(sec_inc, dtype=np.int32)
    return lib().orc_mutual_inductance_gpu_rule(ctypes.byref(prim), ctypes.byref(sec), _ip(pi), _ip(si))


def force_torque_gpu_rule(prim, sec, prim_inc, sec_inc, correction):
    pi = np.ascontiguousarray(prim_inc, dtype=np.int32)
    si = np.ascontiguousarray(sec_inc, dtype=np.int32)
    F = np.zeros(3)
    T = np.zeros(3)
    lib().orc_force_torque_gpu_rule(ctypes.byref(prim), ctypes.byref(sec), _ip(pi), _ip(si), correction,
                                    _dp(F), _dp(T))
    return F, T


def _zaxis(prim, sec, args12, mode, extended):
    args12 = np.ascontiguousarray(args12, dtype=np.int32)
    return lib().orc_zaxis_rule(ctypes.byref(prim), ctypes.byref(sec), _ip(args12), mode, int(extended))


def mutual_inductance_zaxis_fast(prim, sec, args12, extended=False):
    """Coil::calculateMutualInductanceZAxisFast (CalculateMInductanceZAxis.cxx:84-221)."""
    return _zaxis(prim, sec, args12, 0, extended)


def force_zaxis_fast(prim, sec, args12, extended=False):
    """Coil::calculateAmpereForceZAxisFast (CalculateForceZAxis.cxx:82-219): axial force on the secondary."""
    return _zaxis(prim, sec, args12, 1, extended)


def self_inductance(coil, args12, extended=False):
    """Coil::calculateSelfInductance (CalculateMInductanceZAxis.cxx:224-317)."""
    return _zaxis(coil, coil, args12, 2, extended)


def _zaxis_slow_grid(sec, args12):
    """Secondary r-z grid and weights of CalculateMInductanceZAxis.cxx:19-72 / CalculateForceZAxis.cxx:16-69."""
    lb, li = int(args12[6 + 2]), int(args12[6 + 5])
    tb, ti = int(args12[6 + 1]), int(args12[6 + 4])
    lbs, tbs = sec.b / lb, sec.a / tb
    xl, wl = gauss_legendre(li)
    xt, wt = gauss_legendre(ti)
    pts, w = [], []
    for zb in range(lb):
        for rb in range(tb):
            zbp = -(sec.b * 0.5) + lbs * (zb + 0.5)
            rbp = sec.R + tbs * (rb + 0.5)
            for zi in range(li):
                for ri in range(ti):
                    z = sec.pos[2] + zbp + (lbs * 0.5) * xl[zi]
                    r = rbp + (tbs * 0.5) * xt[ri]
                    pts.append((r, 0.0, z))
                    w.append(0.25 * r * wl[zi] * wt[ri])
    return np.array(pts), np.array(w), lb * tb


def mutual_inductance_zaxis_slow(prim, sec, args12, field_args6, axis_ratio=AXIS_RATIO_GPU, extended=False):
    """Coil::calculateMutualInductanceZAxisSlow (CalculateMInductanceZAxis.cxx:17-82): |A| of the primary at the
    secondary's r-z grid, evaluated with field_args6 (the GPU field rule: one block, the primary increments)."""
    pts, w, blocks = _zaxis_slow_grid(sec, args12)
    a, _, _ = coil_fields(prim, field_args6, pts, True, False, False, axis_ratio, extended=extended)
    m = float(np.sum(np.linalg.norm(a, axis=1) * w)) / blocks
    return m * 2 * np.pi * sec.N / prim.I


def force_zaxis_slow(prim, sec, args12, field_args6, axis_ratio=AXIS_RATIO_GPU, extended=False):
    """Coil::calculateAmpereForceZAxisSlow (CalculateForceZAxis.cxx:16-80): B_x of the primary at the grid."""
    pts, w, blocks = _zaxis_slow_grid(sec, args12)
    _, b, _ = coil_fields(prim, field_args6, pts, False, True, False, axis_ratio, extended=extended)
    f = float(np.sum(b[:, 0] * w)) / blocks
    return -f * 2 * np.pi * sec.N * sec.I
