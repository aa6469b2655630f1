"""ctypes front end of the C ABI in include/ccoil_hip.h (libccoil_hip.so), for numpy and torch callers.

This is the thin Python binding a maintainer would add next to the reference's own bindings: it passes
plain pointers and sizes across the boundary and never computes anything itself. Loading fails loudly
(OSError) when libccoil_hip.so has not been built.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libccoil_hip.so")

OK, ERR_ARGUMENT, ERR_OUT_OF_MEMORY, ERR_DEVICE, ERR_NO_DEVICE, ERR_CAPACITY = 0, -1, -2, -3, -4, -5
A, B, E, G, G_JACOBIAN = 1, 2, 4, 8, 16
FP64, FP32C, FP32, FP64_REFERENCE, FP64_IEEE = 0, 1, 2, 3, 4
DEVICE_POINTERS, NO_NEAR_FP64, FORCE_REFERENCE_SCALE, AUTOTUNE, DIRECT, FAR_FIELD = 1, 2, 4, 8, 16, 32
PAIR_M, PAIR_FT, PAIR_DM = 1, 2, 4
HIP_STREAM_LEGACY = 1   # hipStreamLegacy: the default (null) stream as a stream handle

EXPORTED = ["ccoil_create", "ccoil_destroy", "ccoil_last_error", "ccoil_version", "ccoil_set_stream",
            "ccoil_synchronize", "ccoil_gauss_legendre", "ccoil_fields", "ccoil_pairs", "ccoil_field_gradient",
            "ccoil_pairs_zaxis", "ccoil_self_inductance", "ccoil_get_stream", "ccoil_fields_noalloc",
            "ccoil_set_tuning_file", "ccoil_fields_plan", "ccoil_pairs_plan", "ccoil_plan_launch",
            "ccoil_plan_synchronize", "ccoil_plan_destroy", "ccoil_pairs_grad", "ccoil_far_field_radius"]
GRAD_TERMS = ("current", "R", "a", "b", "px", "py", "pz", "alpha", "beta")
PAIR_GRAD_TERMS = ("sec_x", "sec_y", "sec_z", "sec_alpha", "sec_beta", "sec_R", "sec_a", "sec_b",
                   "prim_x", "prim_y", "prim_z", "prim_alpha", "prim_beta", "prim_R", "prim_a", "prim_b")


class Coils(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64)] + [(k, ctypes.POINTER(ctypes.c_int32)) for k in
                                          ("use_fast", "n_ang", "n_thk", "n_len", "blocks")] + \
               [(k, ctypes.POINTER(ctypes.c_double)) for k in
                ("R", "a", "b", "J", "current", "turns", "frequency", "pos", "alpha", "beta")]


class Report(ctypes.Structure):
    _fields_ = [("node_evaluations", ctypes.c_int64), ("near_pairs", ctypes.c_int64), ("kernel_ms", ctypes.c_double),
                ("block_size", ctypes.c_int32), ("split", ctypes.c_int32), ("main_kernel_ms", ctypes.c_double),
                ("record_ms", ctypes.c_double)]


class CCoilError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"ccoil status {status}: {msg}")
        self.status = status


_lib = None


def lib():
    global _lib
    if _lib is None:
        import hip_runtime
        hip_runtime.preload()
        L = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        dp = ctypes.POINTER(ctypes.c_double)
        L.ccoil_create.argtypes = [ctypes.c_int, ctypes.POINTER(vp)]
        L.ccoil_destroy.argtypes = [vp]
        L.ccoil_last_error.argtypes = [vp]
        L.ccoil_last_error.restype = ctypes.c_char_p
        L.ccoil_version.restype = ctypes.c_char_p
        L.ccoil_set_stream.argtypes = [vp, vp]
        L.ccoil_synchronize.argtypes = [vp]
        L.ccoil_gauss_legendre.argtypes = [vp, ctypes.c_int32, dp, dp]
        L.ccoil_fields.argtypes = [vp, ctypes.POINTER(Coils), vp, ctypes.c_int64, ctypes.c_uint32, ctypes.c_int32,
                                   ctypes.c_uint32, vp, vp, vp, vp, ctypes.POINTER(Report)]
        L.ccoil_fields_noalloc.argtypes = L.ccoil_fields.argtypes
        L.ccoil_set_tuning_file.argtypes = [vp, ctypes.c_char_p]
        L.ccoil_fields_plan.argtypes = [vp, ctypes.POINTER(Coils), vp, ctypes.c_int64, ctypes.c_uint32, ctypes.c_int32,
                                        ctypes.c_uint32, vp, vp, vp, vp, ctypes.POINTER(vp)]
        L.ccoil_pairs_plan.argtypes = [vp, ctypes.POINTER(Coils), ctypes.POINTER(Coils), vp, vp, ctypes.c_int64, vp,
                                       vp, ctypes.c_uint32, ctypes.c_int32, ctypes.c_uint32, vp, vp, vp, vp,
                                       ctypes.POINTER(vp)]
        L.ccoil_pairs_grad.argtypes = [vp, ctypes.POINTER(Coils), ctypes.POINTER(Coils), vp, vp, ctypes.c_int64, vp,
                                       vp, ctypes.c_int32, ctypes.c_uint32, vp, vp, ctypes.POINTER(Report)]
        L.ccoil_far_field_radius.argtypes = [ctypes.POINTER(Coils), ctypes.c_int32, dp, dp]
        L.ccoil_plan_launch.argtypes = [vp, vp]
        L.ccoil_plan_synchronize.argtypes = [vp, ctypes.POINTER(Report)]
        L.ccoil_plan_destroy.argtypes = [vp]
        L.ccoil_pairs.argtypes = [vp, ctypes.POINTER(Coils), ctypes.POINTER(Coils), vp, vp, ctypes.c_int64, vp, vp,
                                  ctypes.c_uint32, ctypes.c_int32, ctypes.c_uint32, vp, vp, vp, vp, ctypes.POINTER(Report)]
        L.ccoil_field_gradient.argtypes = [vp, ctypes.POINTER(Coils), vp, ctypes.c_int64, vp, ctypes.c_int32,
                                           ctypes.c_uint32, vp, ctypes.POINTER(Report)]
        L.ccoil_pairs_zaxis.argtypes = [vp, ctypes.POINTER(Coils), ctypes.POINTER(Coils), vp, vp, ctypes.c_int64,
                                        ctypes.c_uint32, ctypes.c_uint32, vp, vp, ctypes.POINTER(Report)]
        L.ccoil_self_inductance.argtypes = [vp, ctypes.POINTER(Coils), vp, ctypes.c_uint32, vp,
                                            ctypes.POINTER(Report)]
        for f in ("ccoil_create", "ccoil_destroy", "ccoil_set_stream", "ccoil_synchronize", "ccoil_gauss_legendre",
                  "ccoil_fields", "ccoil_fields_noalloc", "ccoil_set_tuning_file", "ccoil_fields_plan",
                  "ccoil_pairs_plan", "ccoil_plan_launch", "ccoil_plan_synchronize", "ccoil_plan_destroy", "ccoil_pairs_grad", "ccoil_pairs", "ccoil_field_gradient", "ccoil_pairs_zaxis",
                  "ccoil_self_inductance", "ccoil_far_field_radius"):
            getattr(L, f).restype = ctypes.c_int32
        _lib = L
    return _lib


def version():
    return lib().ccoil_version().decode()


def far_field_radius(coils, precision):
    """(r_max, far-field distance) per coil for field calls of the given precision (host only; distance 0: the coil
    keeps its rule)."""
    n = coils.n
    rmax, dist = np.zeros(n), np.zeros(n)
    dp = ctypes.POINTER(ctypes.c_double)
    st = lib().ccoil_far_field_radius(ctypes.byref(coils.struct), precision, rmax.ctypes.data_as(dp),
                                      dist.ctypes.data_as(dp))
    if st != OK:
        raise CCoilError(st, lib().ccoil_last_error(None).decode())
    return rmax, dist


def _ptr(a):
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        return a.ctypes.data_as(ctypes.c_void_p)
    return ctypes.c_void_p(int(a.data_ptr()))   # torch tensor on the context's device


class CoilSet:
    """SoA coil description (host memory) as the C ABI takes it. Arrays are kept alive by the object."""

    def __init__(self, use_fast, n_ang, n_thk, R, a, b, J, pos, alpha, beta, n_len=None, blocks=None,
                 current=None, turns=None, frequency=None):
        i32 = lambda v: None if v is None else np.ascontiguousarray(v, dtype=np.int32)
        f64 = lambda v: None if v is None else np.ascontiguousarray(v, dtype=np.float64)
        self.arrays = dict(use_fast=i32(use_fast), n_ang=i32(n_ang), n_thk=i32(n_thk), n_len=i32(n_len),
                           blocks=i32(blocks), R=f64(R), a=f64(a), b=f64(b), J=f64(J), current=f64(current),
                           turns=f64(turns), frequency=f64(frequency), pos=f64(pos), alpha=f64(alpha),
                           beta=f64(beta))
        self.n = len(self.arrays["R"])
        s = Coils()
        s.n = self.n
        for k, v in self.arrays.items():
            if v is None:
                continue
            ct = ctypes.c_int32 if v.dtype == np.int32 else ctypes.c_double
            setattr(s, k, v.ctypes.data_as(ctypes.POINTER(ct)))
        self.struct = s

    @staticmethod
    def from_rows(rows):
        """rows: iterable of dicts with keys use_fast, n_ang, n_thk, n_len, blocks(3), R, a, b, J, current, turns,
        frequency, pos(3), alpha, beta."""
        rows = list(rows)
        col = lambda k, d=None: [r.get(k, d) for r in rows]
        return CoilSet(col("use_fast"), col("n_ang"), col("n_thk"), col("R"), col("a"), col("b"), col("J"),
                       np.array(col("pos")).reshape(-1), col("alpha"), col("beta"), n_len=col("n_len", 1),
                       blocks=np.array(col("blocks", (1, 1, 1))).reshape(-1), current=col("current", 1.0),
                       turns=col("turns", 1.0), frequency=col("frequency", 0.0))


class Plan:
    """A field or pair call captured into a HIP graph (ccoil_fields_plan / ccoil_pairs_plan); launch() replays it."""

    def __init__(self, ctx, handle, keep):
        self.ctx, self.h, self.keep = ctx, handle, keep

    def launch(self, stream=None):
        self.ctx._check(lib().ccoil_plan_launch(self.h, ctypes.c_void_p(stream) if stream else None))

    def synchronize(self):
        rep = Report()
        self.ctx._check(lib().ccoil_plan_synchronize(self.h, ctypes.byref(rep)))
        return rep

    def close(self):
        if self.h:
            lib().ccoil_plan_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Context:
    def __init__(self, device=0):
        self.h = ctypes.c_void_p()
        st = lib().ccoil_create(device, ctypes.byref(self.h))
        if st != OK:
            raise CCoilError(st, lib().ccoil_last_error(None).decode())

    def close(self):
        if self.h:
            lib().ccoil_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st):
        if st != OK:
            raise CCoilError(st, lib().ccoil_last_error(self.h).decode())

    def set_stream(self, stream_ptr):
        """Orders the context's work on a HIP stream: None restores the context's own stream; 0 (torch's
        cuda_stream for the default stream) is the default (null) stream, passed as hipStreamLegacy."""
        if stream_ptr is None:
            handle = None
        else:
            handle = ctypes.c_void_p(stream_ptr if stream_ptr else HIP_STREAM_LEGACY)
        self._check(lib().ccoil_set_stream(self.h, handle))

    def set_tuning_file(self, path):
        self._check(lib().ccoil_set_tuning_file(self.h, None if path is None else str(path).encode()))

    def synchronize(self):
        self._check(lib().ccoil_synchronize(self.h))

    def gauss_legendre(self, n):
        x = np.zeros(n)
        w = np.zeros(n)
        dp = ctypes.POINTER(ctypes.c_double)
        self._check(lib().ccoil_gauss_legendre(self.h, n, x.ctypes.data_as(dp), w.ctypes.data_as(dp)))
        return x, w

    def fields(self, coils, points, quantities, precision=FP32C, flags=0, out=None):
        """Host numpy (n,3) points -> dict of numpy outputs; with DEVICE_POINTERS pass torch tensors and `out`."""
        rep = Report()
        dev = flags & DEVICE_POINTERS
        n = points.shape[0]
        if out is None:
            out = {}
            for key, bit, w in (("A", A, 3), ("B", B, 3), ("E", E, 3), ("G", G | G_JACOBIAN, 9)):
                if quantities & bit:
                    out[key] = np.zeros((n, w)) if not dev else None
        if not dev:
            points = np.ascontiguousarray(points, dtype=np.float64)
        self._check(lib().ccoil_fields(self.h, ctypes.byref(coils.struct), _ptr(points), n, quantities, precision,
                                       flags, _ptr(out.get("A")), _ptr(out.get("B")), _ptr(out.get("E")),
                                       _ptr(out.get("G")), ctypes.byref(rep)))
        return out, rep

    def fields_noalloc(self, coils, points, quantities, precision, out):
        """ccoil_fields_noalloc: torch tensors on the context's device (points (n, 3), out: dict of outputs)."""
        rep = Report()
        self._check(lib().ccoil_fields_noalloc(self.h, ctypes.byref(coils.struct), _ptr(points), points.shape[0],
                                               quantities, precision, DEVICE_POINTERS, _ptr(out.get("A")),
                                               _ptr(out.get("B")), _ptr(out.get("E")), _ptr(out.get("G")),
                                               ctypes.byref(rep)))
        return out, rep

    def fields_plan(self, coils, points, quantities, precision, out, flags=0):
        """ccoil_fields_plan over torch tensors on the context's device (points (n, 3), out: dict of outputs)."""
        h = ctypes.c_void_p()
        self._check(lib().ccoil_fields_plan(self.h, ctypes.byref(coils.struct), _ptr(points), points.shape[0],
                                            quantities, precision, flags | DEVICE_POINTERS, _ptr(out.get("A")),
                                            _ptr(out.get("B")), _ptr(out.get("E")), _ptr(out.get("G")),
                                            ctypes.byref(h)))
        return Plan(self, h, (coils, points, out))

    def pairs_plan(self, prim, sec, n_pairs, what, precision, out, flags=0, prim_index=None, sec_index=None,
                   ring_offset=None, half_ring=None):
        """ccoil_pairs_plan; out: dict of torch tensors M (n,), F, T, dM (n, 3) on the context's device."""
        pi = None if prim_index is None else np.ascontiguousarray(prim_index, dtype=np.int64)
        si = None if sec_index is None else np.ascontiguousarray(sec_index, dtype=np.int64)
        ro = None if ring_offset is None else np.ascontiguousarray(ring_offset, dtype=np.float64)
        hr = None if half_ring is None else np.ascontiguousarray(half_ring, dtype=np.int32)
        h = ctypes.c_void_p()
        self._check(lib().ccoil_pairs_plan(self.h, ctypes.byref(prim.struct), ctypes.byref(sec.struct), _ptr(pi),
                                           _ptr(si), n_pairs, _ptr(ro), _ptr(hr), what, precision,
                                           flags | DEVICE_POINTERS, _ptr(out.get("M")), _ptr(out.get("F")),
                                           _ptr(out.get("T")), _ptr(out.get("dM")), ctypes.byref(h)))
        return Plan(self, h, (prim, sec, out))

    def field_gradient(self, coils, points, dL_dB, precision=FP64, flags=0, out=None):
        """Adjoint of the summed B: (n_coils, 9) gradients in GRAD_TERMS order. Host numpy arrays, or torch tensors
        with DEVICE_POINTERS and a preallocated `out`."""
        rep = Report()
        dev = flags & DEVICE_POINTERS
        n = points.shape[0]
        if not dev:
            points = np.ascontiguousarray(points, dtype=np.float64)
            dL_dB = np.ascontiguousarray(dL_dB, dtype=np.float64)
            if dL_dB.shape != points.shape:
                raise ValueError("dL_dB must have the shape of points")
        if out is None:
            out = np.zeros((coils.n, len(GRAD_TERMS)))
        self._check(lib().ccoil_field_gradient(self.h, ctypes.byref(coils.struct), _ptr(points), n, _ptr(dL_dB),
                                               precision, flags, _ptr(out), ctypes.byref(rep)))
        return out, rep

    def pairs(self, prim, sec, n_pairs, what, precision=FP32C, flags=0, prim_index=None, sec_index=None,
              ring_offset=None, half_ring=None, out=None):
        rep = Report()
        if out is None:
            out = {}
            if what & PAIR_M:
                out["M"] = np.zeros(n_pairs)
            if what & PAIR_FT:
                out["F"] = np.zeros((n_pairs, 3))
                out["T"] = np.zeros((n_pairs, 3))
            if what & PAIR_DM:
                out["dM"] = np.zeros((n_pairs, 3))
        pi = None if prim_index is None else np.ascontiguousarray(prim_index, dtype=np.int64)
        si = None if sec_index is None else np.ascontiguousarray(sec_index, dtype=np.int64)
        ro = None if ring_offset is None else np.ascontiguousarray(ring_offset, dtype=np.float64)
        hr = None if half_ring is None else np.ascontiguousarray(half_ring, dtype=np.int32)
        self._check(lib().ccoil_pairs(self.h, ctypes.byref(prim.struct), ctypes.byref(sec.struct), _ptr(pi), _ptr(si),
                                      n_pairs, _ptr(ro), _ptr(hr), what, precision, flags, _ptr(out.get("M")),
                                      _ptr(out.get("F")), _ptr(out.get("T")), _ptr(out.get("dM")),
                                      ctypes.byref(rep)))
        return out, rep

    def pairs_grad(self, prim, sec, n_pairs, precision=FP64, flags=0, prim_index=None, sec_index=None,
                   ring_offset=None, half_ring=None):
        """ccoil_pairs_grad: (M (n,), grad (n, 16) in PAIR_GRAD_TERMS order), host numpy arrays."""
        rep = Report()
        M = np.zeros(n_pairs)
        grad = np.zeros((n_pairs, len(PAIR_GRAD_TERMS)))
        pi = None if prim_index is None else np.ascontiguousarray(prim_index, dtype=np.int64)
        si = None if sec_index is None else np.ascontiguousarray(sec_index, dtype=np.int64)
        ro = None if ring_offset is None else np.ascontiguousarray(ring_offset, dtype=np.float64)
        hr = None if half_ring is None else np.ascontiguousarray(half_ring, dtype=np.int32)
        self._check(lib().ccoil_pairs_grad(self.h, ctypes.byref(prim.struct), ctypes.byref(sec.struct), _ptr(pi),
                                           _ptr(si), n_pairs, _ptr(ro), _ptr(hr), precision, flags, _ptr(M),
                                           _ptr(grad), ctypes.byref(rep)))
        return M, grad, rep

    def pairs_zaxis(self, prim, sec, n_pairs, what, flags=0, prim_index=None, sec_index=None, out=None):
        """Coaxial pairs: {"M": (n,), "Fz": (n,)} (axial force on the secondary)."""
        rep = Report()
        if out is None:
            out = {}
            if what & PAIR_M:
                out["M"] = np.zeros(n_pairs)
            if what & PAIR_FT:
                out["Fz"] = np.zeros(n_pairs)
        pi = None if prim_index is None else np.ascontiguousarray(prim_index, dtype=np.int64)
        si = None if sec_index is None else np.ascontiguousarray(sec_index, dtype=np.int64)
        self._check(lib().ccoil_pairs_zaxis(self.h, ctypes.byref(prim.struct), ctypes.byref(sec.struct), _ptr(pi),
                                            _ptr(si), n_pairs, what, flags, _ptr(out.get("M")), _ptr(out.get("Fz")),
                                            ctypes.byref(rep)))
        return out, rep

    def self_inductance(self, coils, radial, flags=0, out=None):
        """Self-inductance of every coil; radial: (n, 2) secondary thickness (blocks, increments)."""
        rep = Report()
        radial = np.ascontiguousarray(radial, dtype=np.int32).reshape(-1)
        if out is None:
            out = np.zeros(coils.n)
        self._check(lib().ccoil_self_inductance(self.h, ctypes.byref(coils.struct), _ptr(radial), flags, _ptr(out),
                                                ctypes.byref(rep)))
        return out, rep
