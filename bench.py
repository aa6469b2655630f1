#!/usr/bin/env python3
"""Throughput of the MI355X C-Coil field path, one JSON line on stdout.

Timed workload (one step = one pass of the hot path over the whole batch, inputs resident in HBM):
  B and the gradient tensor of a 128-coil mixed CoilGroup (FILAMENT/THIN/FLAT/RECTANGULAR, R ~ U[0.02, 0.5] m,
  a, b ~ U[0.005, 0.2] m, N ~ U{1..2000}, I ~ U[-10, 10] A, centres ~ U[-1, 1]^3 m, random orientation) at
  2^20 points ~ U[-1.5, 1.5]^3 m, PrecisionFactor 5, through one ccoil_fields call with device pointers, by default
  in CCOIL_FP64 with CCOIL_FAR_FIELD: the quadrature rule in double (cancellation-free integrands, Newton-refined
  hardware reciprocals) for the (coil, point) pairs inside each coil's far-field distance, and beyond it the coil's
  exterior zonal-harmonic series in double, which reproduces the rule to ~1e-14 there (the fastest mode that meets
  the 1e-6 relative-error bar; --precision picks another).
value = coil·point evaluations of B + gradient per second (coils x points / seconds per step).

Reported beside it in the same line:
  side         the same pass in CCOIL_FP64 (the rule at every pair), CCOIL_FP32C (float rule inside the far-field
               distance, FP64 series beyond, FP64 near-winding pairs), CCOIL_FP32C + CCOIL_DIRECT (the float rule at
               every pair), CCOIL_FP64_IEEE and CCOIL_FP64_REFERENCE (the reference's integrands in double: the
               plain-FP64 comparator): speed and roofline fraction;
  accuracy     every other mode against CCOIL_FP64 at every point (max / p99 / p99.9, points inside the
               near-winding shell of any coil and the rest), relative to the field and relative to the sum of the
               coils' magnitudes at the point (the superposition's conditioning);
  pairs        mutual inductance + force + torque of 65,536 random coil pairs with per-dimension increments
               ~ U{4, 8, ..., 80} (one ccoil_pairs call) in CCOIL_FP64 and CCOIL_FP32C: pairs/s, node-pairs/s and the
               FP32C error;
  roofline     field_kernel's algorithmic flops per launch (frozen per-node model of the rule and per-term model of
               the series, roofline_model.py, with the kernel's own split of the pairs) over its hipEvent-timed
               duration on the library's stream, against the MI355X vector peak of each part's type;
  adjoint      the coil-design workload (64 THIN/RECTANGULAR coils near the axis, 4,096 points in a small
               cylinder): forward B (ccoil_fields) and the adjoint of a field loss (ccoil_field_gradient, all 9
               parameters of every coil), FP64 and FP32C, the adjoint / forward time ratio, and the forward's
               per-call wall time eager and replayed from a captured plan (HIP graph);
  coaxial      4,096 coaxial THIN/RECTANGULAR pairs through the z-axis fast rules (ccoil_pairs_zaxis, M + Fz);
  io           small-call cost through the c_coil module: Vector3Array API vs numpy overload vs DLPack device tensors,
               and through the C ABI on device pointers, eager and replayed from a captured plan;
  cpu_baseline the FP64 CPU restatement of the reference's CPU_MT field path (-O3 -mavx2 -ffast-math, point chunks on
               a thread pool), timed on a bounded sample of the same workload on this host.
With --large (not part of the default run): B+gradient of a 1,024-coil group at 2^22 points, and the full
1,024 x 1,024 mutual-inductance matrix with its position gradient (CCOIL_PAIR_DM) checked against F / (I_i I_j).
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "c-coil_amd")
if PKG not in sys.path:
    sys.path.insert(0, PKG)

METRIC = "coil·point B+∇B evals/s and M-pairs/s, FP32-comp vs FP64 at ≤1e-6 rel err"
# committed rocprofv3 --pmc summaries per mode (HBM bytes per launch of each kernel they measure)
PMC_FILES = {"fp64": os.path.join(ROOT, "profiles", "field_kernel_fp64_group128_points1M_pmc.md"),
             "fp64_far_field": os.path.join(ROOT, "profiles", "field_pass_fp64_far_field_group128_points1M_pmc.md")}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def rel_err(x, ref):
    x = np.asarray(x).reshape(len(ref), -1)
    ref = np.asarray(ref).reshape(len(ref), -1)
    den = np.linalg.norm(ref, axis=1)
    num = np.linalg.norm(x - ref, axis=1)
    ok = den > 0
    return num[ok] / den[ok], ok


def stats(e):
    if e.size == 0:
        return {"max": None, "p99": None, "p999": None, "n": 0}
    return {"max": float(np.max(e)), "p99": float(np.percentile(e, 99)), "p999": float(np.percentile(e, 99.9)),
            "n": int(e.size)}


def pmc_traffic(mode, kernel):
    """HBM bytes per launch of a mode's kernel from its committed rocprofv3 --pmc summary (lines
    `traffic_bytes_per_launch[kernel] = bytes`), if present."""
    path = PMC_FILES.get(mode)
    if path is None or not os.path.exists(path):
        return None
    key = f"traffic_bytes_per_launch[{kernel}]"
    with open(path) as f:
        for line in f:
            if line.startswith(key):
                return float(line.split("=")[1].split()[0])
    return None


def cpu_baseline(params, pts, rows, n_sample):
    """FP64 CPU restatement of the reference's CPU_MT path, timed on this host (bounded sample)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    threads = min(16, len(os.sched_getaffinity(0)))
    coils = [O.make_coil(p["R"], p["a"], p["b"], int(p["N"]), p["I"], 50.0, tuple(float(x) for x in p["pos"]),
                         p["alpha"], p["beta"]) for p in params]
    args = [np.array([1, 1, 1, r["n_ang"], r["n_thk"], r["n_len"]], dtype=np.int32) for r in rows]
    sample = np.ascontiguousarray(pts[:n_sample])
    O.group_fields(coils[:2], args[:2], sample[:256], False, True, True, axis_ratio=O.AXIS_RATIO_GPU,
                   threads=threads, fast=True)
    t0 = time.perf_counter()
    O.group_fields(coils, args, sample, False, True, True, axis_ratio=O.AXIS_RATIO_GPU, threads=threads, fast=True)
    dt = time.perf_counter() - t0
    return {"value": len(coils) * len(sample) / dt, "unit": "coil·point B+∇B evals/s", "cores": threads,
            "kind": "port", "seconds": dt,
            "sample": f"all {len(coils)} coils x the first {len(sample)} points of the timed workload, FP64, "
                      f"oracle/liboracle_fast.so (-O3 -mavx2 -ffast-math), point chunks on {threads} threads"}


def cpu_coaxial_baseline(cpp, csp, cpr, csr, n_sample):
    """The oracle's fast coaxial rules (the reference's CPU_ST calculateMutualInductanceZAxisFast and
    calculateAmpereForceZAxisFast, -O3 -mavx2 -ffast-math) on the first n_sample pairs of the coaxial leg, one core."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    O.lib(True)
    mk = lambda p: O.make_coil(p["R"], p["a"], p["b"], int(p["N"]), p["I"], 50.0, tuple(float(x) for x in p["pos"]),
                               p["alpha"], p["beta"])
    a12 = lambda rp, rs: np.array(list(rp["blocks"]) + [rp["n_ang"], rp["n_thk"], rp["n_len"]] + list(rs["blocks"]) +
                                  [rs["n_ang"], rs["n_thk"], rs["n_len"]], dtype=np.int32)
    L = O.lib(True)
    work = [(mk(p), mk(s), a12(rp, rs)) for p, s, rp, rs in zip(cpp[:n_sample], csp[:n_sample], cpr, csr)]
    t0 = time.perf_counter()
    for p, s, a in work:
        L.orc_zaxis_rule(ctypes.byref(p), ctypes.byref(s), a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), 0, 0)
        L.orc_zaxis_rule(ctypes.byref(p), ctypes.byref(s), a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), 1, 0)
    dt = time.perf_counter() - t0
    return {"value": len(work) / dt, "unit": "coaxial pairs (M + Fz)/s", "cores": 1, "kind": "port",
            "seconds": dt, "sample": f"the first {len(work)} pairs of the coaxial leg, oracle/liboracle_fast.so"}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--coils", type=int, default=128)
    ap.add_argument("--points", type=int, default=1 << 20)
    ap.add_argument("--pairs", type=int, default=65536)
    ap.add_argument("--side-steps", type=int, default=3, help="timed steps of each secondary field mode")
    ap.add_argument("--cpu-points", type=int, default=1 << 19)
    ap.add_argument("--precision", choices=("fp64_far_field", "fp64", "fp32c", "fp32c_direct"),
                    default="fp64_far_field", help="mode of the timed steps")
    ap.add_argument("--no-side", action="store_true", help="skip the secondary field modes and the accuracy report")
    ap.add_argument("--no-pairs", action="store_true", help="skip the pair leg")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint leg")
    ap.add_argument("--no-coaxial", action="store_true", help="skip the coaxial-pair leg")
    ap.add_argument("--coaxial-pairs", type=int, default=4096)
    ap.add_argument("--no-io", action="store_true", help="skip the small-call I/O leg")
    ap.add_argument("--large", action="store_true", help="also run the 1,024-coil field and M-matrix legs")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 512 MiB device-to-device copy 3 times (calibrates the HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        log("bench.py measures one GPU: the field and pair paths shard by points / pairs with no exchange, and "
            "this benchmark has no multi-GPU mode. Run with --gpus 1.")
        sys.exit(2)

    import torch
    import ccoil_abi as abi
    import roofline_model as RM
    import workloads as W

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    if args.copy_calibration:
        src = torch.ones(64 << 20, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": src.numel() * 8, "copy_bytes_written": dst.numel() * 8, "launches": 3}))
        return
    params, pts = W.group_workload(args.coils, args.points, seed=0)
    rows = W.group_rows(params)
    cs = abi.CoilSet.from_rows(rows)
    n = len(pts)
    log(f"workload: {len(rows)} coils x {n} points, mean nodes/coil "
        f"{np.mean([r['n_ang'] * r['n_thk'] for r in rows]):.1f}")

    ctx = abi.Context(0)
    stream = torch.cuda.current_stream(dev)
    ctx.set_stream(stream.cuda_stream)
    d_pts = torch.from_numpy(np.ascontiguousarray(pts)).to(dev)
    q = abi.B | abi.G
    flops, trans = RM.field_work(rows, n, with_gradient=True)

    def new_out():
        return {"B": torch.empty((n, 3), dtype=torch.float64, device=dev),
                "G": torch.empty((n, 9), dtype=torch.float64, device=dev)}

    def timed(prec, steps, warmup, out, flags=0):
        for _ in range(warmup):
            ctx.fields(cs, d_pts, q, prec, abi.DEVICE_POINTERS | flags, out=out)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        reps = [ctx.fields(cs, d_pts, q, prec, abi.DEVICE_POINTERS | flags, out=out)[1] for _ in range(steps)]
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / steps
        main_ms = float(np.mean([r.main_kernel_ms for r in reps]))
        all_ms = float(np.mean([r.kernel_ms for r in reps]))
        rec_ms = float(np.mean([r.record_ms for r in reps]))
        return dt, main_ms, all_ms, reps[-1], rec_ms

    # field modes: (precision, flags, label); the series modes take each coil's far-field series beyond its
    # distance (ccoil_far_field_radius) and the rule inside it
    modes = {"fp64_far_field": (abi.FP64, abi.FAR_FIELD, "CCOIL_FP64 + CCOIL_FAR_FIELD"),
             "fp64": (abi.FP64, 0, "CCOIL_FP64"),
             "fp32c": (abi.FP32C, 0, "CCOIL_FP32C (far-field series)"),
             "fp32c_direct": (abi.FP32C, abi.DIRECT, "CCOIL_FP32C + CCOIL_DIRECT"),
             "fp64_ieee": (abi.FP64_IEEE, 0, "CCOIL_FP64_IEEE"),
             "fp64_reference": (abi.FP64_REFERENCE, 0, "CCOIL_FP64_REFERENCE")}
    works = {}

    def mode_work(name):
        """Algorithmic work of one call of the mode, by kernel: {part: (fp32 flops, fp64 flops, special ops)}. The
        series modes at this size run field_kernel for the series alone ("series") and the record pass for every pair
        inside the far-field distance ("records": the FP64 rule, or FP32C's float rule plus its FP64 near pairs);
        the others run the rule in field_kernel ("field")."""
        if name not in works:
            prec, flags, _ = modes[name]
            series = (prec == abi.FP32C and not flags & abi.DIRECT) or (prec == abi.FP64 and flags & abi.FAR_FIELD)
            if series:
                rmax, dist = abi.far_field_radius(cs, prec)
                w = RM.series_field_work(rows, pts, dist, rmax, RM.LOG_TOL_FP32C if prec == abi.FP32C
                                         else RM.LOG_TOL_FP64, with_gradient=True, near_pass=prec == abi.FP32C)
                rec = (w["rule_flops"], w["near_flops"], w["rule_special"]) if prec == abi.FP32C else \
                    (0.0, w["rule_flops"], w["rule_special"])
                works[name] = ({"series": (0.0, w["series_flops"], w["series_special"]), "records": rec}, w)
            else:
                f32 = prec in (abi.FP32C, abi.FP32)
                works[name] = ({"field": (flops if f32 else 0.0, 0.0 if f32 else flops, trans)}, None)
        return works[name]

    def part_roofline(kernel, work, ms):
        f32, f64, sp = work
        t = ms * 1e-3
        at_peak = f32 / (RM.PEAK_FP32_TFLOPS * 1e12) + f64 / (RM.PEAK_FP64_TFLOPS * 1e12)
        achieved = (f32 + f64) / t / 1e12
        peak = (f32 + f64) / at_peak / 1e12
        return {"bound": "valu", "achieved": achieved, "peak": peak, "unit": "TFLOP/s", "frac": achieved / peak,
                "kernel": kernel, "kernel_ms": ms, "algorithmic_flops_per_launch": f32 + f64, "fp32_flops": f32,
                "fp64_flops": f64, "transcendental": {"achieved_ops": sp / t, "peak_ops": RM.PEAK_TRANS_OPS,
                                                      "frac": sp / t / RM.PEAK_TRANS_OPS}}

    def roofline(name, main_ms, rec_ms, all_ms):
        """The dominant kernel's roofline, with the other part and the whole call (all kernels) beside it."""
        parts, w = mode_work(name)
        if "field" in parts:
            rl = part_roofline("field_kernel", parts["field"], main_ms)
        else:
            ser = part_roofline("field_series_kernel", parts["series"], main_ms)
            rec = part_roofline("near_eval_kernel (rule records)", parts["records"], rec_ms)
            rl = dict(rec if rec_ms >= main_ms else ser)
            tot = tuple(a + b for a, b in zip(parts["series"], parts["records"]))
            rl["parts"] = {"series": ser, "records": rec, "call": part_roofline("all kernels of the call", tot, all_ms)}
            rl["pairs"] = {"rule": w["rule_pairs"], "series": w["series_pairs"], "near_pass": w["near_pairs"]}
        # field_kernel: the points in, B and the gradient out; near_eval_kernel: per record its key, point and slot
        # index and its point in, its B and gradient out
        records = (w["rule_pairs"] + w["near_pairs"]) if w is not None else 0
        rl["hbm_algorithmic_bytes_per_launch"] = records * (4 + 8 + 24 + 96) if rl["kernel"].startswith("near_eval") \
            else n * (24 + 24 + 72)
        return rl

    head_name = args.precision
    head_prec, head_flags, head_label = modes[head_name]
    head = new_out()
    dt, main_ms, all_ms, rep, rec_ms = timed(head_prec, args.steps, args.warmup, head, head_flags)
    value = len(rows) * n / dt
    log(f"{head_name}: {dt * 1e3:.2f} ms/step, field kernel {main_ms:.2f} ms, record pass {rec_ms:.2f} ms, "
        f"all kernels {all_ms:.2f} ms")
    rl = roofline(head_name, main_ms, rec_ms, all_ms)
    rl["traffic"] = pmc_traffic(head_name, rl["kernel"].split()[0])
    result = {"metric": METRIC, "value": value, "unit": "coil·point evals/s", "n_gpus": 1,
              "steps": args.steps, "warmup": args.warmup, "ms_per_step": dt * 1e3, "higher_is_better": True,
              "scaling": "weak", "vs_baseline": None, "dtype": "f32" if head_prec == abi.FP32C else "f64",
              "data": "synthetic",
              "config": {"workload": f"B+gradient of a {len(rows)}-coil mixed CoilGroup at {n} points",
                         "coils": len(rows), "points": n, "precision_factor": 5.0, "quantities": "B+G",
                         "precision": head_label, "seed": 0},
              "roofline": rl, "kernel_ms": {"field_kernel": main_ms, "record_pass": rec_ms, "all": all_ms}}

    if not args.no_side:
        side, outs = {}, {head_name: head}
        for name, (prec, flags, label) in modes.items():
            if name == head_name:
                continue
            outs[name] = new_out()
            sdt, sms, sall, srep, srec = timed(prec, args.side_steps, 1, outs[name], flags)
            srl = roofline(name, sms, srec, sall)
            side[name] = {"mode": label, "value": len(rows) * n / sdt, "ms_per_step": sdt * 1e3, "kernel_ms": sms,
                          "record_ms": srec, "all_ms": sall, "roofline_kernel": srl["kernel"],
                          "roofline_frac": srl["frac"], "near_pairs": int(srep.near_pairs),
                          "value_over_headline": (len(rows) * n / sdt) / value}
            if "parts" in srl:
                side[name]["call_frac"] = srl["parts"]["call"]["frac"]
            log(f"{name}: {sdt * 1e3:.1f} ms/step, field kernel {sms:.2f} ms, record pass {srec:.2f} ms, all kernels "
                f"{sall:.2f} ms")
        result["side"] = side
        b64, g64 = outs["fp64"]["B"].cpu().numpy(), outs["fp64"]["G"].cpu().numpy()
        near = W.near_mask(rows, pts)
        acc = {"against": "CCOIL_FP64 (the rule in double) on the same points", "near_points": int(near.sum())}
        # the superposition's own conditioning: sum over coils of |B_k| and |G_k| (Frobenius) per point, from one
        # CCOIL_FP64 launch per coil
        mb = torch.zeros(n, dtype=torch.float64, device=dev)
        mg = torch.zeros(n, dtype=torch.float64, device=dev)
        one = new_out()
        for r in rows:
            ctx.fields(abi.CoilSet.from_rows([r]), d_pts, q, abi.FP64, abi.DEVICE_POINTERS, out=one)
            mb += torch.linalg.vector_norm(one["B"], dim=1)
            mg += torch.linalg.vector_norm(one["G"], dim=1)
        mb, mg = mb.cpu().numpy(), mg.cpu().numpy()
        acc["conditioned"] = ("error over the sum of the coils' magnitudes at the point, sum_k |B_k| "
                              "(sum_k |G_k| for the gradient)")
        for name in ("fp64_far_field", "fp32c", "fp32c_direct", "fp64_ieee"):
            ob, og = outs[name]["B"].cpu().numpy(), outs[name]["G"].cpu().numpy()
            eb, okb = rel_err(ob, b64)
            eg, okg = rel_err(og, g64)
            cb = np.linalg.norm(ob - b64, axis=1) / mb
            cg = np.linalg.norm(og - g64, axis=1) / mg
            acc[name] = {"B": {"far": stats(eb[~near[okb]]), "near": stats(eb[near[okb]]),
                               "far_conditioned": stats(cb[~near]), "near_conditioned": stats(cb[near])},
                         "G": {"far": stats(eg[~near[okg]]), "near": stats(eg[near[okg]]),
                               "far_conditioned": stats(cg[~near]), "near_conditioned": stats(cg[near])}}
        # A and E (E = 2 pi f A per coil) on the first 2^18 points, every other mode against CCOIL_FP64
        ns = min(n, 1 << 18)
        qae = abi.A | abi.E
        ae = {}
        for name, (prec, flags, label) in modes.items():
            if name == "fp64_reference":
                continue
            o = {"A": torch.empty((ns, 3), dtype=torch.float64, device=dev),
                 "E": torch.empty((ns, 3), dtype=torch.float64, device=dev)}
            ctx.fields(cs, d_pts[:ns], qae, prec, abi.DEVICE_POINTERS | flags, out=o)
            ae[name] = {k: o[k].cpu().numpy() for k in ("A", "E")}
        nm = near[:ns]
        acc["A_E"] = {"points": ns}
        for name in ae:
            if name == "fp64":
                continue
            acc["A_E"][name] = {}
            for k in ("A", "E"):
                e, ok = rel_err(ae[name][k], ae["fp64"][k])
                acc["A_E"][name][k] = {"far": stats(e[~nm[ok]]), "near": stats(e[nm[ok]])}
        result["accuracy"] = acc
        log(f"accuracy: {json.dumps(acc)}")
        if "fp64_reference" in side:
            result["speedup_vs_plain_fp64"] = value / side["fp64_reference"]["value"]

    if not args.no_pairs:
        prow, srow = W.pair_workload(args.pairs, seed=0)
        P, S = abi.CoilSet.from_rows(prow), abi.CoilSet.from_rows(srow)
        what = abi.PAIR_M | abi.PAIR_FT
        pf, ptrans, pnodes = RM.pair_work(prow, srow)
        pres = {}
        # the rule at every node (fp64, fp32c_direct: node-pair rates and algorithmic flops are the rule's), and the
        # primaries' series beyond their far-field distance (fp64_far_field, fp32c: pairs/s)
        for name, prec, flags in (("fp64", abi.FP64, 0), ("fp64_far_field", abi.FP64, abi.FAR_FIELD),
                                  ("fp32c", abi.FP32C, 0), ("fp32c_direct", abi.FP32C, abi.DIRECT)):
            ctx.pairs(P, S, len(prow), what, prec, flags)
            t0 = time.perf_counter()
            o, prep = ctx.pairs(P, S, len(prow), what, prec, flags)
            dt_p = time.perf_counter() - t0
            pres[name] = {"pairs_per_s": len(prow) / dt_p, "ms_per_step": dt_p * 1e3,
                          "kernel_ms": prep.main_kernel_ms, "out": o}
            if name in ("fp64", "fp32c_direct"):
                pres[name]["node_pairs_per_s"] = pnodes / (prep.main_kernel_ms * 1e-3)
                pres[name]["tflops_algorithmic"] = pf / (prep.main_kernel_ms * 1e-3) / 1e12
                peak = RM.PEAK_FP64_TFLOPS if prec == abi.FP64 else RM.PEAK_FP32_TFLOPS
                pres[name]["frac"] = pres[name]["tflops_algorithmic"] / peak
            log(f"pairs {name}: {dt_p * 1e3:.1f} ms/step, kernel {prep.main_kernel_ms:.1f} ms")
        o64, o32 = pres["fp64"].pop("out"), pres["fp32c"].pop("out")
        ofar, o32d = pres["fp64_far_field"].pop("out"), pres["fp32c_direct"].pop("out")
        em = np.abs(o32["M"] - o64["M"]) / np.abs(o64["M"])
        ef, _ = rel_err(o32["F"], o64["F"])
        et, _ = rel_err(o32["T"], o64["T"])
        # M of a random pose can be a small remainder of cancelling contributions; |M| + |grad M| |d| (grad M =
        # F / (I1 I2), d the centre distance) is the size of M over the pose's neighbourhood
        cur = np.array([p.get("current", 1.0) * s.get("current", 1.0) for p, s in zip(prow, srow)])
        dist = np.linalg.norm(np.array([s["pos"] for s in srow]) - np.array([p["pos"] for p in prow]), axis=1)
        mscale = np.abs(o64["M"]) + np.linalg.norm(o64["F"], axis=1) / np.abs(cur) * dist
        emc = np.abs(o32["M"] - o64["M"]) / mscale
        pres["accuracy_fp32c_vs_fp64"] = {"M": stats(em), "M_conditioned": stats(emc), "F": stats(ef), "T": stats(et)}
        for name, o in (("fp64_far_field", ofar), ("fp32c_direct", o32d)):
            pres[f"accuracy_{name}_vs_fp64"] = {
                "M": stats(np.abs(o["M"] - o64["M"]) / np.abs(o64["M"])),
                "M_conditioned": stats(np.abs(o["M"] - o64["M"]) / mscale),
                "F": stats(rel_err(o["F"], o64["F"])[0]), "T": stats(rel_err(o["T"], o64["T"])[0])}
        pres["workload"] = (f"{len(prow)} random mixed pairs, increments ~ U{{4,...,80}} per active dimension, "
                            f"{pnodes:.3g} node-pairs, M+F+T")
        result["pairs"] = pres

    if not args.no_adjoint:
        aparams, apts = W.adjoint_workload(64, 4096, seed=0)
        arows = W.group_rows(aparams)
        acs = abi.CoilSet.from_rows(arows)
        ad_pts = torch.from_numpy(apts).to(dev)
        aB = torch.empty((len(apts), 3), dtype=torch.float64, device=dev)
        agrad = torch.empty((len(arows), len(abi.GRAD_TERMS)), dtype=torch.float64, device=dev)
        gB = torch.from_numpy(np.random.default_rng(1).standard_normal(apts.shape)).to(dev)
        adj = {"workload": "64 THIN/RECTANGULAR coils x 4,096 points, B and the 9-parameter adjoint of a field loss"}
        grads = {}
        for name, prec in (("fp64", abi.FP64), ("fp32c", abi.FP32C)):
            fw, bw = [], []
            for it in range(4):
                _, r1 = ctx.fields(acs, ad_pts, abi.B, prec, abi.DEVICE_POINTERS, out={"B": aB})
                _, r2 = ctx.field_gradient(acs, ad_pts, gB, prec, abi.DEVICE_POINTERS, out=agrad)
                if it:
                    fw.append(r1.kernel_ms)
                    bw.append(r2.kernel_ms)
            grads[name] = agrad.cpu().numpy().copy()
            adj[name] = {"forward_ms": float(np.mean(fw)), "adjoint_ms": float(np.mean(bw)),
                         "adjoint_over_forward": float(np.mean(bw) / np.mean(fw))}
            # per-call wall time of the forward: eager library calls, and replays of a captured plan (HIP graph)
            plan = ctx.fields_plan(acs, ad_pts, abi.B, prec, {"B": aB})
            for key, fn in (("forward_eager_call_ms",
                             lambda: ctx.fields(acs, ad_pts, abi.B, prec, abi.DEVICE_POINTERS, out={"B": aB})),
                            ("forward_plan_replay_ms", plan.launch)):
                fn()
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                for _ in range(50):
                    fn()
                torch.cuda.synchronize()
                adj[name][key] = (time.perf_counter() - t0) / 50 * 1e3
            plan.close()
        # FP32C adjoint against FP64, per parameter kind, relative to that parameter's largest |gradient| over coils
        g64, g32 = grads["fp64"], grads["fp32c"]
        scale = np.max(np.abs(g64), axis=0)
        adj["accuracy_fp32c_vs_fp64"] = {t: stats(np.abs(g32[:, k] - g64[:, k]) / scale[k])
                                         for k, t in enumerate(abi.GRAD_TERMS)}
        result["adjoint"] = adj
        log(f"adjoint: {json.dumps(adj)}")

    if not args.no_coaxial:
        cpp, csp, cpr, csr = W.coaxial_workload(args.coaxial_pairs, seed=0)
        CP, CS = abi.CoilSet.from_rows(cpr), abi.CoilSet.from_rows(csr)
        ctx.pairs_zaxis(CP, CS, len(cpr), abi.PAIR_M | abi.PAIR_FT)
        t0 = time.perf_counter()
        co, crep = ctx.pairs_zaxis(CP, CS, len(cpr), abi.PAIR_M | abi.PAIR_FT)
        cdt = time.perf_counter() - t0
        result["coaxial"] = {"workload": f"{len(cpr)} coaxial THIN/RECTANGULAR pairs, z-axis fast rules, M + Fz, "
                                         f"planned at PrecisionFactor 5 (GPU method)",
                             "pairs_per_s": len(cpr) / cdt, "ms": cdt * 1e3, "kernel_ms": crep.main_kernel_ms,
                             "node_evaluations": int(crep.node_evaluations),
                             "node_evals_per_s": crep.node_evaluations / (crep.main_kernel_ms * 1e-3)}
        log(f"coaxial: {json.dumps(result['coaxial'])}")

    if not args.no_io:
        # small calls through the c_coil module: the reference's Vector3Array API, the numpy overload and DLPack
        # device tensors (compute_all_fields_into), one coil, B at 4,096 points
        import c_coil
        C = c_coil.coil
        C.Coil.set_gpu_precision(C.GPU_FP64)
        icoil = W.make_coil(params[0])
        ipts = np.ascontiguousarray(pts[:4096])
        va = c_coil.tensor.Vector3Array()
        for p in ipts:
            va.append(float(p[0]), float(p[1]), float(p[2]))
        dpts = torch.from_numpy(ipts).to(dev)
        dB = torch.empty((len(ipts), 3), dtype=torch.float64, device=dev)

        def per_call(fn, reps=50):
            fn()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                fn()
            torch.cuda.synchronize()
            return (time.perf_counter() - t0) / reps * 1e3

        def v3():
            r = icoil.compute_all_B_field_vectors(va, C.GPU)
            return np.stack([np.array(r.x()), np.array(r.y()), np.array(r.z())], axis=1)

        io = {"workload": "one coil, B at 4,096 points, CCOIL_FP64, per call including conversions",
              "vector3array_ms": per_call(v3),
              "numpy_ms": per_call(lambda: icoil.compute_all_B_field_vectors(ipts, C.GPU)),
              "dlpack_device_ms": per_call(lambda: icoil.compute_all_fields_into(dpts, B=dB))}
        # the same call through the C ABI on device pointers: eager, and replayed from a captured plan (HIP graph)
        ics = abi.CoilSet.from_rows([W.coil_row(icoil)])
        io["abi_eager_ms"] = per_call(lambda: ctx.fields(ics, dpts, abi.B, abi.FP64, abi.DEVICE_POINTERS,
                                                         out={"B": dB}))
        iplan = ctx.fields_plan(ics, dpts, abi.B, abi.FP64, {"B": dB})
        io["plan_replay_ms"] = per_call(iplan.launch)
        iplan.close()
        C.Coil.set_gpu_precision(C.GPU_FP32C)
        result["io"] = io
        log(f"io: {json.dumps(io)}")

    if args.large:
        lparams, lpts = W.group_workload(1024, 1 << 22, seed=0)
        lrows = W.group_rows(lparams)
        lcs = abi.CoilSet.from_rows(lrows)
        lp = torch.from_numpy(lpts).to(dev)
        lout = {"B": torch.empty((len(lpts), 3), dtype=torch.float64, device=dev),
                "G": torch.empty((len(lpts), 9), dtype=torch.float64, device=dev)}
        large = {}
        for name, prec, lflags in (("fp64", abi.FP64, 0), ("fp64_far_field", abi.FP64, abi.FAR_FIELD),
                                   ("fp32c", abi.FP32C, 0)):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            _, lrep = ctx.fields(lcs, lp, q, prec, abi.DEVICE_POINTERS | lflags, out=lout)
            torch.cuda.synchronize()
            ldt = time.perf_counter() - t0
            large["fields_" + name] = {"value": len(lrows) * len(lpts) / ldt, "ms": ldt * 1e3,
                                       "kernel_ms": lrep.main_kernel_ms}
            log(f"large fields {name}: {ldt * 1e3:.0f} ms")
        prim = abi.CoilSet.from_rows(lrows)
        sec = abi.CoilSet.from_rows(W.secondary_rows(lparams))
        pi, si = W.matrix_pairs(len(lrows))
        t0 = time.perf_counter()
        mo, mrep = ctx.pairs(prim, sec, len(pi), abi.PAIR_M | abi.PAIR_FT | abi.PAIR_DM, abi.FP64,
                             prim_index=pi, sec_index=si)
        mdt = time.perf_counter() - t0
        I = np.array([p["I"] for p in lparams])
        force = mo["F"] / (I[pi] * I[si])[:, None]
        e = np.linalg.norm(mo["dM"] - force, axis=1) / np.linalg.norm(force, axis=1)
        # separation of the coils' bounding spheres: dM (derivative of the A-rule) and F (the B-rule) agree up to
        # the quadrature error, which is large for touching or interlocked coils
        ctr = np.array([p["pos"] for p in lparams])
        rad = np.array([np.hypot(p["R"] + p["a"], 0.5 * p["b"]) for p in lparams])
        sep = np.linalg.norm(ctr[pi] - ctr[si], axis=1) / (rad[pi] + rad[si])
        large["m_matrix_fp64"] = {"pairs": int(len(pi)), "ms": mdt * 1e3, "kernel_ms": mrep.main_kernel_ms,
                                  "pairs_per_s": len(pi) / mdt, "node_pairs": int(mrep.node_evaluations),
                                  "dM_vs_F_over_II": {"all": stats(e), "separated_1": stats(e[sep > 1.0]),
                                                      "separated_2": stats(e[sep > 2.0])}}
        log(f"large M matrix: {mdt * 1e3:.0f} ms, dM vs F/(II) {json.dumps(large['m_matrix_fp64']['dM_vs_F_over_II'])}")
        # the same matrix with the primaries' far-field series, against the rule at every node
        t0 = time.perf_counter()
        so, srep = ctx.pairs(prim, sec, len(pi), abi.PAIR_M | abi.PAIR_FT | abi.PAIR_DM, abi.FP64, abi.FAR_FIELD,
                             prim_index=pi, sec_index=si)
        sdt = time.perf_counter() - t0
        rel = lambda k: (np.linalg.norm((so[k] - mo[k]).reshape(len(pi), -1), axis=1)
                         / np.linalg.norm(mo[k].reshape(len(pi), -1), axis=1))
        large["m_matrix_fp64_far_field"] = {"pairs": int(len(pi)), "ms": sdt * 1e3, "kernel_ms": srep.main_kernel_ms,
                                            "pairs_per_s": len(pi) / sdt,
                                            "vs_rule": {k: stats(rel(k)) for k in ("M", "F", "dM")}}
        log(f"large M matrix with the series: {sdt * 1e3:.0f} ms")
        result["large"] = large

    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(params, pts, rows, min(args.cpu_points, n))
        log(f"cpu baseline: {result['cpu_baseline']['value']:.3g} coil-points/s")
        if not args.no_coaxial:
            result["cpu_baseline"]["coaxial"] = cpu_coaxial_baseline(cpp, csp, cpr, csr, 1024)
            log(f"cpu baseline coaxial: {result['cpu_baseline']['coaxial']['value']:.3g} pairs/s")
    else:
        result["cpu_baseline"] = None
    ctx.close()
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
