"""Synthetic workloads of the benchmark configurations, and packing of coils for the C ABI.

Distributions (numpy PCG64, float64, drawn in the order written):
  single RECTANGULAR coil    R = a = b = 0.1 m, N = 100, I = 1 A at the origin; points uniform in the
                             cylinder r <= 0.3 m, |z| <= 0.3 m (r = 0.3 sqrt(u0), phi = 2 pi u1, z = 0.6 u2 - 0.3)
  mixed coil group           type ~ U{FILAMENT, THIN, FLAT, RECTANGULAR}; R ~ U[0.02, 0.5]; a ~ U[0.005, 0.2] for
                             FLAT/RECTANGULAR else 0; b ~ U[0.005, 0.2] for THIN/RECTANGULAR else 0;
                             N ~ U{1..2000}; I ~ U[-10, 10]; centre ~ U[-1, 1]^3; yaw ~ U[0, 2 pi); pitch ~ U[0, pi]
                             points ~ U[-1.5, 1.5]^3
  uneven pairs               both coils of a pair drawn as in the mixed group; per active quadrature dimension
                             increments ~ U{4, 8, ..., 80}
Planning (PrecisionArguments) comes from the c_coil module, i.e. the same C++ code the drop-in API uses.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)

import ccoil_abi  # noqa: E402

FILAMENT, THIN, FLAT, RECTANGULAR = "FILAMENT", "THIN", "FLAT", "RECTANGULAR"
TYPES = (FILAMENT, THIN, FLAT, RECTANGULAR)


def cylinder_points(rng, n, r_max=0.3, z_half=0.3):
    u = rng.random((n, 3))
    r = r_max * np.sqrt(u[:, 0])
    phi = 2 * np.pi * u[:, 1]
    z = 2 * z_half * u[:, 2] - z_half
    return np.stack([r * np.cos(phi), r * np.sin(phi), z], axis=1)


def cube_points(rng, n, half=1.5):
    return rng.uniform(-half, half, size=(n, 3))


def random_coil_params(rng, n):
    """Mixed-type coils; returns a list of dicts (R, a, b, N, I, pos, alpha, beta, type)."""
    out = []
    for _ in range(n):
        t = TYPES[rng.integers(4)]
        R = rng.uniform(0.02, 0.5)
        a = rng.uniform(0.005, 0.2) if t in (FLAT, RECTANGULAR) else 0.0
        b = rng.uniform(0.005, 0.2) if t in (THIN, RECTANGULAR) else 0.0
        N = int(rng.integers(1, 2001))
        I = rng.uniform(-10, 10)
        pos = rng.uniform(-1, 1, size=3)
        beta = rng.uniform(0, 2 * np.pi)   # zAxisAngle (yaw)
        alpha = rng.uniform(0, np.pi)      # yAxisAngle (pitch)
        out.append(dict(type=t, R=R, a=a, b=b, N=N, I=I, pos=pos, alpha=alpha, beta=beta))
    return out


def _c_coil():
    import c_coil
    return c_coil


def make_coil(p, pf=5.0):
    c = _c_coil()
    V = c.tensor.Vector3
    return c.coil.Coil(p["R"], p["a"], p["b"], int(p["N"]), p.get("I", 1.0), p.get("freq", 50.0),
                       c.coil.PrecisionFactor(pf), 8, V(*[float(x) for x in p["pos"]]), p["alpha"], p["beta"])


def coil_row(coil, args=None, pos=None, alpha=None, beta=None):
    """One C ABI row from a c_coil Coil and its PrecisionArguments (GPU defaults if None)."""
    if args is None:
        args = coil.get_precision_settings_GPU()
    P = coil.get_position_vector()
    ang = coil.get_rotation_angles()
    return dict(use_fast=int(coil.is_using_fast_method()), n_ang=args.angular_increment_count,
                n_thk=args.thickness_increment_count, n_len=args.length_increment_count,
                blocks=(args.angular_block_count, args.thickness_block_count, args.length_block_count),
                R=coil.get_inner_radius(), a=coil.get_thickness(), b=coil.get_length(), J=coil.get_current_density(),
                current=coil.get_current(), turns=float(coil.get_num_of_turns()), frequency=coil.get_sine_frequency(),
                pos=(P.x, P.y, P.z) if pos is None else tuple(pos), alpha=ang[0] if alpha is None else alpha,
                beta=ang[1] if beta is None else beta)


def coilset(coils, pf=5.0):
    """CoilSet of c_coil Coils at their GPU PrecisionArguments for the given group PrecisionFactor."""
    c = _c_coil()
    rows = [coil_row(k, c.coil.PrecisionArguments.get_coil_precision_arguments_GPU(k, c.coil.PrecisionFactor(pf)))
            for k in coils]
    return ccoil_abi.CoilSet.from_rows(rows)


def single_rect_workload(n_points=65536, seed=0):
    rng = np.random.default_rng(seed)
    pts = cylinder_points(rng, n_points)
    params = [dict(type=RECTANGULAR, R=0.1, a=0.1, b=0.1, N=100, I=1.0, pos=np.zeros(3), alpha=0.0, beta=0.0)]
    return params, pts


def group_workload(n_coils=128, n_points=1 << 20, seed=0):
    rng = np.random.default_rng(seed)
    params = random_coil_params(rng, n_coils)
    pts = cube_points(rng, n_points)
    return params, pts


def pair_workload(n_pairs=65536, seed=0):
    """Rows for ccoil_pairs: every pair its own primary and secondary, increments ~ U{4,...,80} per active
    dimension (primary: angular, thickness; secondary: angular, thickness, length)."""
    rng = np.random.default_rng(seed)
    prim = random_coil_params(rng, n_pairs)
    sec = random_coil_params(rng, n_pairs)
    incs = lambda: int(4 * rng.integers(1, 21))
    prim_rows, sec_rows = [], []
    for p, s in zip(prim, sec):
        cp, cs = make_coil(p), make_coil(s)
        pa = (incs(), incs() if p["type"] in (FLAT, RECTANGULAR) else 1, 1)
        sa = (incs(), incs() if s["type"] in (FLAT, RECTANGULAR) else 1, incs() if s["type"] in (THIN, RECTANGULAR) else 1)
        rp = coil_row(cp)
        rp.update(n_ang=pa[0], n_thk=pa[1], n_len=pa[2], blocks=(1, 1, 1))
        rs = coil_row(cs)
        rs.update(n_ang=sa[0], n_thk=sa[1], n_len=sa[2], blocks=(1, 1, 1))
        prim_rows.append(rp)
        sec_rows.append(rs)
    return prim_rows, sec_rows


def group_rows(params, pf=5.0):
    """C ABI rows of coils at their GPU PrecisionArguments for a group PrecisionFactor (the product's planning)."""
    c = _c_coil()
    rows = []
    for p in params:
        k = make_coil(p, pf)
        a = c.coil.PrecisionArguments.get_coil_precision_arguments_GPU(k, c.coil.PrecisionFactor(pf))
        rows.append(coil_row(k, a))
    return rows


def rotation(alpha, beta):
    """The coil's transformation matrix (CalculateAttributes.cxx:65-78) as a 3x3 numpy array. Its entries
    equal the product Rz(beta) Ry(alpha) Rz(beta); T e_z = (cos b sin a, sin b sin a, cos a) is the coil axis."""
    ca, sa, cb, sb = np.cos(alpha), np.sin(alpha), np.cos(beta), np.sin(beta)
    rz = np.array([[cb, -sb, 0.0], [sb, cb, 0.0], [0.0, 0.0, 1.0]])
    ry = np.array([[ca, 0.0, sa], [0.0, 1.0, 0.0], [-sa, 0.0, ca]])
    return rz @ ry @ rz


def near_mask(rows, pts, width_floor=1e-3, chunk=1 << 18):
    """Points within one cross-section width (max(a, b), floored at 1e-3 R) of any coil's winding rectangle, the
    shell where the FP32 path hands (coil, point) pairs to the FP64 evaluator."""
    near = np.zeros(len(pts), dtype=bool)
    for s in range(0, len(pts), chunk):
        P = pts[s:s + chunk]
        m = near[s:s + chunk]
        for r in rows:
            T = rotation(r["alpha"], r["beta"])
            loc = (P - np.asarray(r["pos"])) @ T      # T^T (x - pos), row vectors
            rr = np.hypot(loc[:, 0], loc[:, 1])
            dr = np.maximum(np.maximum(r["R"] - rr, rr - (r["R"] + r["a"])), 0.0)
            dz = np.maximum(np.abs(loc[:, 2]) - 0.5 * r["b"], 0.0)
            w = max(r["a"], r["b"], width_floor * r["R"])
            m |= dr * dr + dz * dz < w * w
    return near


def adjoint_workload(n_coils=64, n_points=4096, seed=0):
    """Coil-design workload: THIN / RECTANGULAR coils near the z axis and a target region on the axis.
    type ~ U{THIN, RECTANGULAR}; R ~ U[0.05, 0.3]; b ~ U[0.01, 0.1]; a ~ U[0.01, 0.1] for RECTANGULAR (0 for THIN);
    N = 500; I ~ U[0.5, 5]; x, y ~ U[-0.05, 0.05], z ~ U[-0.5, 0.5]; tilt alpha ~ U[0, 5 deg], beta ~ U[0, 2 pi).
    Points: z ~ U[-0.2, 0.2], uniform in the disk r <= 0.02."""
    rng = np.random.default_rng(seed)
    params = []
    for _ in range(n_coils):
        t = (THIN, RECTANGULAR)[rng.integers(2)]
        R = rng.uniform(0.05, 0.3)
        b = rng.uniform(0.01, 0.1)
        a = rng.uniform(0.01, 0.1) if t == RECTANGULAR else 0.0
        I = rng.uniform(0.5, 5.0)
        pos = np.array([rng.uniform(-0.05, 0.05), rng.uniform(-0.05, 0.05), rng.uniform(-0.5, 0.5)])
        alpha = rng.uniform(0.0, np.deg2rad(5.0))
        beta = rng.uniform(0.0, 2 * np.pi)
        params.append(dict(type=t, R=R, a=a, b=b, N=500, I=I, pos=pos, alpha=alpha, beta=beta))
    u = rng.random((n_points, 3))
    r = 0.02 * np.sqrt(u[:, 0])
    phi = 2 * np.pi * u[:, 1]
    z = 0.4 * u[:, 2] - 0.2
    pts = np.stack([r * np.cos(phi), r * np.sin(phi), z], axis=1)
    return params, pts


def secondary_rows(params, pf=5.0):
    """C ABI rows of coils as pair secondaries: PrecisionArguments.getSecondaryCoilPrecisionArgumentsGPU."""
    c = _c_coil()
    rows = []
    for p in params:
        k = make_coil(p, pf)
        a = c.coil.PrecisionArguments.get_secondary_coil_precision_arguments_GPU(k, c.coil.PrecisionFactor(pf))
        rows.append(coil_row(k, a))
    return rows


def matrix_pairs(n):
    """(prim_index, sec_index) of every off-diagonal entry of an n x n matrix, row-major."""
    i, j = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    off = i != j
    return i[off].astype(np.int64), j[off].astype(np.int64)


def coaxial_workload(n_pairs=4096, pf=5.0, seed=0):
    """Coaxial THIN/RECTANGULAR pairs (every pair its own coils on the z axis, z ~ U[-1, 1] m, rotations about z only)
    with the arguments Coil::computeMutualInductance plans for them (GPU method, z-axis case). Returns
    (params_prim, params_sec, prim_rows, sec_rows)."""
    c = _c_coil()
    rng = np.random.default_rng(seed)

    def draw():
        thin = rng.random() < 0.3
        return dict(R=rng.uniform(0.02, 0.5), a=0.0 if thin else rng.uniform(0.005, 0.2), b=rng.uniform(0.005, 0.2),
                    N=int(rng.integers(1, 2001)), I=float(rng.uniform(0.5, 10.0) * rng.choice([-1.0, 1.0])),
                    pos=np.array([0.0, 0.0, rng.uniform(-1.0, 1.0)]), alpha=0.0, beta=float(rng.uniform(0, 2 * np.pi)))
    pp, sp, pr, sr = [], [], [], []
    for _ in range(n_pairs):
        p, s = draw(), draw()
        cp, cs = make_coil(p, pf), make_coil(s, pf)
        a = c.coil.CoilPairArguments.get_appropriate_coil_pair_arguments(cp, cs, c.coil.PrecisionFactor(pf),
                                                                         c.coil.GPU, True, False)
        pp.append(p)
        sp.append(s)
        pr.append(coil_row(cp, a.primary_precision))
        sr.append(coil_row(cs, a.secondary_precision))
    return pp, sp, pr, sr
