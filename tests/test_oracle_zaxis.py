"""Pins the oracle's coaxial (z-axis) rules against independent results.

  mutual inductance, fast rule (CalculateMInductanceZAxis.cxx:84-221)
      - equals the general rule (CalculateMInductanceGeneral.cxx, its own oracle restatement, pinned by the field
        tests) for separated coaxial coils, to 1e-12;
      - the double evaluation stays within its cancellation bound of the long double one.
  axial force, fast rule (CalculateForceZAxis.cxx:82-219)
      - equals the general rule's z force for separated coaxial coils, to 1e-12;
      - equals I1 I2 dM/dz of the fast mutual inductance (central differences in long double), to 1e-8.
  self-inductance (CalculateMInductanceZAxis.cxx:224-317)
      - converges to Nagaoka's closed form for a thin solenoid, L = mu0 pi R^2 N^2 / l * K(2R/l), with
        K = 4 / (3 pi k') ((k'^2 / k^2)(K(k) - E(k)) + E(k) - k), k^2 = 4R^2 / (4R^2 + l^2);
      - equals the fast mutual-inductance rule of the coil with itself (same nodes), to 1e-12.
  slow rules (CalculateMInductanceZAxis.cxx:17-82, CalculateForceZAxis.cxx:16-80)
      - filament loops: Maxwell's closed form M = mu0 N1 N2 sqrt(R r) ((2/k - k) K(k) - 2/k E(k)),
        k^2 = 4 R r / ((R + r)^2 + d^2), to 1e-12;
      - rectangular coils: the slow and fast rules agree (two different quadratures of one integral), to 1e-6;
      - slow force = I1 I2 dM/dz of the slow mutual inductance.
"""
import numpy as np
import pytest
from scipy.special import ellipe, ellipk

import oracle as O

MU0 = 4e-7 * np.pi


def _nagaoka(R, l, N):
    k2 = 4 * R * R / (4 * R * R + l * l)
    k, kp = np.sqrt(k2), np.sqrt(1 - k2)
    K = 4 / (3 * np.pi * kp) * ((kp * kp / k2) * (ellipk(k2) - ellipe(k2)) + ellipe(k2) - k)
    return MU0 * np.pi * R * R * N * N / l * K


def _maxwell(R, r, d, N1, N2):
    k2 = 4 * R * r / ((R + r) ** 2 + d * d)
    k = np.sqrt(k2)
    return MU0 * N1 * N2 * np.sqrt(R * r) * ((2 / k - k) * ellipk(k2) - 2 / k * ellipe(k2))


PAIRS = [  # (primary R, a, b, N, I), (secondary R, a, b, N, I), secondary z, primary z
    ((0.1, 0.05, 0.1, 100, 2.0), (0.08, 0.04, 0.06, 50, 1.5), 0.3, 0.0),
    ((0.1, 0.05, 0.1, 100, 2.0), (0.08, 0.04, 0.06, 50, 1.5), 1.0, 0.2),
    ((0.05, 0.0, 0.2, 300, 1.0), (0.12, 0.03, 0.0001, 20, -2.0), -0.4, 0.0),   # THIN primary, short secondary
    ((0.2, 0.1, 0.02, 40, 3.0), (0.2, 0.1, 0.02, 40, 3.0), 0.05, 0.0),          # equal discs, close
]


@pytest.mark.parametrize("case", range(len(PAIRS)))
def test_fast_rules_match_general_rule(case):
    pp, sp, zs, zp = PAIRS[case]
    p = O.make_coil(*pp, pos=(0, 0, zp))
    s = O.make_coil(*sp, pos=(0, 0, zs))
    za = O.pair_precision_args(p, s, 7.0, False, True, False)
    ga = O.pair_precision_args(p, s, 7.0, False, False, False)
    m_ld = O.mutual_inductance_zaxis_fast(p, s, za, extended=True)
    f_ld = O.force_zaxis_fast(p, s, za, extended=True)
    assert abs(O.mutual_inductance_general(p, s, ga) - m_ld) <= 1e-12 * abs(m_ld)
    F, T = O.force_torque_general(p, s, ga)
    assert abs(F[2] - f_ld) <= 1e-12 * abs(f_ld)
    assert np.all(np.abs(F[:2]) <= 1e-12 * abs(f_ld)) and np.all(np.abs(T) <= 1e-12 * abs(f_ld) * 1.0)
    # the reference's own double arithmetic: within 1e-10 of the long double evaluation of the same rule (the
    # four length terms cancel: a 1e-4 m long secondary next to a 0.2 m primary loses 3e-11)
    assert abs(O.mutual_inductance_zaxis_fast(p, s, za) - m_ld) <= 1e-10 * abs(m_ld)
    assert abs(O.force_zaxis_fast(p, s, za) - f_ld) <= 1e-10 * abs(f_ld)


@pytest.mark.parametrize("case", range(len(PAIRS)))
def test_fast_force_is_current_times_dM_dz(case):
    pp, sp, zs, zp = PAIRS[case]
    p = O.make_coil(*pp, pos=(0, 0, zp))
    s = O.make_coil(*sp, pos=(0, 0, zs))
    za = O.pair_precision_args(p, s, 5.0, False, True, False)
    h = 1e-5 * max(abs(zs - zp), 0.05)
    mz = lambda z: O.mutual_inductance_zaxis_fast(p, O.make_coil(*sp, pos=(0, 0, z)), za, extended=True)
    fd = (mz(zs + h) - mz(zs - h)) / (2 * h) * pp[4] * sp[4]
    f = O.force_zaxis_fast(p, s, za, extended=True)
    assert abs(f - fd) <= 1e-8 * abs(f)


@pytest.mark.parametrize("R,l", [(0.1, 0.2), (0.05, 0.4), (0.1, 0.05)])
def test_self_inductance_thin_solenoid_nagaoka(R, l):
    c = O.make_coil(R, 0.0, l, 100, 1.0)
    assert c.type == O.THIN
    ref = _nagaoka(R, l, 100)
    errs = []
    for pf in (5.0, 9.0):
        a = O.pair_precision_args(c, c, pf, False, True, False)
        errs.append(abs(O.self_inductance(c, a) / ref - 1))
    # algebraic convergence of the log-singular kernel: 3e-4 at precision factor 5, 2e-5 at 9
    assert errs[0] < 3e-4 and errs[1] < 2e-5 and errs[1] < errs[0] / 4


@pytest.mark.parametrize("dims", [(0.1, 0.05, 0.1, 100, 2.0), (0.03, 0.02, 0.3, 500, 0.5)])
def test_self_inductance_is_fast_mutual_with_itself(dims):
    c = O.make_coil(*dims)
    a = O.pair_precision_args(c, c, 6.0, False, True, False)
    ld = O.self_inductance(c, a, extended=True)
    assert abs(O.mutual_inductance_zaxis_fast(c, c, a, extended=True) - ld) <= 1e-12 * ld
    assert abs(O.self_inductance(c, a) - ld) <= 1e-12 * ld
    assert ld > 0


@pytest.mark.parametrize("R1,R2,d", [(0.1, 0.07, 0.05), (0.2, 0.2, 0.01), (0.05, 0.3, -0.4)])
def test_slow_rule_filaments_maxwell(R1, R2, d):
    p = O.make_coil(R1, 0.0, 0.0, 3, 2.0)
    s = O.make_coil(R2, 0.0, 0.0, 5, 1.0, pos=(0, 0, d))
    a = O.pair_precision_args(p, s, 5.0, True, True, False)
    fa = np.array([1, 1, 1, a[3], a[4], a[5]], dtype=np.int32)
    m = O.mutual_inductance_zaxis_slow(p, s, a, fa, extended=True)
    assert abs(m / _maxwell(R1, R2, d, 3, 5) - 1) < 1e-12


def test_slow_and_fast_rules_agree_and_slow_force_is_dM_dz():
    pp, sp = (0.1, 0.05, 0.1, 100, 2.0), (0.08, 0.04, 0.06, 50, 1.5)
    p = O.make_coil(*pp)
    s = O.make_coil(*sp, pos=(0, 0, 0.3))
    a = O.pair_precision_args(p, s, 9.0, True, True, False)
    fa = np.array([1, 1, 1, a[3], a[4], a[5]], dtype=np.int32)
    fast = O.mutual_inductance_zaxis_fast(p, s, a, extended=True)
    slow = O.mutual_inductance_zaxis_slow(p, s, a, fa)
    assert abs(slow / fast - 1) < 1e-6
    h = 1e-5
    ms = lambda z: O.mutual_inductance_zaxis_slow(p, O.make_coil(*sp, pos=(0, 0, z)), a, fa, extended=True)
    fd = (ms(0.3 + h) - ms(0.3 - h)) / (2 * h) * pp[4] * sp[4]
    f = O.force_zaxis_slow(p, s, a, fa, extended=True)
    assert abs(f - fd) <= 1e-6 * abs(f)
    assert abs(f / O.force_zaxis_fast(p, s, a, extended=True) - 1) < 1e-5


def test_mirror_symmetry_and_attraction():
    """Reflecting the secondary through the primary's mid-plane keeps M and reverses the force; co-directed
    currents attract. A sign error in the displacement or the force fails here."""
    pp, sp, zs, _ = PAIRS[0]
    p = O.make_coil(*pp)
    s = O.make_coil(*sp, pos=(0, 0, zs))
    za = O.pair_precision_args(p, s, 7.0, False, True, False)
    m = O.mutual_inductance_zaxis_fast(p, s, za, extended=True)
    f = O.force_zaxis_fast(p, s, za, extended=True)
    # mirror image: same M, opposite force (so a sign error in the force is visible)
    s_m = O.make_coil(*sp, pos=(0, 0, -zs))
    assert abs(O.mutual_inductance_zaxis_fast(p, s_m, za, extended=True) - m) <= 1e-12 * m
    assert abs(O.force_zaxis_fast(p, s_m, za, extended=True) + f) <= 1e-12 * abs(f)
    # attraction of co-directed currents: the force on a secondary above the primary points down
    assert f < 0 < m


def test_self_inductance_brooks_coil():
    """The Brooks coil (inner radius = thickness = length = c) has L = 2.5491e-6 c N^2 H (Grover's tables, five
    significant figures): an independent pin of the thick-coil rule, which the Nagaoka test does not reach."""
    c = O.make_coil(0.1, 0.1, 0.1, 1, 1.0)
    for pf in (7.0, 9.0):
        a = O.pair_precision_args(c, c, pf, False, True, False)
        assert abs(O.self_inductance(c, a) / (2.5491e-6 * 0.1) - 1) < 3e-5


@pytest.mark.parametrize("R1,R2,d", [(0.5, 0.3, 0.4), (0.1, 0.07, 0.05), (0.2, 0.25, -0.1)])
def test_slow_force_filaments_maxwell_derivative(R1, R2, d):
    """Axial force between coaxial loops = I1 I2 dM/dd of Maxwell's closed form (central differences of the
    elliptic-integral formula), the reference's own filament comparison (CompForce.cxx:8-20) at other sizes."""
    I1, I2 = 100.0, 200.0
    p = O.make_coil(R1, 0.0, 0.0, 1, I1)
    s = O.make_coil(R2, 0.0, 0.0, 1, I2, pos=(0, 0, d))
    a = O.pair_precision_args(p, s, 5.0, True, True, False)
    fa = np.array([1, 1, 1, a[3], a[4], a[5]], dtype=np.int32)
    h = 1e-6 * abs(d)
    ref = (_maxwell(R1, R2, d + h, 1, 1) - _maxwell(R1, R2, d - h, 1, 1)) / (2 * h) * I1 * I2
    assert abs(O.force_zaxis_slow(p, s, a, fa, extended=True) / ref - 1) < 1e-8
