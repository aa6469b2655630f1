"""Pins the oracle's pair integrals (CalculateMInductanceGeneral.cxx, CalculateForceGeneral.cxx and the GPU
pair rule of MInductance.cu / Force.cu) without trusting its formulas:
  - Maxwell's closed form for coaxial filament loops (complete elliptic integrals);
  - F = I1 I2 grad M and torque = I1 I2 dM/d(angle) by central finite differences of the oracle's own M
    (catches a wrong sign, a dropped term or a swapped index in the force kernel);
  - reciprocity M12 = M21 on the reference's own pair geometries (tests/golden);
  - the GPU pair rule converges to the general rule.
"""
import os

import numpy as np
import pytest
from scipy.special import ellipe, ellipk

import oracle as O

MU0 = 4e-7 * np.pi
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def maxwell(R1, R2, d):
    m = 4 * R1 * R2 / ((R1 + R2) ** 2 + d ** 2)
    k = np.sqrt(m)
    return MU0 * np.sqrt(R1 * R2) * ((2 / k - k) * ellipk(m) - 2 / k * ellipe(m))


@pytest.mark.parametrize("R1,R2,d", [(0.1, 0.3, 0.2), (0.05, 0.05, 0.01), (0.2, 0.1, 0.5), (0.03, 0.03, 0.2)])
def test_coaxial_filaments_maxwell(R1, R2, d):
    p = O.make_coil(R1, 0.0, 0.0, 1)
    s = O.make_coil(R2, 0.0, 0.0, 1, pos=(0.0, 0.0, d))
    args = O.pair_precision_args(p, s, 10.0, False, False, False)
    assert O.mutual_inductance_general(p, s, args) == pytest.approx(maxwell(R1, R2, d), rel=1e-12)


def _pose(pos, al, be, I1=2.0, I2=3.0):
    p = O.make_coil(0.1, 0.05, 0.08, 100, I1)
    s = O.make_coil(0.07, 0.03, 0.04, 50, I2, pos=tuple(pos), alpha=al, beta=be)
    return p, s


def test_force_is_current_product_times_gradient_of_m():
    pos, al, be = np.array([0.05, 0.02, 0.15]), 0.3, 0.7
    p, s = _pose(pos, al, be)
    args = O.pair_precision_args(p, s, 9.0, False, False, False)
    F, T = O.force_torque_general(p, s, args)
    h = 1e-5
    grad = np.zeros(3)
    for k in range(3):
        e = np.zeros(3)
        e[k] = h
        grad[k] = (O.mutual_inductance_general(*_pose(pos + e, al, be), args) -
                   O.mutual_inductance_general(*_pose(pos - e, al, be), args)) / (2 * h)
    assert np.linalg.norm(F - 6.0 * grad) <= 1e-7 * np.linalg.norm(F)
    # torque about the secondary centre: rotation about z changes beta, about u = (-sin b, cos b, 0) alpha
    h = 1e-6
    dmb = (O.mutual_inductance_general(*_pose(pos, al, be + h), args) -
           O.mutual_inductance_general(*_pose(pos, al, be - h), args)) / (2 * h)
    dma = (O.mutual_inductance_general(*_pose(pos, al + h, be), args) -
           O.mutual_inductance_general(*_pose(pos, al - h, be), args)) / (2 * h)
    u = np.array([-np.sin(be), np.cos(be), 0.0])
    assert T[2] == pytest.approx(6.0 * dmb, rel=1e-6)
    assert T @ u == pytest.approx(6.0 * dma, rel=1e-6)


def _golden(name):
    return np.loadtxt(os.path.join(GOLDEN, name))


def test_reciprocity_on_reference_z_axis_geometries():
    rows = _golden("values_MInductance_zAxis.txt")
    assert rows.shape[1] == 9 and len(rows) >= 10
    for R1, a1, b1, N1, R2, a2, b2, N2, d in rows:
        p = O.make_coil(R1, a1, b1, int(N1))
        s = O.make_coil(R2, a2, b2, int(N2), pos=(0.0, 0.0, d))
        s0 = O.make_coil(R2, a2, b2, int(N2))
        p0 = O.make_coil(R1, a1, b1, int(N1), pos=(0.0, 0.0, -d))
        m12 = O.mutual_inductance_general(p, s, O.pair_precision_args(p, s, 5.0, False, False, False))
        m21 = O.mutual_inductance_general(s0, p0, O.pair_precision_args(s0, p0, 5.0, False, False, False))
        assert np.isfinite(m12) and m12 != 0.0
        assert m12 == pytest.approx(m21, rel=1e-6), (R1, R2, d)


def test_general_geometries_converge():
    rows = _golden("values_MInductance_general.txt")
    assert rows.shape[1] == 10 and len(rows) >= 10
    for R1, a1, b1, N1, R2, a2, b2, N2, dz, dr in rows[:12]:
        p = O.make_coil(R1, a1, b1, int(N1))
        s = O.make_coil(R2, a2, b2, int(N2), pos=(dr, 0.0, dz))
        m5 = O.mutual_inductance_general(p, s, O.pair_precision_args(p, s, 5.0, False, False, False))
        m9 = O.mutual_inductance_general(p, s, O.pair_precision_args(p, s, 9.0, False, False, False))
        assert m5 == pytest.approx(m9, rel=1e-4, abs=1e-15), (R1, R2, dz, dr)


def test_gpu_pair_rule_converges_to_general_rule():
    p, s = _pose((0.05, 0.02, 0.25), 0.3, 0.7)
    general = O.mutual_inductance_general(p, s, O.pair_precision_args(p, s, 12.0, False, False, False))
    gpu = O.mutual_inductance_gpu_rule(p, s, (80, 40, 1), (60, 20, 20))
    assert gpu == pytest.approx(general, rel=1e-8)
    Fg, Tg = O.force_torque_general(p, s, O.pair_precision_args(p, s, 12.0, False, False, False))
    F, T = O.force_torque_gpu_rule(p, s, (80, 40, 1), (60, 20, 20), 2 * np.pi * 50 * 3.0)
    assert np.linalg.norm(F - Fg) <= 1e-7 * np.linalg.norm(Fg)
    assert np.linalg.norm(T - Tg) <= 1e-7 * np.linalg.norm(Tg)
