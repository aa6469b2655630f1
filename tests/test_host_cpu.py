"""CPU tests of the host side: planning (PrecisionArguments / CoilPairArguments), coil attributes and
transformation matrices, the c_coil API's error behaviour, and the C ABI library's exports.

Planning anchors (exact integers, from the reference's planning of two coils at PrecisionFactor 5):
  RECTANGULAR R = a = b = 0.1:  CPU (1,1,1) 59/27/27, GPU 59/27/27
  THIN R = 0.1, b = 0.1:        CPU (2,1,1) 81/1/90,  GPU (1,1,1) 80/1/45
The product's planning (c_coil, C++) and the oracle's (C) are separate restatements of
PrecisionArguments.cxx:69-288 and CoilPairArguments.cxx:20-255; they must agree on every case of a seeded grid.
"""
import ctypes
import os
import re

import numpy as np
import pytest

import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def cc():
    import c_coil
    return c_coil


def _args6(a):
    return [a.angular_block_count, a.thickness_block_count, a.length_block_count, a.angular_increment_count,
            a.thickness_increment_count, a.length_increment_count]


@pytest.mark.parametrize("geom,cpu,gpu", [((0.1, 0.1, 0.1), [1, 1, 1, 59, 27, 27], [1, 1, 1, 59, 27, 27]),
                                          ((0.1, 0.0, 0.1), [2, 1, 1, 81, 1, 90], [1, 1, 1, 80, 1, 45])])
def test_planning_anchors(cc, geom, cpu, gpu):
    C = cc.coil
    c = C.Coil(*geom, 100)
    pf = C.PrecisionFactor(5.0)
    assert _args6(C.PrecisionArguments.get_coil_precision_arguments_CPU(c, pf)) == cpu
    assert _args6(C.PrecisionArguments.get_coil_precision_arguments_GPU(c, pf)) == gpu
    oc = O.make_coil(*geom, 100)
    assert list(O.coil_precision_args(oc, 5.0, False)) == cpu
    assert list(O.coil_precision_args(oc, 5.0, True)) == gpu


def _grid(rng, n):
    for _ in range(n):
        R = float(10 ** rng.uniform(-3, 0.5))
        kind = rng.integers(4)
        a = 0.0 if kind in (0, 1) else float(R * 10 ** rng.uniform(-3, 1))
        b = 0.0 if kind in (0, 2) else float(R * 10 ** rng.uniform(-3, 1.5))
        yield R, a, b, int(rng.integers(1, 5000))


def test_planning_product_matches_oracle(cc):
    C = cc.coil
    rng = np.random.default_rng(0)
    bad = 0
    for R, a, b, N in _grid(rng, 400):
        c, oc = C.Coil(R, a, b, N), O.make_coil(R, a, b, N)
        for pf in (1.0, 2.5, 5.0, 9.0, 15.0):
            P = C.PrecisionFactor(pf)
            bad += _args6(C.PrecisionArguments.get_coil_precision_arguments_CPU(c, P)) != \
                list(O.coil_precision_args(oc, pf, False))
            bad += _args6(C.PrecisionArguments.get_coil_precision_arguments_GPU(c, P)) != \
                list(O.coil_precision_args(oc, pf, True))
            bad += _args6(C.PrecisionArguments.get_secondary_coil_precision_arguments_GPU(c, P)) != \
                list(O.secondary_precision_args_gpu(oc, pf))
    assert bad == 0


def test_planning_invariants():
    """Properties of the planning that hold whatever the details of the balancing: GPU arguments never exceed the
    cap and use one block per dimension; dimensions a coil type does not have get one increment; more precision
    never means fewer GPU nodes; and the measures sqrt(2 pi (R + a/2)), sqrt(2 a), sqrt(2 b) are homogeneous of one
    degree, so scaling a coil by 4 (exact in floating point) leaves every argument unchanged. CPU arguments are
    not monotone: the validating constructor (PrecisionArguments.cxx:36-43) replaces an increment count above 100
    (a layer of 200 split over 2 blocks gives 101) by the default 20; a FLAT coil at PrecisionFactor 9 shows it."""
    rng = np.random.default_rng(7)
    bad = []
    for R, a, b, N in _grid(rng, 300):
        oc = O.make_coil(R, a, b, N)
        big = O.make_coil(4 * R, 4 * a, 4 * b, N)
        prev_gpu = 0
        for pf in (1.0, 2.0, 3.5, 5.0, 7.0, 9.0, 12.0, 15.0):
            cpu = list(O.coil_precision_args(oc, pf, False))
            gpu = list(O.coil_precision_args(oc, pf, True))
            sec = list(O.secondary_precision_args_gpu(oc, pf))
            if gpu[:3] != [1, 1, 1] or max(gpu[3:]) > 80 or sec[:3] != [1, 1, 1] or max(sec[3:]) > 80:
                bad.append(("cap", R, a, b, pf, gpu, sec))
            if a == 0.0 and (cpu[1] * cpu[4] != 1 or gpu[4] != 1 or sec[4] != 1):
                bad.append(("no thickness", R, a, b, pf, cpu, gpu, sec))
            if b == 0.0 and sec[5] != 1:
                bad.append(("no length", R, a, b, pf, sec))
            if min(cpu[:3]) < 1 or min(cpu[3:]) < 1 or max(cpu[3:]) > 100:
                bad.append(("range", R, a, b, pf, cpu))
            n_gpu = int(np.prod(gpu[3:]))
            if n_gpu < prev_gpu:
                bad.append(("monotone", R, a, b, pf, gpu))
            prev_gpu = n_gpu
            if (list(O.coil_precision_args(big, pf, False)) != cpu or list(O.coil_precision_args(big, pf, True)) != gpu
                    or list(O.secondary_precision_args_gpu(big, pf)) != sec):
                bad.append(("scale", R, a, b, pf))
    assert not bad, bad[:5]
    # by hand at PrecisionFactor 9: 25,600 increments, measures 0.70078 (angular) and 0.45028 (thickness), linear
    # measure sqrt(0.70078 * 0.45028 / 25600) = 3.5109e-3, layers round(199.6) = 200 and round(128.25) = 128, so
    # 2 blocks of 200 / 2 + 1 = 101 -> 20 and 2 blocks of 128 / 2 + 1 = 65 increments
    flat = O.make_coil(0.027467899781113806, 0.10137763355974024, 0.0, 100)
    assert list(O.coil_precision_args(flat, 8.0, False)) == [2, 1, 1, 71, 91, 1]
    assert list(O.coil_precision_args(flat, 9.0, False)) == [2, 2, 1, 20, 65, 1]
    # GPU spill, by hand: RECTANGULAR R = 0.5, a = b = 0.01 at PF 5: 1,600 increments, measures 1.7813 and 0.14142,
    # angular layer round(141.96) = 142 > 80 -> 80, the rest ceil(1600 / 80) = 20 to the thickness, and the length
    # layer round(80 * 0.14142 / 1.7813) = round(6.35) = 6
    assert list(O.coil_precision_args(O.make_coil(0.5, 0.01, 0.01, 100), 5.0, True)) == [1, 1, 1, 80, 20, 6]
    # secondary GPU rule, by hand: THIN R = b = 0.1 at PF 5 balances angle and length: 1,600 increments, measures
    # 0.79267 and 0.44721, linear measure 0.014885, layers round(53.25) = 53 and round(30.04) = 30
    assert list(O.secondary_precision_args_gpu(O.make_coil(0.1, 0.0, 0.1, 100), 5.0)) == [1, 1, 1, 53, 1, 30]


def test_pair_planning_product_matches_oracle(cc):
    C = cc.coil
    rng = np.random.default_rng(1)
    geoms = list(_grid(rng, 120))
    bad = 0
    for k in range(0, len(geoms), 2):
        g1, g2 = geoms[k], geoms[k + 1]
        c1, c2 = C.Coil(*g1), C.Coil(*g2)
        o1, o2 = O.make_coil(*g1), O.make_coil(*g2)
        for pf in (1.0, 4.0, 7.5, 12.0):
            for method, gpu in ((C.CPU_ST, False), (C.GPU, True)):
                for z_axis in (False, True):
                    for pure in ((False, True) if gpu else (False,)):
                        a = C.CoilPairArguments.get_appropriate_coil_pair_arguments(c1, c2, C.PrecisionFactor(pf),
                                                                                   method, z_axis, pure)
                        got = _args6(a.primary_precision) + _args6(a.secondary_precision)
                        ref = list(O.pair_precision_args(o1, o2, pf, gpu, z_axis, pure))
                        bad += got != ref
    assert bad == 0


def test_precision_factor_clamp(cc):
    C = cc.coil
    for pf, want in ((0.5, 5.0), (1.0, 1.0), (7.25, 7.25), (15.0, 15.0), (15.5, 5.0)):
        assert C.PrecisionFactor(pf).relative_precision == want
        assert O.precision_factor(pf) == want


def test_coil_attributes(cc):
    C = cc.coil
    c = C.Coil(0.1, 0.05, 0.2, 300, 2.0)
    assert c.get_coil_type() == C.RECTANGULAR and c.is_using_fast_method()
    assert c.get_current_density() == pytest.approx(300 * 2.0 / (0.05 * 0.2), rel=1e-15)
    f = C.Coil(0.1, 0.0, 0.0, 1)
    assert f.get_coil_type() == C.FILAMENT and not f.is_using_fast_method()
    assert 0 < f.get_thickness() < 1e-20 and 0 < f.get_length() < 1e-20   # thin clamp (Coil.cxx:27-28)
    assert C.Coil(0.1, 0.0, 0.1, 10).get_coil_type() == C.THIN
    assert C.Coil(0.1, 0.1, 0.0, 10).get_coil_type() == C.FLAT


def test_transformation_matrix_closed_form():
    """T (CalculateAttributes.cxx:65-78) equals Rz(beta) Ry(alpha) Rz(beta), is orthonormal with det 1, maps e_z to
    the coil axis (cos b sin a, sin b sin a, cos a), and the stored inverse is its transpose."""
    rng = np.random.default_rng(3)
    for _ in range(50):
        al, be = rng.uniform(0, np.pi), rng.uniform(0, 2 * np.pi)
        c = O.make_coil(0.1, 0.1, 0.1, 10, alpha=al, beta=be)
        T = np.array(c.T).reshape(3, 3)
        Ti = np.array(c.Tinv).reshape(3, 3)
        rz = np.array([[np.cos(be), -np.sin(be), 0], [np.sin(be), np.cos(be), 0], [0, 0, 1]])
        ry = np.array([[np.cos(al), 0, np.sin(al)], [0, 1, 0], [-np.sin(al), 0, np.cos(al)]])
        assert np.allclose(T, rz @ ry @ rz, atol=1e-15)
        assert np.allclose(T @ T.T, np.eye(3), atol=1e-15)
        assert np.linalg.det(T) == pytest.approx(1.0, abs=1e-14)
        assert np.allclose(Ti, T.T, atol=0)
        assert np.allclose(T[:, 2], [np.cos(be) * np.sin(al), np.sin(be) * np.sin(al), np.cos(al)], atol=1e-15)


def test_cpu_methods_and_missing_device_raise(cc):
    C = cc.coil
    V = cc.tensor.Vector3
    c = C.Coil(0.1, 0.1, 0.1, 100)
    with pytest.raises(RuntimeError, match="CPU"):
        c.compute_B_field_vector(V(0.0, 0.0, 0.2))
    pts = cc.tensor.Vector3Array()
    pts.append(0.0, 0.0, 0.2)
    with pytest.raises(RuntimeError, match="CPU"):
        c.compute_all_B_field_vectors(pts, C.CPU_MT)
    import torch
    if torch.cuda.is_available():
        pytest.skip("a HIP device is present")
    with pytest.raises(RuntimeError, match="GPU functions are unavailable"):
        c.compute_all_B_field_vectors(pts, C.GPU)


def test_group_member_errors(cc):
    G = cc.coil_group.CoilGroup()
    G.add_coil(0.1, 0.1, 0.1, 100)
    with pytest.raises((IndexError, ValueError, RuntimeError)):
        G.remove_coil(5)
    assert len(G.get_member_coils()) == 1


def test_abi_library_exports_header_symbols():
    """libccoil_hip.so loads without a device and exports every function include/ccoil_hip.h declares."""
    import ccoil_abi
    with open(os.path.join(ROOT, "include", "ccoil_hip.h")) as f:
        header = f.read()
    declared = sorted(set(re.findall(r"^\s*(?:ccoil_status|const char \*)\s*(ccoil_\w+)\s*\(", header, re.M)))
    assert declared == sorted(ccoil_abi.EXPORTED)
    L = ctypes.CDLL(ccoil_abi.LIB_PATH)
    for name in declared:
        assert hasattr(L, name), name
    assert ccoil_abi.version()


def test_abi_without_device_reports_no_device():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a HIP device is present")
    import ccoil_abi
    with pytest.raises(ccoil_abi.CCoilError) as e:
        ccoil_abi.Context(0)
    assert e.value.status == ccoil_abi.ERR_NO_DEVICE
    L = ccoil_abi.lib()
    assert L.ccoil_fields(None, None, None, 0, 0, 0, 0, None, None, None, None, None) == ccoil_abi.ERR_ARGUMENT
    assert L.ccoil_pairs(None, None, None, None, None, 0, None, None, 0, 0, 0, None, None, None, None, None) == \
        ccoil_abi.ERR_ARGUMENT
    assert L.ccoil_field_gradient(None, None, None, 0, None, 0, 0, None, None) == ccoil_abi.ERR_ARGUMENT


def test_zero_copy_argument_checks(cc):
    """The numpy / DLPack entry points check shapes, dtypes, writability and sides before any device work."""
    import torch
    C = cc.coil
    for src in (C.Coil(0.1, 0.05, 0.08, 100), cc.coil_group.CoilGroup([C.Coil(0.1, 0.05, 0.08, 100)])):
        with pytest.raises(ValueError, match=r"shape \(n, 3\)"):
            src.compute_all_fields_into(np.zeros((4, 2)), B=np.zeros((4, 3)))
        with pytest.raises(ValueError, match="as many rows"):
            src.compute_all_fields_into(np.zeros((4, 3)), B=np.zeros((5, 3)))
        with pytest.raises(ValueError, match="no output"):
            src.compute_all_fields_into(np.zeros((4, 3)))
        with pytest.raises(ValueError, match=r"\(n, 9\) or \(n, 3, 3\)"):
            src.compute_all_fields_into(np.zeros((4, 3)), G=np.zeros((4, 3)))
        with pytest.raises(ValueError, match="writable C-contiguous float64"):
            src.compute_all_fields_into(np.zeros((4, 3)), B=np.zeros((4, 3), dtype=np.float32))
        with pytest.raises(ValueError, match="writable C-contiguous float64"):
            src.compute_all_fields_into(np.zeros((4, 3)), B=np.zeros((3, 4)).T)
        with pytest.raises(ValueError, match="numpy array or a DLPack array"):
            src.compute_all_fields_into([[0.0, 0.0, 0.2]], B=np.zeros((1, 3)))
        with pytest.raises(ValueError, match="float64"):
            src.compute_all_fields_into(torch.zeros(4, 3, dtype=torch.float32), B=np.zeros((4, 3)))
        with pytest.raises(ValueError, match="contiguous"):
            src.compute_all_fields_into(torch.zeros(3, 4, dtype=torch.float64).T, B=np.zeros((4, 3)))
        with pytest.raises(RuntimeError, match="CPU"):
            src.compute_all_B_field_vectors(np.zeros((4, 3)))
