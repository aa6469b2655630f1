"""GPU tests of the zero-copy entry points of the c_coil module: numpy overloads of Coil / CoilGroup
compute_all_{A_potential,B_field,E_field}_vectors and compute_all_B_gradient_matrices, and compute_all_fields_into
with numpy arrays and DLPack arrays (torch tensors on the CPU and on the GPU).

The zero-copy paths run the same ccoil_fields launch as the Vector3Array API for the same quantities, so the results
must be bit-identical to it; the Vector3Array API itself is checked against the oracle in test_gpu_dropin.py.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cc():
    import c_coil
    c_coil.coil.Coil.set_gpu_precision(c_coil.coil.GPU_FP64)
    yield c_coil
    c_coil.coil.Coil.set_gpu_precision(c_coil.coil.GPU_FP32C)


def _va(cc, pts):
    va = cc.tensor.Vector3Array()
    for p in pts:
        va.append(float(p[0]), float(p[1]), float(p[2]))
    return va


def _np3(va):
    return np.stack([np.array(va.x()), np.array(va.y()), np.array(va.z())], axis=1)


def _np9(ma):
    return np.stack([np.array(getattr(ma, k)()) for k in ("xx", "xy", "xz", "yx", "yy", "yz", "zx", "zy", "zz")],
                    axis=1)


def _sources(cc):
    import workloads as W
    C, V = cc.coil, cc.tensor.Vector3
    coil = C.Coil(0.1, 0.05, 0.08, 100, 2.0, 60.0, C.PrecisionFactor(5.0), 8, V(0.1, -0.05, 0.02), 0.4, 1.1)
    params, _ = W.group_workload(6, 8, seed=81)
    group = cc.coil_group.CoilGroup()
    for p in params:
        group.add_coil(p["R"], p["a"], p["b"], int(p["N"]), p["I"], C.PrecisionFactor(5.0), 8, V(*p["pos"]),
                       p["alpha"], p["beta"])
    return [coil, group]


def test_numpy_overloads_match_vector3array(cc):
    C = cc.coil
    rng = np.random.default_rng(82)
    pts = rng.uniform(-0.5, 0.5, size=(777, 3))
    va = _va(cc, pts)
    for src in _sources(cc):
        for name in ("compute_all_A_potential_vectors", "compute_all_B_field_vectors", "compute_all_E_field_vectors"):
            ref = _np3(getattr(src, name)(va, C.GPU))
            out = getattr(src, name)(pts, C.GPU)
            assert out.shape == (777, 3) and np.array_equal(out, ref), name
        ref = _np9(src.compute_all_B_gradient_matrices(va, C.GPU))
        out = src.compute_all_B_gradient_matrices(pts, C.GPU)
        assert out.shape == (777, 3, 3) and np.array_equal(out.reshape(-1, 9), ref)
        # float32 and Fortran-ordered inputs are converted, not rejected
        out32 = src.compute_all_B_field_vectors(np.asfortranarray(pts.astype(np.float32)), C.GPU)
        assert np.array_equal(out32, _np3(src.compute_all_B_field_vectors(_va(cc, pts.astype(np.float32)), C.GPU)))


def test_fields_into_numpy_torch_cpu_and_device(cc):
    import torch
    C = cc.coil
    rng = np.random.default_rng(83)
    n = 1000
    pts = rng.uniform(-0.5, 0.5, size=(n, 3))
    for src in _sources(cc):
        ref = {k: getattr(src, m)(pts, C.GPU) for k, m in
               (("A", "compute_all_A_potential_vectors"), ("B", "compute_all_B_field_vectors"),
                ("E", "compute_all_E_field_vectors"), ("G", "compute_all_B_gradient_matrices"))}
        out = {k: np.zeros((n, 3)) for k in "ABE"}
        out["G"] = np.zeros((n, 9))
        src.compute_all_fields_into(pts, **out)
        # one launch for all four quantities: another instantiation of the field kernel, equal to the last few ulp
        for k in "ABE":
            assert np.max(np.linalg.norm(out[k] - ref[k], axis=1) / np.linalg.norm(ref[k], axis=1)) < 1e-13, k
        g = ref["G"].reshape(n, 9)
        assert np.max(np.linalg.norm(out["G"] - g, axis=1) / np.linalg.norm(g, axis=1)) < 1e-13
        # torch tensors on the CPU (DLPack, host side)
        tp = torch.from_numpy(pts.copy())
        tb = torch.zeros(n, 3, dtype=torch.float64)
        src.compute_all_fields_into(tp, B=tb)
        assert np.array_equal(tb.numpy(), ref["B"])
        # torch tensors on the GPU (DLPack, device side): G as (n, 3, 3)
        dp = torch.from_numpy(pts).cuda()
        dB = torch.zeros(n, 3, dtype=torch.float64, device="cuda")
        dG = torch.zeros(n, 3, 3, dtype=torch.float64, device="cuda")
        src.compute_all_fields_into(dp, B=dB, G=dG)
        hB, hG = np.zeros((n, 3)), np.zeros((n, 3, 3))
        src.compute_all_fields_into(pts, B=hB, G=hG)
        assert np.array_equal(dB.cpu().numpy(), hB)
        assert np.array_equal(dG.cpu().numpy(), hG)
        # host and device buffers are not mixed
        with pytest.raises(ValueError, match="all be host or all device"):
            src.compute_all_fields_into(dp, B=np.zeros((n, 3)))


def test_device_inputs_are_ordered_after_the_producer(cc):
    """Points written by torch on its own stream just before the call are the points the kernels read: the DLPack
    handshake makes the context's stream wait for torch's pending work."""
    import torch
    C = cc.coil
    coil = _sources(cc)[0]
    n = 1 << 16
    with torch.cuda.stream(torch.cuda.Stream()):
        big = torch.randn(3072, 3072, dtype=torch.float64, device="cuda")
        for _ in range(4):
            big = big @ big.T.contiguous() / 3072.0   # keeps the stream busy ahead of the points
        pts = (torch.linspace(-0.4, 0.4, 3 * n, dtype=torch.float64, device="cuda").reshape(n, 3) +
               torch.nan_to_num(big[0, 0] * 0))
        B = torch.zeros(n, 3, dtype=torch.float64, device="cuda")
        coil.compute_all_fields_into(pts, B=B)
        got = B.cpu().numpy()
    ref = coil.compute_all_B_field_vectors(pts.cpu().numpy(), C.GPU)
    assert np.array_equal(got, ref)
