"""Build recipe for the gfx950 library and the c_coil Python module (in-tree, hipcc + g++).

Outputs (git-ignored, they travel with the source tree):
  c-coil_amd/libccoil_hip.so   HIP kernels + the C ABI of include/ccoil_hip.h
  c-coil_amd/libccoil.so       the C++ Coil / CoilGroup / vec3 API (links libccoil_hip.so)
  c-coil_amd/_c_coil*.so       pybind11 module behind c_coil.py (tensor, coil, coil_group submodules)
"""
import concurrent.futures as cf
import os
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
OBJ = os.path.join(HERE, "_build")
ARCH = os.environ.get("CCOIL_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_SOURCES = ["csrc/field_kernels.hip", "csrc/pair_kernels.hip", "csrc/adjoint_kernels.hip", "csrc/api.cpp",
               "csrc/pairs_api.cpp", "csrc/adjoint_api.cpp", "csrc/zaxis_kernels.hip", "csrc/zaxis_api.cpp", "csrc/near_kernels.hip",
               "csrc/noalloc.hip", "csrc/fields_api.cpp", "csrc/tune.cpp",
               "csrc/plan_api.cpp", "csrc/multipole_kernels.hip"]
HOST_SOURCES = ["host/Tensor.cpp", "host/Coil.cpp", "host/CoilGroup.cpp", "host/Planning.cpp", "host/GpuBridge.cpp"]
PY_SOURCES = ["python/bindings.cpp"]

# Device code: no packed FP32 math (v_pk_* are avoided on this hardware pool); hardware FP64 atomics.
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang",
             "-packed-fp32-ops", "-munsafe-fp-atomics", "-I" + os.path.join(ROOT, "include"),
             "-I" + os.path.join(HERE, "csrc")]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(HERE, "host")]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build step failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r


def _obj(src):
    return os.path.join(OBJ, src.replace("/", "_") + ".o")


def build(jobs=8, verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    steps = []
    for s in HIP_SOURCES:
        steps.append([HIPCC] + HIP_FLAGS + ["-c", os.path.join(HERE, s), "-o", _obj(s)])
    for s in HOST_SOURCES:
        steps.append(["g++"] + HOST_FLAGS + ["-c", os.path.join(HERE, s), "-o", _obj(s)])
    import pybind11
    py_inc = sysconfig.get_paths()["include"]
    for s in PY_SOURCES:
        steps.append(["g++"] + HOST_FLAGS + ["-fvisibility=hidden", "-I" + pybind11.get_include(), "-I" + py_inc,
                                             "-c", os.path.join(HERE, s), "-o", _obj(s)])
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for r in ex.map(_run, steps):
            if verbose and (r.stdout or r.stderr):
                print(r.stdout + r.stderr)
    lib_hip = os.path.join(HERE, "libccoil_hip.so")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", lib_hip] + [_obj(s) for s in HIP_SOURCES])
    lib_host = os.path.join(HERE, "libccoil.so")
    _run(["g++", "-shared", "-fPIC", "-o", lib_host] + [_obj(s) for s in HOST_SOURCES] +
         ["-L" + HERE, "-lccoil_hip", "-Wl,-rpath,$ORIGIN"])
    ext = sysconfig.get_config_var("EXT_SUFFIX")
    mod = os.path.join(HERE, "_c_coil" + ext)
    _run(["g++", "-shared", "-fPIC", "-o", mod] + [_obj(s) for s in PY_SOURCES] +
         ["-L" + HERE, "-lccoil", "-lccoil_hip", "-Wl,-rpath,$ORIGIN"])
    return lib_hip, lib_host, mod


if __name__ == "__main__":
    for p in build(verbose="-v" in sys.argv):
        print(p)
