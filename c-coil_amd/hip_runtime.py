"""One HIP runtime per process.

PyTorch-ROCm wheels ship their own libamdhip64 / libhsa-runtime64 under torch/lib. If the system runtime
(/opt/rocm, which libccoil_hip.so is linked against) is loaded first and torch second, the process holds two HSA
runtimes and the second one sees no GPU. Loading torch's runtime first (it carries the same soname) makes
libccoil_hip.so bind to it, so ccoil and torch share one runtime, one device context and each other's streams and
device pointers. Without torch installed the system runtime is used.
"""
import ctypes
import importlib.util
import os

_done = False


def preload():
    global _done
    if _done:
        return
    _done = True
    spec = importlib.util.find_spec("torch")
    if spec is None or not spec.origin:
        return
    lib = os.path.join(os.path.dirname(spec.origin), "lib", "libamdhip64.so")
    if os.path.exists(lib):
        ctypes.CDLL(lib, mode=ctypes.RTLD_GLOBAL)
