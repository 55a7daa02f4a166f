"""The drop-in boundary as compiled code (CPU side): the reference's own
rad_utils.c built with GPUAXL links emma_amd/dropin/rad_utils_hip.c, which
defines the reference's GPU entry points (rad_utils_gpu.h:1-6, interface.h:
16-18) over libemma_rt.so.  Here: every such build loads and exports those
symbols; the GPU tests run them (test_gpu_dropin.py)."""
import ctypes
import os

import pytest

from oracle import REF_SO_GPU, REF_SO_GPU_CR

SYMBOLS = ["advanceradGPU", "create_radstencil_GPU", "create_pinned_stencil_rad", "destroy_radstencil_GPU",
           "destroy_pinned_stencil_rad", "create_param_GPU", "countdevices", "initlocaldevice", "checkdevice"]

BUILDS = sorted(set(REF_SO_GPU.values()) | set(REF_SO_GPU_CR.values()))


@pytest.mark.parametrize("path", BUILDS, ids=[os.path.basename(p) for p in BUILDS])
def test_gpuaxl_build_exports_reference_symbols(path):
    if not os.path.exists(path):
        pytest.skip("GPUAXL reference build (oracle/_ref) not present")
    lib = ctypes.CDLL(path)  # resolves libemma_rt.so through its rpath; no call is made
    for name in SYMBOLS:
        assert hasattr(lib, name), name
    assert lib.refdrv_gpuaxl() == 1
    assert lib.refdrv_coarserad() == (1 if path in REF_SO_GPU_CR.values() else 0)
