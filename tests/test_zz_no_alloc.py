"""sg_unet_forward_noalloc: the U-Net forward pass on device buffers with every intermediate
activation in the library's fixed per-GPU scratch area (no allocation by the call), against the
allocating entry point sg_unet_forward on the same weights and inputs: the same kernels on the
same data, so the outputs must be bit-identical. Two calls in a row, the second smaller (the
scratch keeps the first call's contents), and a size beyond the scratch maximum rejected.

Device buffers for the test itself come from the HIP runtime the library is linked to (the
library's own hipMalloc, through ctypes); the entry point under test allocates nothing.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import simple_gpu as sg  # noqa: E402
import unet_ref as U  # noqa: E402

H2D, D2H = 1, 2  # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


def _hip():
    sg.gpu()  # loads libsimple_gpu.so and with it its HIP runtime
    path = None
    with open("/proc/self/maps") as f:
        for line in f:
            if "libamdhip64" in line:
                path = line.split()[-1]
                break
    assert path, "the library's HIP runtime is not mapped"
    h = C.CDLL(path)
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipFree.argtypes = [C.c_void_p]
    return h


def _inputs(seed, n, h, w):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((n, h, w, U.IN_PAD)).astype(np.float32)
    x[..., U.IN_CH:] = 0
    return x


def test_forward_noalloc_matches_forward():
    hip = _hip()
    rng = np.random.default_rng(7)
    theta = (U.init_params(3) + 0.01 * rng.standard_normal(U.n_params())).astype(np.float32)
    net = sg.UNet(theta)
    cap = 2 * 64 * 64  # pixels of the first (larger) call
    dx, dy = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(dx), cap * U.IN_PAD * 4) == 0
    assert hip.hipMalloc(C.byref(dy), cap * 4) == 0
    try:
        for n, h, w in [(2, 64, 64), (1, 32, 40)]:
            print(f"forward_noalloc {n}x{h}x{w}", flush=True)
            x = _inputs(n * h + w, n, h, w)
            ref = net.forward(x)
            assert hip.hipMemcpy(dx, x.ctypes.data, x.nbytes, H2D) == 0
            net.forward_noalloc(n, h, w, dx.value, dy.value)  # null stream
            y = np.empty((n, h, w), np.float32)
            assert hip.hipMemcpy(y.ctypes.data, dy, y.nbytes, D2H) == 0  # orders after it
            assert np.array_equal(y.view(np.uint32), ref.view(np.uint32)), \
                (n, h, w, float(np.abs(y - ref).max()))
        with pytest.raises(sg.GpuError, match="exceeds the scratch maximum of 16384"):
            net.forward_noalloc(1, 136, 128, dx.value, dy.value)
        with pytest.raises(sg.GpuError, match="multiples of 8"):
            net.forward_noalloc(1, 12, 16, dx.value, dy.value)
    finally:
        hip.hipFree(dx)
        hip.hipFree(dy)
        net.close()
