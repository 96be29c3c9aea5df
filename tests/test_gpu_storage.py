"""Host <-> device transfers (from_cpu_vec / to_cpu_vec, storage/creation.rs:3-34, storage/device.rs:23-57):
every path (HIP runtime, pinned double-buffered staging, automatic choice) returns the bytes it was given,
at sizes around the 32 MiB chunk and the 8 MiB staging threshold, and at odd offsets. Bit-exact."""
import ctypes as C

import numpy as np
import pytest

import rstsr_amd as rt
from rstsr_amd import abi

pytestmark = pytest.mark.gpu

MiB = 1 << 20


@pytest.mark.parametrize("nbytes", [1, 4097, 8 * MiB - 1, 8 * MiB, 32 * MiB, 32 * MiB + 3, 65 * MiB + 11,
                                    97 * MiB + 5])
@pytest.mark.parametrize("mode", [0, 1, 2])
def test_transfer_roundtrip(dev, nbytes, mode):
    L = abi.lib()
    ctx = C.c_void_p(dev._ctx_ptr())
    rng = np.random.default_rng(nbytes % 1000 + mode)
    src = rng.integers(0, 256, size=nbytes + 7, dtype=np.uint8)
    d = C.c_void_p()
    abi.check(L.rstsr_hip_malloc(ctx, nbytes + 64, C.byref(d)))
    try:
        # unaligned host source (offset 3) and unaligned device destination (offset 5)
        dst_dev = C.c_void_p(d.value + 5)
        abi.check(L.rstsr_hip_memcpy_h2d_mode(ctx, dst_dev, C.c_void_p(src.ctypes.data + 3), nbytes, mode))
        back = np.zeros(nbytes + 9, dtype=np.uint8)
        abi.check(L.rstsr_hip_memcpy_d2h_mode(ctx, C.c_void_p(back.ctypes.data + 1), dst_dev, nbytes, mode))
        assert np.array_equal(back[1:1 + nbytes], src[3:3 + nbytes])
        assert back[0] == 0 and np.all(back[1 + nbytes:] == 0)
    finally:
        abi.check(L.rstsr_hip_free(ctx, d))


def test_transfer_rejects_out_of_buffer(dev):
    L = abi.lib()
    ctx = C.c_void_p(dev._ctx_ptr())
    d = C.c_void_p()
    abi.check(L.rstsr_hip_malloc(ctx, 1024, C.byref(d)))
    h = np.zeros(4096, dtype=np.uint8)
    try:
        assert L.rstsr_hip_memcpy_h2d(ctx, d, h.ctypes.data_as(C.c_void_p), 2048) != 0
        assert L.rstsr_hip_memcpy_d2h_mode(ctx, h.ctypes.data_as(C.c_void_p), d, 1024, 3) != 0
    finally:
        abi.check(L.rstsr_hip_free(ctx, d))


def test_tensor_roundtrip_large(dev):
    """rt.asarray / to_numpy of a 200 MiB array go through the staged path."""
    x = np.random.default_rng(0).standard_normal(25 * MiB)
    t = rt.asarray(x, device=dev)
    assert np.array_equal(t.to_numpy(), x)
