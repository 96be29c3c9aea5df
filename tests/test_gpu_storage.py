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


def test_display_of_device_tensors(dev):
    """str(tensor) copies the covered storage to the host and prints the reference's layout walk."""
    import json
    import os

    ref = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_cases.json")))["display"]
    a = rt.arange(8, device=dev).reshape([2, 2, 2])
    assert str(a) == ref["cases"][0]["out"]
    assert str(a[::-1]) == ref["cases"][1]["out"]          # flip(0)
    assert str(a[:, ::-1]) == ref["cases"][2]["out"]       # flip(1)
    assert str(rt.asarray(np.array([-1, 2, -3]), device=dev).unary("abs")) == "[ 1 2 3]"
    assert str(rt.asarray(np.array([0.0, 1.0, 4.0]), device=dev).unary("sqrt")) == "[ 0 1 2]"
    x = rt.asarray(np.array([0.0, np.nan, np.inf]), device=dev)
    assert str(x.unary("is_nan")) == "[ false true false]"
    big = rt.arange(1800, device=dev).reshape([10, 12, 15]).transpose([2, 0, 1])
    from rstsr_amd import format_tensor
    host = np.arange(1800).reshape(10, 12, 15).transpose(2, 0, 1)
    assert str(big) == format_tensor(np.arange(1800), [15, 10, 12], [1, 180, 15], 0)
    assert "..." in str(big) and str(big).startswith("[[[ 0 15 30")
    assert host[0, 0, 1] == 15


def test_to_device_keeps_layout(dev):
    dev2 = rt.DeviceHIP(0)
    x = np.arange(600.0).reshape(20, 30)
    t = rt.asarray(x, device=dev)[2:18:3, ::-2]
    u = t.to_device(dev2)
    assert u.shape == t.shape and u.stride == t.stride
    assert np.array_equal(u.to_numpy(), x[2:18:3, ::-2])
    # the copy is independent of the source
    t.fill(0.0)
    assert np.array_equal(u.to_numpy(), x[2:18:3, ::-2])
    z = rt.empty((0, 4), device=dev).to_device(dev2)
    assert z.shape == (0, 4) and z.size == 0
