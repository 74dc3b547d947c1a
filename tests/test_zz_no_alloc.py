"""The allocation-free gradient entry point (wnlqg_mlp_gradient_noalloc, scratch in a static device array)
against wnlqg_mlp_gradient on the same samples: two calls in a row, the second with fewer samples, so that the
second call runs on scratch the first one left behind.  The test's device buffers come from hipMalloc of the
HIP runtime the library itself links (plumbing of the test, not of the call under test)."""
import ctypes
import sys

import numpy as np
import pytest

from wanniernlqg_amd.engine import load_library
from wanniernlqg_amd.mlp import INPUTS, PARAMS, ProposalNetwork

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2   # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


def hip_runtime():
    load_library()                                   # loads the library's libamdhip64.so.7 into the process
    hip = ctypes.CDLL("libamdhip64.so.7")            # the same, already-loaded runtime
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    for f in (hip.hipMalloc, hip.hipFree, hip.hipMemcpy, hip.hipDeviceSynchronize):
        f.restype = ctypes.c_int
    return hip


def device_copy(hip, a: np.ndarray) -> ctypes.c_void_p:
    p = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(p), a.nbytes) == 0
    assert hip.hipMemcpy(p, a.ctypes.data_as(ctypes.c_void_p), a.nbytes, H2D) == 0
    return p


def test_noalloc_gradient_equals_gradient_on_two_calls():
    print("[no-alloc] gradient vs allocation-free gradient, n = 50000 then 3001", flush=True)
    sys.stdout.flush()
    hip = hip_runtime()
    net = ProposalNetwork(seed=11)
    rng = np.random.default_rng(4)
    for n in (50000, 3001):
        x = rng.uniform(-1, 1, (n, INPUTS)).astype(np.float32)
        t = (np.sin(2 * x[:, 0]) + x[:, 1] * x[:, 2] - 0.5 * x[:, 7]).astype(np.float32)
        g_ref, loss_ref = net.gradient(x, t)
        poisoned = np.full(PARAMS + 1, np.nan, dtype=np.float32)
        xd, td, gd = device_copy(hip, x), device_copy(hip, t), device_copy(hip, poisoned)
        try:
            net.gradient_noalloc(xd.value, td.value, n, gd.value)
            got = np.empty(PARAMS + 1, dtype=np.float32)
            assert hip.hipMemcpy(got.ctypes.data_as(ctypes.c_void_p), gd, got.nbytes, D2H) == 0
        finally:
            for p in (xd, td, gd):
                hip.hipFree(p)
        assert np.array_equal(got[:PARAMS], g_ref), n
        assert abs(got[PARAMS] / n - loss_ref) <= 1e-6 * loss_ref
    net.close()


def test_noalloc_gradient_rejects_invalid_sizes():
    net = ProposalNetwork(seed=1)
    # both are rejected before any device access (the addresses are never dereferenced)
    with pytest.raises(Exception):
        net.gradient_noalloc(1, 1, 0, 1)
    with pytest.raises(Exception):
        net.gradient_noalloc(1, 1, (1 << 40) + 1, 1)
    net.close()
