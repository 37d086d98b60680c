"""zmpm_sort_keys_noalloc: the radix sort of zmpm_sort_keys on the library's static device
scratch, with no context and no device allocation.  On the GPU its output equals
zmpm_sort_keys' bit for bit on two calls in a row, the second smaller (the scratch keeps the
first call's contents), and numpy's stable argsort; argument errors come back as status codes
before any device call."""
import ctypes

import numpy as np
import pytest

from zeno_amd import zmpm

MAX_KEYS = 1 << 20


def test_noalloc_sort_argument_errors_without_device():
    L = zmpm.load_library()
    P = ctypes.c_void_p(0x1000)  # never dereferenced: every case fails before a device call
    assert L.zmpm_sort_keys_noalloc(P, MAX_KEYS + 1, 20, P, P, None) == 3  # ZMPM_ERR_CAPACITY
    assert b"1048576" in L.zmpm_last_error()
    assert L.zmpm_sort_keys_noalloc(P, -1, 20, P, P, None) == 3
    assert L.zmpm_sort_keys_noalloc(P, 10, 0, P, P, None) == 1
    assert L.zmpm_sort_keys_noalloc(P, 10, 33, P, P, None) == 1
    assert L.zmpm_sort_keys_noalloc(None, 10, 20, P, P, None) == 1
    assert L.zmpm_sort_keys_noalloc(P, 10, 20, P, ctypes.c_void_p(0x2000), None) == 1  # keys_out aliases keys
    assert L.zmpm_sort_keys_noalloc(None, 0, 20, None, None, None) == 0  # nothing to sort


@pytest.mark.gpu
def test_noalloc_sort_equals_context_sort():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    L = zmpm.load_library()
    dev = "cuda:0"
    torch.cuda.set_device(0)
    ctx = zmpm.Context(1.0 / 64, 64, MAX_KEYS)
    stream = torch.cuda.current_stream().cuda_stream
    rng = np.random.default_rng(12)
    for n, bits in ((MAX_KEYS, 26), (300_001, 13)):  # the second call is smaller
        print(f"noalloc sort: n={n} key_bits={bits}", flush=True)
        keys = rng.integers(0, 1 << bits, n, dtype=np.uint32)
        keys[: n // 3] = keys[0]  # long runs of equal keys: stability matters
        k = torch.from_numpy(keys.view(np.int32)).to(dev)
        ko, po = torch.empty_like(k), torch.empty_like(k)
        rc = L.zmpm_sort_keys_noalloc(ctypes.c_void_p(k.data_ptr()), n, bits, ctypes.c_void_p(ko.data_ptr()),
                                      ctypes.c_void_p(po.data_ptr()), ctypes.c_void_p(stream))
        assert rc == 0, L.zmpm_last_error()
        torch.cuda.synchronize()
        rk, rp = ctx.sort_keys(k, bits)
        torch.cuda.synchronize()
        assert torch.equal(ko, rk) and torch.equal(po, rp)
        order = np.argsort(keys, kind="stable")
        np.testing.assert_array_equal(po.cpu().numpy(), order.astype(np.int32))
        np.testing.assert_array_equal(ko.cpu().numpy().view(np.uint32), keys[order])
    ctx.close()
