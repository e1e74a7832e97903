"""nlfem_mlp_forward_noalloc (the allocation-free surrogate forward pass) against
nlfem_mlp_forward, through the C ABI.

The caller owns every device buffer (allocated here with torch, which is only plumbing for
device memory and the stream); the library keeps the packed network in its fixed device
array.  Two calls in a row, the second with fewer samples and the other network of the
default ensemble, so the second call must repack the array rather than reuse what the first
left there.  Both use the same packing and forward kernels as nlfem_mlp_forward, so the
outputs must be identical bit for bit.
"""
import ctypes as C

import numpy as np
import pytest

from nlfem_amd import _lib, surrogate

pytestmark = pytest.mark.gpu


def _inputs(n, seed):
    rng = np.random.default_rng(seed)
    return (rng.normal(size=(n, 12)) * 0.5).astype(np.float32)


def _noalloc(torch, params, norm, X):
    dev = torch.device("cuda:0")
    dp = torch.from_numpy(np.ascontiguousarray(params, dtype=np.float32)).to(dev)
    dn = torch.from_numpy(np.ascontiguousarray(norm, dtype=np.float32)).to(dev)
    dX = torch.from_numpy(np.ascontiguousarray(X, dtype=np.float32)).to(dev)
    dY = torch.full((X.shape[0], 9), float("nan"), dtype=torch.float32, device=dev)
    stream = torch.cuda.current_stream(dev)
    torch.cuda.synchronize(dev)
    rc = _lib.lib().nlfem_mlp_forward_noalloc(C.c_void_p(dp.data_ptr()), C.c_void_p(dn.data_ptr()),
                                              C.c_void_p(dX.data_ptr()), X.shape[0], C.c_void_p(dY.data_ptr()),
                                              C.c_void_p(stream.cuda_stream))
    _lib.check(rc)
    torch.cuda.synchronize(dev)
    return dY.cpu().numpy()


def test_noalloc_forward_matches_allocating_entry_point():
    torch = pytest.importorskip("torch")
    assert torch.cuda.is_available()
    params, norm, meta = surrogate.load()
    print("[noalloc] call 1: n=5000, network 0", flush=True)
    X1 = _inputs(5000, 1)
    Y1 = _noalloc(torch, params[0], norm[0], X1)
    R1 = surrogate.forward(params[0], norm[0], X1)
    assert np.array_equal(Y1, R1), float(np.nanmax(np.abs(Y1 - R1)))
    print("[noalloc] call 2: n=777, network 1", flush=True)
    X2 = _inputs(777, 2)
    Y2 = _noalloc(torch, params[1], norm[1], X2)
    R2 = surrogate.forward(params[1], norm[1], X2)
    assert np.array_equal(Y2, R2), float(np.nanmax(np.abs(Y2 - R2)))
    # the two networks differ, so a stale pack would have shown
    assert not np.array_equal(surrogate.forward(params[0], norm[0], X2), R2)


def test_noalloc_rejects_bad_sizes():
    L = _lib.lib()
    one = C.c_void_p(1)
    assert L.nlfem_mlp_forward_noalloc(one, one, one, -1, one, None) != 0
    assert L.nlfem_mlp_forward_noalloc(one, one, one, 1 << 31, one, None) != 0
    assert "2^31" in L.nlfem_last_error().decode()
    assert L.nlfem_mlp_forward_noalloc(None, one, one, 4, one, None) != 0
