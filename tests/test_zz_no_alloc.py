"""efm_port_overlap_noalloc (results staged in the library's fixed device array, no
allocation) against efm_port_overlap on the same inputs, called through the C ABI: two
calls in a row, the second with fewer ports, fewer right-hand sides and a shorter vector;
results must be bit-identical (same kernel, same fixed-order reduction)."""
import ctypes as C
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(torch, rng, n, P, L, R):
    gi = np.full((P, L), -1, np.int32)
    for j in range(P):
        k = rng.integers(1, L + 1)
        gi[j, :k] = rng.choice(n, size=k, replace=False)
    gv = rng.standard_normal((P, L)) + 1j * rng.standard_normal((P, L))
    X = rng.standard_normal((n, R)) + 1j * rng.standard_normal((n, R))
    dev = torch.device("cuda:0")
    return (torch.from_numpy(gi).to(dev), torch.from_numpy(gv).to(dev), torch.from_numpy(X).to(dev), gi, gv, X)


def _call(lib, fn, n, P, L, R, gi, gv, X):
    from edgefem_amd import _capi
    out = np.zeros(P * R, np.complex128)
    rc = getattr(lib, fn)(n, P, L, _capi.dptr(gi), _capi.dptr(gv), _capi.dptr(X), R,
                          out.ctypes.data_as(C.c_void_p), _capi.stream_ptr())
    assert rc == 0, lib.efm_last_error()
    return out.reshape(P, R)


def test_port_overlap_noalloc_matches_allocating_entry_point():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    torch.cuda.set_device(0)
    from edgefem_amd import _capi
    lib = _capi.load_library()
    rng = np.random.default_rng(42)
    for (n, P, L, R) in [(20000, 12, 700, 8), (5000, 3, 300, 2)]:
        print(f"[gpu] port overlap noalloc n={n} ports={P} len={L} nrhs={R}", flush=True)
        sys.stdout.flush()
        gi_d, gv_d, X_d, gi, gv, X = _case(torch, rng, n, P, L, R)
        ref = _call(lib, "efm_port_overlap", n, P, L, R, gi_d, gv_d, X_d)
        got = _call(lib, "efm_port_overlap_noalloc", n, P, L, R, gi_d, gv_d, X_d)
        assert np.array_equal(got.view(np.float64), ref.view(np.float64))
        # the values themselves, against numpy on the host
        exp = np.zeros((P, R), np.complex128)
        for j in range(P):
            m = gi[j] >= 0
            exp[j] = np.conj(gv[j, m]) @ X[gi[j, m]]
        np.testing.assert_allclose(got, exp, rtol=1e-12, atol=1e-12)
    # a size beyond the fixed array is rejected before any device work
    out = np.zeros(1, np.complex128)
    rc = lib.efm_port_overlap_noalloc(10, 4097, 1, _capi.dptr(gi_d), _capi.dptr(gv_d), _capi.dptr(X_d), 1,
                                      out.ctypes.data_as(C.c_void_p), _capi.stream_ptr())
    assert rc == 1 and b"4096" in lib.efm_last_error()
