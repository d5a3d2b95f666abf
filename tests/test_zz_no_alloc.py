"""The allocation-free apply, hifde_apply_device_noalloc, through the C ABI.

It runs the single-vector sweeps of hifde_apply_device with its scratch in a
fixed per-GPU array of the library instead of buffers it allocates, so its
results must equal hifde_apply_device's bit for bit.  Two calls in a row,
the second on a smaller grid, check that nothing depends on what the shared
scratch held before.
"""
import ctypes

import numpy as np
import pytest

from hifde3d_amd import problems as pb

pytestmark = pytest.mark.gpu

ERR_SIZE, ERR_STATE = -3, -4


@pytest.fixture(scope="module")
def H():
    import hifde3d_amd
    hifde3d_amd.lib()
    return hifde3d_amd


def _factored(H, N, P, precision, seed):
    NC = N + 1
    f = H.HIFFactor(N, P, 1e-6, precision=precision)
    f.set_stencil(pb.pow10_coeff(NC, seed), pb.zeros_coeff(NC), True)
    assert f.Factor() == 0
    return f


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_noalloc_apply_equals_apply_device(H, precision):
    import torch
    L = H.lib()
    # The HIP runtime sets up some device memory of its own lazily, on the
    # first use of a path (queues, signals); one call on a throwaway object
    # does that before free memory is compared around calls on fresh ones.
    w = _factored(H, 15, 4, precision, 9)
    uw = torch.zeros(w.size, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    assert L.hifde_apply_device_noalloc(w._h, ctypes.c_void_p(uw.data_ptr()), 1) == 0
    # second problem smaller than the first: the scratch keeps the first
    # call's contents when the second runs
    for N, P, seed in [(31, 4, 3), (15, 4, 4)]:
        f = _factored(H, N, P, precision, seed)
        u = torch.from_numpy(pb.rhs_normal(N + 1, 10 + seed)).cuda()
        torch.cuda.synchronize()
        for inverse in (1, 0, 1):
            a, b = u.clone(), u.clone()
            torch.cuda.synchronize()
            free0 = torch.cuda.mem_get_info()[0]
            assert L.hifde_apply_device_noalloc(f._h, ctypes.c_void_p(b.data_ptr()), inverse) == 0
            torch.cuda.synchronize()
            # nothing was allocated on the device by the call
            assert torch.cuda.mem_get_info()[0] == free0
            assert L.hifde_apply_device(f._h, ctypes.c_void_p(a.data_ptr()), 1, f.size,
                                        inverse) == 0
            torch.cuda.synchronize()
            assert torch.equal(a, b), (N, precision, inverse)
            assert float(torch.linalg.norm(a)) > 0
        f.close()
    w.close()


def test_noalloc_apply_rejects_large_grid_and_unfactored(H):
    import torch
    L = H.lib()
    u = torch.zeros(8, dtype=torch.float64, device="cuda")
    big = H.HIFFactor(159, 5, 1e-3)  # (N+1)^3 = 160^3 > 128^3; not factored
    assert L.hifde_apply_device_noalloc(big._h, ctypes.c_void_p(u.data_ptr()), 1) == ERR_SIZE
    assert b"scratch limit" in L.hifde_last_error()
    small = H.HIFFactor(15, 4, 1e-3)
    assert L.hifde_apply_device_noalloc(small._h, ctypes.c_void_p(u.data_ptr()), 1) == ERR_STATE
    big.close()
    small.close()
