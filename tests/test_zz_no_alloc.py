"""The allocation-free exec (sfft_exec_sparse_noalloc) through the C ABI.

It must give exactly what sfft_exec_sparse gives for the same plan and
draws (bit for bit: same kernels, same arithmetic), on consecutive calls of
different sizes sharing the one scratch region, reject a plan whose
workspace exceeds that region before drawing anything, and reject v3 plans
and host buffers.
"""
import ctypes as C

import numpy as np
import pytest

import sfft_amd
from sfft_amd import _lib
from sfft_amd.signals import signal_np

pytestmark = pytest.mark.gpu


def _noalloc(torch, plan, x):
    L = _lib.load()
    cap = plan.capacity
    fr = torch.full((cap,), -1, dtype=torch.int32, device="cuda")
    va = torch.zeros(cap, dtype=torch.complex128, device="cuda")
    out = _lib.SfftOutput(0, cap, fr.data_ptr(), va.data_ptr())
    rc = L.sfft_exec_sparse_noalloc(plan.sfft_plan, x.data_ptr(), C.byref(out))
    torch.cuda.synchronize()
    assert rc == 0, _lib.last_error()
    c = int(out.count)
    return fr[:c].cpu().numpy(), va[:c].cpu().numpy()


# the second call is smaller: the scratch region keeps the first call's
# contents, and nothing of it may leak into the second
CALLS = [(2, 1 << 22, 64, 40.0, "fp32_refine"), (1, 1 << 16, 50, None, "fp64"),
         (2, 1 << 20, 64, 30.0, "fp64"), (1, 1 << 14, 50, 20.0, "fp32_refine")]


def test_noalloc_equals_exec_sparse(torch_cuda):
    torch = torch_cuda
    for i, (v, n, k, snr, prec) in enumerate(CALLS):
        print(f"case noalloc v{v} n={n} k={k} {prec}", flush=True)
        x, _, _ = signal_np(n, k, seed=70 + i, snr_db=snr)
        xd = torch.from_numpy(x).cuda()
        a = sfft_amd.sfft(n, k, v, strict=False, precision=prec, seed=(3 + i, 9 + i))
        b = sfft_amd.sfft(n, k, v, strict=False, precision=prec, seed=(3 + i, 9 + i))
        for rep in range(2):  # the plan's draws advance identically in both
            f1, v1 = _noalloc(torch, a, xd)
            f2, v2 = b.execute_sparse_host(x)
            assert len(f1) > 0
            assert np.array_equal(f1, f2)
            assert np.array_equal(v1, v2)


def test_noalloc_rejects(torch_cuda):
    torch = torch_cuda
    L = _lib.load()
    # v1 with B = 32768 in the FP32 form (FP64 location and FP32 estimation
    # buckets, their transform scratch, 2.7 M vote slots) needs ~43 MB of
    # workspace: above the 32 MiB region
    n, k = 1 << 24, 2000
    big = sfft_amd.sfft(n, k, 1, strict=False, precision="fp32_refine", seed=(1, 1))
    ref = sfft_amd.sfft(n, k, 1, strict=False, precision="fp32_refine", seed=(1, 1))
    xd = torch.zeros(n, dtype=torch.complex128, device="cuda")
    cap = big.capacity
    fr = torch.empty(cap, dtype=torch.int32, device="cuda")
    va = torch.empty(cap, dtype=torch.complex128, device="cuda")
    out = _lib.SfftOutput(0, cap, fr.data_ptr(), va.data_ptr())
    assert L.sfft_exec_sparse_noalloc(big.sfft_plan, xd.data_ptr(), C.byref(out)) == -1
    assert "MiB scratch" in _lib.last_error()
    # the rejected call drew nothing: the plan's next exec matches a fresh one
    small = torch.zeros(n, dtype=torch.complex128, device="cuda")
    a = big.execute_many_sparse([small], *big.make_sparse_outputs(1))
    b = ref.execute_many_sparse([small], *ref.make_sparse_outputs(1))
    assert np.array_equal(a, b)
    # host buffers and v3 plans
    n, k = 1 << 16, 50
    p2 = sfft_amd.sfft(n, k, 2, strict=False, seed=(1, 1))
    host = np.zeros(n, np.complex128)
    cap = p2.capacity
    fr = torch.empty(cap, dtype=torch.int32, device="cuda")
    va = torch.empty(cap, dtype=torch.complex128, device="cuda")
    out = _lib.SfftOutput(0, cap, fr.data_ptr(), va.data_ptr())
    assert L.sfft_exec_sparse_noalloc(p2.sfft_plan, host.ctypes.data, C.byref(out)) == -1
    p3 = sfft_amd.sfft(n, k, 3, seed=(1, 1))
    xd = torch.zeros(n, dtype=torch.complex128, device="cuda")
    out3 = _lib.SfftOutput(0, p3.capacity, fr.data_ptr(), va.data_ptr())
    assert L.sfft_exec_sparse_noalloc(p3.sfft_plan, xd.data_ptr(), C.byref(out3)) == -1
    assert "v1/v2" in _lib.last_error()
