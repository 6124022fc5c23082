"""The reference's calling conventions through the C ABI on the GPU: a C
program (tests/c/sfft_caller.c, the sequence of the reference's
test/test.c), and sfft_exec_many with host pointer arrays and dense
outputs (src/sfft.cc:139-147).  Both are checked bit for bit against the
oracle run on the plan's own filters with the same glibc draws."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import oracle
import sfft_amd
from sfft_amd import _lib
from sfft_amd.signals import signal_np
from test_abi import _build_c_caller

pytestmark = pytest.mark.gpu


def _libc():
    libc = C.CDLL(None)
    libc.drand48.restype = C.c_double
    libc.random.restype = C.c_long
    return libc


def _filters(n, k, version):
    plan = sfft_amd.sfft(n, k, version, strict=False)
    lt, lf = plan.filter(0)
    et, ef = plan.filter(1)
    return lt, lf, et, ef


def test_c_caller_matches_oracle(torch_cuda, tmp_path):
    exe = _build_c_caller(tmp_path)
    fin, fout = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    print("case C caller v2 n=8192 k=50", flush=True)
    r = subprocess.run([exe, fin, fout, "4242"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    n, k = 8192, 50
    x = np.fromfile(fin, np.complex128)
    out = np.fromfile(fout, np.complex128)
    assert len(x) == n and len(out) == n
    # the program's draws: srand(17), srand48(4242), k tone positions from
    # drand48(), then the exec's random() / drand48() draws
    filters = _filters(n, k, 2)
    P = oracle.V12(n, k, True)
    P.set_filters(*filters)
    oracle.seed_glibc(17, 4242)
    libc = _libc()
    freqs = [int(np.floor(libc.drand48() * n)) for _ in range(k)]
    a, ai, off = P.draw()
    # the exec consumed exactly the oracle's draws
    words = r.stdout.split()
    assert int(words[1]) == libc.random() and float(words[2]) == libc.drand48()
    ref = P.exec(x, a, ai, off)
    assert np.array_equal(out, ref)
    assert int(words[-1]) == np.count_nonzero(ref)
    assert np.all(np.isin(freqs, np.nonzero(out)[0]))  # every tone located


@pytest.mark.parametrize("version", [1, 2])
def test_exec_many_host_pointers_dense(torch_cuda, version):
    """sfft_exec_many(plan, 4, complex_t **in, complex_t **out) with host
    memory, the reference's own convention: every dense, zero-filled output
    equals the oracle's with the draws consumed in signal order."""
    n, k, S = 1 << 16, 50, 4
    print(f"case exec_many v{version} n={n} k={k} S={S}", flush=True)
    xs = [signal_np(n, k, seed=600 + i, snr_db=30.0)[0] for i in range(S)]
    outs = [np.full(n, np.nan + 0j, np.complex128) for _ in range(S)]  # must be overwritten
    plan = sfft_amd.sfft(n, k, version, strict=False, seed=(17, 4242))
    L = _lib.load()
    pin = (C.c_void_p * S)(*[x.ctypes.data for x in xs])
    pout = (C.c_void_p * S)(*[o.ctypes.data for o in outs])
    L.sfft_exec_many(plan.sfft_plan, S, pin, pout)
    assert _lib.last_error() == ""
    oracle.seed_glibc(17, 4242)
    P = oracle.V12(n, k, version == 2)
    lt, lf = plan.filter(0)
    et, ef = plan.filter(1)
    P.set_filters(lt, lf, et, ef)
    for i in range(S):
        a, ai, off = P.draw()
        assert np.array_equal(outs[i], P.exec(xs[i], a, ai, off)), i


def test_exec_many_host_pointers_dense_v3(torch_cuda):
    """sfft_exec_many with host memory on a v3 plan: every dense output
    equals the oracle's (same filters, draws consumed in signal order) in
    its nonzero set and within 1e-9 of the largest value, zero elsewhere
    (src/sfft.cc:121-123, 139-147)."""
    n, k, S = 1 << 16, 50, 4
    print(f"case exec_many v3 n={n} k={k} S={S}", flush=True)
    xs = [signal_np(n, k, seed=650 + i)[0] for i in range(S)]
    outs = [np.full(n, np.nan + 0j, np.complex128) for _ in range(S)]
    plan = sfft_amd.sfft(n, k, 3, seed=(17, 4242))
    L = _lib.load()
    pin = (C.c_void_p * S)(*[x.ctypes.data for x in xs])
    pout = (C.c_void_p * S)(*[o.ctypes.data for o in outs])
    L.sfft_exec_many(plan.sfft_plan, S, pin, pout)
    assert _lib.last_error() == ""
    oracle.seed_glibc(17, 4242)
    P = oracle.V3(n, k)
    t1, f1 = plan.filter(0)
    t2, f2 = plan.filter(1)
    P.set_filters(t1, f1, t2, f2)
    for i in range(S):
        ref = P.exec(xs[i], P.draw())
        assert not np.isnan(outs[i]).any()
        if ref["early"]:
            assert np.count_nonzero(outs[i]) == 0
            continue
        assert np.array_equal(np.nonzero(outs[i])[0], np.nonzero(ref["out"])[0]), i
        assert np.abs(outs[i] - ref["out"]).max() <= 1e-9 * np.abs(ref["out"]).max(), i
