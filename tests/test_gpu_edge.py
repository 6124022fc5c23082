"""GPU parity on edge inputs: the all-zero signal (every magnitude ties, so
the top-2k tie rule decides: > cutoff none, == cutoff in index order), k=1,
the smallest feasible n, tones at frequency 0 and n-1, and the reference
harness's own input (X[f] = 1 at k draws with replacement, so fewer than k
distinct tones).  v1/v2 compare bit for bit with the oracle run on the
plan's filters; v3 as in test_gpu_v3.
"""
import numpy as np
import pytest

import oracle
import sfft_amd
from sfft_amd.signals import reference_input, signal_np

pytestmark = pytest.mark.gpu


def _v12_check(version, n, k, x):
    plan = sfft_amd.sfft(n, k, version, strict=False, seed=(17, 4242))
    lt, lf = plan.filter(0)
    et, ef = plan.filter(1)
    oracle.seed_glibc(17, 4242)
    P = oracle.V12(n, k, version == 2)
    P.set_filters(lt, lf, et, ef)
    a, ai, off = P.draw()
    tr = P.exec(x, a, ai, off, trace=True)
    ref = tr["out"]
    # the located set itself (values may be 0): sparse results vs the
    # oracle's hit list
    located = np.unique(tr["hits"][:tr["hits_prefill"]] if version == 2 else tr["hits"])
    fr, va = plan.execute_sparse_host(x)
    assert np.array_equal(np.unique(fr), located), (len(fr), len(located))
    assert np.array_equal(va, ref[fr])
    return ref


def _v3_check(n, k, x):
    plan = sfft_amd.sfft(n, k, 3, seed=(17, 4242))
    t1, f1 = plan.filter(0)
    t2, f2 = plan.filter(1)
    oracle.seed_glibc(17, 4242)
    P = oracle.V3(n, k)
    P.set_filters(t1, f1, t2, f2)
    ref = P.exec(x, P.draw())
    fr, va = plan.execute_sparse_host(x)
    if ref["early"]:
        assert len(fr) == 0
        return fr, va
    assert np.array_equal(fr, ref["freqs"])
    if len(fr):
        assert np.abs(va - ref["vals"]).max() <= 1e-9 * max(1.0, np.abs(ref["vals"]).max())
    return fr, va


@pytest.mark.parametrize("version", [1, 2])
def test_zero_signal_ties(torch_cuda, version):
    n, k = 1 << 14, 50
    print(f"case v{version} zero signal n={n} k={k}", flush=True)
    out = _v12_check(version, n, k, np.zeros(n, np.complex128))
    assert np.all(out == 0)


def test_zero_signal_v3(torch_cuda):
    n, k = 1 << 16, 50
    print(f"case v3 zero signal n={n} k={k}", flush=True)
    fr, va = _v3_check(n, k, np.zeros(n, np.complex128))
    assert np.all(va == 0)


@pytest.mark.parametrize("version,n,k", [(1, 1 << 14, 1), (2, 1 << 14, 1), (1, 1 << 10, 10),
                                         (2, 1 << 10, 10), (1, 1 << 16, 2)])
def test_tiny_k_and_n(torch_cuda, version, n, k):
    print(f"case v{version} n={n} k={k}", flush=True)
    x, f, v = signal_np(n, k, seed=31 + k)
    out = _v12_check(version, n, k, x)
    assert np.all(np.isin(f, np.nonzero(out)[0]))


@pytest.mark.parametrize("version", [1, 2])
def test_tones_at_band_edges(torch_cuda, version):
    n, k = 1 << 16, 4
    f = np.array([0, 1, n - 2, n - 1])
    X = np.zeros(n, np.complex128)
    X[f] = np.exp(2j * np.pi * np.array([0.1, 0.2, 0.3, 0.4]))
    x = np.fft.ifft(X) * n
    print(f"case v{version} band-edge tones n={n}", flush=True)
    out = _v12_check(version, n, k, x)
    assert np.all(np.isin(f, np.nonzero(out)[0]))
    assert np.abs(out[f] - np.conj(X[f])).max() < 1e-6


def test_tones_at_band_edges_v3(torch_cuda):
    n, k = 1 << 16, 8
    f = np.array([0, 1, 2, 3, n - 4, n - 3, n - 2, n - 1])
    X = np.zeros(n, np.complex128)
    X[f] = np.exp(2j * np.pi * np.arange(8) / 8.3)
    x = np.fft.ifft(X) * n
    print(f"case v3 band-edge tones n={n}", flush=True)
    _v3_check(n, k, x)


@pytest.mark.parametrize("version", [1, 2, 3])
def test_reference_harness_input(torch_cuda, version):
    """src/simulation.cc:95-112: X[f] = 1 at k draws with replacement."""
    n, k = 1 << 16, 50
    x, X, _ = reference_input(n, k, seed=2024)
    print(f"case v{version} reference input n={n} k={k} distinct={np.count_nonzero(X)}",
          flush=True)
    if version == 3:
        fr, va = _v3_check(n, k, x)
        return
    out = _v12_check(version, n, k, x)
    sup = np.nonzero(X)[0]
    assert np.all(np.isin(sup, np.nonzero(out)[0]))
    assert np.abs(out[sup] - 1.0).max() < 0.1   # src/verification.cc:37-58


def test_host_input_after_freed_device_memory(torch_cuda):
    """Host signals after device buffers the library has already seen were
    freed: the freed address ranges can come back as host memory, so the
    library must not remember them as device memory across calls."""
    torch = torch_cuda
    n, k = 1 << 20, 50
    warm = sfft_amd.sfft(n, k, 1, strict=False, seed=(17, 4242))
    for rep in range(3):
        print(f"case freed device memory, round {rep}", flush=True)
        xd = torch.zeros((8, n), dtype=torch.complex128, device="cuda")  # 128 MiB
        fr, va = warm.make_sparse_outputs(8)
        warm.execute_many_sparse(xd, fr, va)
        torch.cuda.synchronize()
        del xd, fr, va
        torch.cuda.empty_cache()
        for i in range(2):
            x = signal_np(n, k, seed=90 + 2 * rep + i)[0]
            on_host = sfft_amd.sfft(n, k, 1, strict=False, seed=(17, 4242))
            on_dev = sfft_amd.sfft(n, k, 1, strict=False, seed=(17, 4242))
            fh, vh = on_host.execute_sparse_host(x)
            f1, v1 = on_dev.make_sparse_outputs(1)
            c = on_dev.execute_many_sparse([torch.from_numpy(x).cuda()], f1, v1)
            torch.cuda.synchronize()
            assert np.array_equal(fh, f1[0, :int(c[0])].cpu().numpy())
            assert np.array_equal(vh, v1[0, :int(c[0])].cpu().numpy())
