"""Synthetic sparse-spectrum signals for tests and benchmarks.

Tone model: k distinct frequencies drawn uniformly without replacement, each
X[f] = e^{i theta} with theta uniform, and the time signal
x[t] = sum_f X[f] e^{+2 pi i f t / n} (unnormalised inverse DFT, the
convention of the reference's harness, src/simulation.cc:95-112).  Optional
complex Gaussian noise of per-sample variance k * 10^(-SNR/10).

The spectrum X is known by construction, so the noiseless oracle is exact.
"""
import numpy as np


def tones(n, k, seed):
    rng = np.random.default_rng(seed)
    f = np.sort(rng.choice(n, size=k, replace=False))
    vals = np.exp(2j * np.pi * rng.random(k))
    return f.astype(np.int64), vals


def signal_np(n, k, seed, snr_db=None):
    """Host signal (numpy complex128) and its sparse spectrum (freqs, vals)."""
    f, v = tones(n, k, seed)
    X = np.zeros(n, np.complex128)
    X[f] = v
    x = np.fft.ifft(X) * n
    if snr_db is not None:
        rng = np.random.default_rng(seed + 7919)
        sigma2 = k * 10.0 ** (-snr_db / 10.0)
        x += np.sqrt(sigma2 / 2) * (rng.standard_normal(n) + 1j * rng.standard_normal(n))
    return np.ascontiguousarray(x), f, v


def signals_device(n, k, seeds, snr_db=None, device="cuda"):
    """Device signals [len(seeds), n] complex128 built on the GPU with torch
    (inverse FFT of the sparse spectrum plus noise); returns (x, freqs, vals)."""
    import torch
    S = len(seeds)
    x = torch.zeros((S, n), dtype=torch.complex128, device=device)
    fs, vs = [], []
    for i, sd in enumerate(seeds):
        f, v = tones(n, k, sd)
        fs.append(f)
        vs.append(v)
        X = torch.zeros(n, dtype=torch.complex128, device=device)
        X[torch.from_numpy(f).to(device)] = torch.from_numpy(v).to(device)
        x[i] = torch.fft.ifft(X) * n
        if snr_db is not None:
            g = torch.Generator(device=device)
            g.manual_seed(int(sd) + 7919)
            sigma = float(np.sqrt(k * 10.0 ** (-snr_db / 10.0) / 2))
            x[i] += sigma * torch.complex(
                torch.randn(n, dtype=torch.float64, device=device, generator=g),
                torch.randn(n, dtype=torch.float64, device=device, generator=g))
    return x, fs, vs


def reference_input(n, k, seed, std_noise=0.0):
    """The reference harness's input (src/simulation.cc:95-112): k
    frequencies drawn with replacement as floor(drand48() * n), X[f] = 1,
    x = unnormalised inverse DFT, plus AWGN noise of standard deviation
    std_noise (src/utils.cc:250-280), from a drand48 stream seeded like
    srand48(seed).  Returns (x, X, snr)."""
    import ctypes as C
    from . import _lib
    x = np.empty(n, np.complex128)
    X = np.empty(n, np.complex128)
    snr = C.c_double(0)
    if _lib.load().sfft_generate_input(n, k, seed, std_noise, x.ctypes.data_as(C.c_void_p),
                                       X.ctypes.data_as(C.c_void_p), C.byref(snr)) != 0:
        raise ValueError(_lib.last_error())
    return x, X, snr.value
