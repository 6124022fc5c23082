"""CPU models of the index arithmetic of two GPU building blocks, checked
against their plain formulations before any GPU run:

* the LDS FFT's pass schedule (``lds_fft_pass`` in device_common.hpp): up
  to three radix-2 DIT stages per pass on groups of 2^R points i0 + m*h must
  perform exactly the butterflies of one stage per pass, in an order that
  gives the same bits, and touch every point once per pass;
* the top-num cutoff search (``block_top_num``): byte-wise radix selection
  with a 64-lane prefix-sum/ballot bin search and an exact-rank finish once
  at most 128 keys share the prefix must find the (num+1)-th largest key,
  which with the tie rule of src/utils.cc:130-158 fixes the selected set.
"""
import numpy as np
import pytest


def _cmul(a, b):
    return complex(a.real * b.real - a.imag * b.imag, a.real * b.imag + a.imag * b.real)


def _radix2(buf, logN, tw, tws):
    N = 1 << logN
    buf = list(buf)
    for s in range(1, logN + 1):
        h = 1 << (s - 1)
        ts = (N >> s) * tws
        for b in range(N >> 1):
            j = b & (h - 1)
            i0 = ((b >> (s - 1)) << s) + j
            w = tw[j * ts]
            u, t = buf[i0], _cmul(buf[i0 + h], w)
            buf[i0], buf[i0 + h] = u + t, u - t
    return buf


def _fused(buf, logN, tw, tws):
    N = 1 << logN
    buf = list(buf)

    def one_pass(s, R):
        M, h = 1 << R, 1 << (s - 1)
        seen = set()
        for g in range(N >> R):
            j = g & (h - 1)
            i0 = ((g >> (s - 1)) << (s - 1 + R)) + j
            idx = [i0 + m * h for m in range(M)]
            assert 0 <= min(idx) and max(idx) < N
            assert not seen.intersection(idx)
            seen.update(idx)
            v = [buf[q] for q in idx]
            for r in range(R):
                ts = (N >> (s + r)) * tws
                for m in range(M):
                    if m & (1 << r):
                        continue
                    m1 = m + (1 << r)
                    ti = (j + (m & ((1 << r) - 1)) * h) * ts
                    assert ti < len(tw)
                    u, t = v[m], _cmul(v[m1], tw[ti])
                    v[m], v[m1] = u + t, u - t
            for m, q in enumerate(idx):
                buf[q] = v[m]
        assert len(seen) == N

    s = 1
    while logN - s + 1 >= 3:
        one_pass(s, 3)
        s += 3
    if logN - s + 1 == 2:
        one_pass(s, 2)
    elif logN - s + 1 == 1:
        one_pass(s, 1)
    return buf


@pytest.mark.parametrize("logN", range(0, 11))
def test_fused_fft_passes_equal_radix2(logN):
    rng = np.random.default_rng(logN)
    for extra in (0, 2):  # twiddle table built for a larger transform
        N, Nmax = 1 << logN, 1 << (logN + extra)
        tws = Nmax // N
        tw = [complex(np.cos(-2 * np.pi * j / Nmax), np.sin(-2 * np.pi * j / Nmax))
              for j in range(max(1, Nmax // 2))]
        x = [complex(a, b) for a, b in rng.standard_normal((N, 2))]
        assert _fused(x, logN, tw, tws) == _radix2(x, logN, tw, tws)  # same bits
        if logN:
            rev = [int(format(i, "0%db" % logN)[::-1], 2) for i in range(N)]
            out = _fused([x[rev[i]] for i in range(N)], logN, tw, tws)
            assert np.allclose(out, np.fft.fft(x))


def _cutoff_model(keys, num):
    """block_top_num's search for the (num+1)-th largest key."""
    N = len(keys)
    rank = N - num - 1
    prefix, hi_mask = 0, 0
    for p in range(7, -1, -1):
        shift = 8 * p
        match = keys[(keys & np.uint64(hi_mask)) == np.uint64(prefix)]
        hist = np.bincount(((match >> np.uint64(shift)) & np.uint64(255)).astype(np.int64),
                           minlength=256)
        # lane l holds bins 4l..4l+3; first lane whose inclusive sum > rank
        lane_sum = hist.reshape(64, 4).sum(axis=1)
        incl = np.cumsum(lane_sum)
        lane = int(np.argmax(rank < incl))
        cum, q = int(incl[lane] - lane_sum[lane]), 0
        while q < 3 and rank >= cum + hist[4 * lane + q]:
            cum += int(hist[4 * lane + q])
            q += 1
        d = 4 * lane + q
        prefix |= d << shift
        hi_mask |= 255 << shift
        rank -= cum
        cand = int(hist[d])
        if p > 0 and cand <= 128:
            ks = keys[(keys & np.uint64(hi_mask)) == np.uint64(prefix)]
            assert len(ks) == cand
            for kj in ks:
                lt, le = int((ks < kj).sum()), int((ks <= kj).sum())
                if lt <= rank < le:
                    return int(kj)
            raise AssertionError("no key holds the rank")
    return prefix


@pytest.mark.parametrize("kind", ["noise", "ties", "zeros", "few_values"])
def test_top_num_cutoff_model(kind):
    rng = np.random.default_rng(7)
    for N, num in [(4096, 128), (2048, 100), (512, 10), (129, 128)]:
        if kind == "noise":
            mag = rng.exponential(size=N) * 10.0 ** rng.integers(-3, 3, size=N)
        elif kind == "ties":
            mag = np.round(rng.exponential(size=N), 2)
        elif kind == "zeros":
            mag = np.zeros(N)
        else:
            mag = rng.choice([0.5, 1.0, 2.0, 1e-300], size=N)
        keys = mag.astype(np.float64).view(np.uint64)
        cut = _cutoff_model(keys, num)
        assert cut == int(np.sort(keys)[N - num - 1])
