"""CPU checks of the FP32-form model (oracle/fp32_model.c), the checker the
GPU's FP32 buckets are compared with bit for bit.

The model is not the reference (the reference's buckets are FP64); these
tests pin it to the FP64 oracle from outside: its buckets after the FFT
must equal the oracle's FP64 x_samp to FP32 accuracy, whatever the segment
size, and cutting the rows into segments only reorders the sums.
"""
import numpy as np
import pytest

import oracle
from sfft_amd.signals import signal_np


def _fp64_and_model(n, k, version, seg_rows, snr=30.0):
    x, _, _ = signal_np(n, k, seed=321, snr_db=snr)
    oracle.seed_glibc(17, 4242)
    P = oracle.V12(n, k, version == 2)
    a, ai, off = P.draw()
    tr = P.exec(x, a, ai, off, trace=True)
    f = P.filters()
    p = P.p
    nmax = max(p.B_loc, p.B_est, p.W_comb if version == 2 else 0)
    out, o = [], 0
    for l in range(P.loops):
        B = p.B_loc if l < p.loc_loops else p.B_est
        taps = f["loc_time"] if l < p.loc_loops else f["est_time"]
        got = oracle.fp32_buckets(x, n, B, taps, ai[l], seg_rows, nmax)
        out.append((got, tr["x_samp"][o:o + B]))
        o += B
    return out


@pytest.mark.parametrize("version,n,k", [(2, 1 << 16, 50), (1, 1 << 16, 50), (2, 1 << 18, 64)])
@pytest.mark.parametrize("seg_rows", [32, 512])
def test_model_matches_fp64_buckets(version, n, k, seg_rows):
    for got, ref in _fp64_and_model(n, k, version, seg_rows):
        scale = np.abs(ref).max()
        # compensated FP32 sums: errors at the FP32 rounding of inputs and taps
        assert np.abs(got - ref).max() <= 3e-6 * scale


def test_segments_only_reorder():
    """Four segments (128 rows, seg_rows 32) against one (seg_rows 512): the
    same terms in another order, so the buckets differ by rounding only
    (a dropped or doubled segment would move them by whole samples)."""
    a = _fp64_and_model(1 << 16, 50, 2, 32)
    b = _fp64_and_model(1 << 16, 50, 2, 512)
    for (ga, ref), (gb, _) in zip(a, b):
        scale = np.abs(ref).max()
        assert np.abs(ga - gb).max() <= 3e-6 * scale
