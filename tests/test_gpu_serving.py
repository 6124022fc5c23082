"""GPU tests of the serving loop (include/sfft_server.h).

Each request carries its own RNG seeds, so its result must be the same
whether it runs alone (a plan of its own, seeded the same way) or batched
with other requests while other batches share the device: the stated
tolerance is exact equality of the located set and of every value, in both
precisions (no arithmetic depends on the batch).
"""
import numpy as np
import pytest

import sfft_amd
from sfft_amd.server import Server, output_capacity
from sfft_amd.signals import signals_device

pytestmark = pytest.mark.gpu

# (version, n, k, snr_db, precision)
MIX = [
    (2, 1 << 18, 30, 30.0, "fp32_refine"),
    (2, 1 << 20, 64, 30.0, "fp32_refine"),
    (2, 1 << 20, 64, 30.0, "fp64"),
    (1, 1 << 18, 100, 30.0, "fp32_refine"),
    (1, 1 << 16, 50, None, "fp64"),
    (3, 1 << 20, 100, None, "fp64"),
    (3, 1 << 18, 50, None, "fp64"),
    (2, 1 << 22, 64, 40.0, "fp32_refine"),
]


def _requests(torch, copies):
    reqs = []
    for i, (v, n, k, snr, prec) in enumerate(MIX):
        seeds = [100 * i + c for c in range(copies)]
        x, _, _ = signals_device(n, k, seeds, snr_db=snr)
        for c in range(copies):
            reqs.append((v, n, k, prec, x[c].contiguous(), (1000 + 7 * i + c, 77 + i * 3 + c)))
    return reqs


def _alone(torch, v, n, k, prec, x, seed):
    plan = sfft_amd.sfft(n, k, v, strict=False, precision=prec, seed=seed)
    fr, va = plan.make_sparse_outputs(1)
    counts = plan.execute_many_sparse([x], fr, va)
    torch.cuda.synchronize()
    c = int(counts[0])
    return fr[0, :c].cpu().numpy(), va[0, :c].cpu().numpy()


def test_mix_equals_alone(torch_cuda):
    torch = torch_cuda
    reqs = _requests(torch, copies=3)
    order = np.random.default_rng(5).permutation(len(reqs))
    with Server(workers=3, max_batch=4) as srv:
        tickets = {}
        for i in order:  # interleaved keys, submitted back to back
            v, n, k, prec, x, seed = reqs[i]
            tickets[i] = srv.submit(x, k, v, prec, seed)
        got = {i: srv.wait(t) for i, t in tickets.items()}
        st = srv.stats()
    assert st["completed"] == len(reqs) and st["failed"] == 0
    assert st["batched_requests"] == len(reqs)
    assert st["batches"] < len(reqs)            # some requests shared a batch
    assert st["max_concurrent_batches"] >= 1
    for i, (v, n, k, prec, x, seed) in enumerate(reqs):
        print(f"case v{v} n={n} k={k} {prec} seed={seed}", flush=True)
        fr, va, lat = got[i]
        f1, v1 = _alone(torch, v, n, k, prec, x, seed)
        assert lat > 0
        assert np.array_equal(fr.cpu().numpy(), f1)
        assert np.array_equal(va.cpu().numpy(), v1)


def test_small_budget_cuts_batches(torch_cuda):
    """Under a tight HBM budget every request still completes, batches are
    cut to fit and the accounted footprint stays within the budget."""
    torch = torch_cuda
    n, k = 1 << 20, 64
    x, _, _ = signals_device(n, k, list(range(8)), snr_db=30.0)
    probe = sfft_amd.sfft(n, k, 2, strict=False, precision="fp32_refine")
    tables, per = probe.memory()
    del probe
    budget = int(tables + 3.5 * per)   # room for 3 signals at a time
    with Server(hbm_budget=budget, workers=1, max_batch=8) as srv:
        t = [srv.submit(x[i].contiguous(), k, 2, "fp32_refine", (i + 1, i + 2)) for i in range(8)]
        res = [srv.wait(tt) for tt in t]
        st = srv.stats()
    assert st["completed"] == 8
    assert st["hbm_peak"] <= budget
    assert st["batches"] >= 3 and st["batched_requests"] == 8
    for i, (fr, va, lat) in enumerate(res):
        f1, v1 = _alone(torch, 2, n, k, "fp32_refine", x[i].contiguous(), (i + 1, i + 2))
        assert np.array_equal(fr.cpu().numpy(), f1)
        assert np.array_equal(va.cpu().numpy(), v1)


def test_rejects_infeasible(torch_cuda):
    torch = torch_cuda
    with Server(workers=1) as srv:
        x = torch.zeros(1 << 22, dtype=torch.complex128, device="cuda")
        with pytest.raises(sfft_amd.SfftError):
            srv.submit(x, 1000, 2)   # v2 with W_Comb < 2k+1
        st = srv.stats()
    assert st["rejected"] == 1 and st["submitted"] == 0
    with pytest.raises(sfft_amd.SfftError):
        output_capacity(1 << 22, 1000, 2)


def test_v3_keys_share_a_launch(torch_cuda):
    """Queued v3 requests of different (n, k) run as one multi-plan launch
    and give what each gives alone."""
    torch = torch_cuda
    keys = [(1 << 18, 50), (1 << 20, 100), (1 << 20, 300)]
    xs = []
    for g, (n, k) in enumerate(keys):
        x, _, _ = signals_device(n, k, [40 + g, 50 + g])
        xs += [(n, k, x[0].contiguous()), (n, k, x[1].contiguous())]
    xv, _, _ = signals_device(1 << 22, 64, [9], snr_db=30.0)
    with Server(workers=1, max_batch=16) as srv:
        for n, k in keys:
            srv.prepare(n, k, 3, "fp64")
        srv.prepare(1 << 22, 64, 2, "fp32_refine")
        tv = srv.submit(xv[0].contiguous(), 64, 2, "fp32_refine", (5, 6))
        tk = [srv.submit(x, k, 3, "fp64", (100 + i, 200 + i)) for i, (n, k, x) in enumerate(xs)]
        srv.wait(tv)
        got = [srv.wait(t) for t in tk]
        st = srv.stats()
    assert st["completed"] == 7 and st["batches"] == 2
    for i, (n, k, x) in enumerate(xs):
        f1, v1 = _alone(torch, 3, n, k, "fp64", x, (100 + i, 200 + i))
        assert np.array_equal(got[i][0].cpu().numpy(), f1)
        assert np.array_equal(got[i][1].cpu().numpy(), v1)


def test_rejects_host_pointers(torch_cuda):
    """Requests are checked before they are queued: an input or output in
    host memory (the kernels would fault on it) is rejected with an error,
    and nothing runs.  (The check reads allocation ranges; a tensor carved
    from a larger allocation is judged by that allocation.)"""
    import ctypes as C
    from sfft_amd import _lib
    torch = torch_cuda
    n, k = 1 << 16, 50
    L = _lib.load()
    host_in = np.zeros(n, np.complex128)
    cap = output_capacity(n, k, 2)
    dev_in = torch.zeros(n, dtype=torch.complex128, device="cuda")
    fr = torch.empty(cap, dtype=torch.int32, device="cuda")
    va = torch.empty(cap, dtype=torch.complex128, device="cuda")
    host_fr = np.zeros(cap, np.int32)
    with Server(workers=1) as srv:
        cases = [(host_in.ctypes.data, fr.data_ptr(), va.data_ptr(), "input"),
                 (dev_in.data_ptr(), host_fr.ctypes.data, va.data_ptr(), "freq")]
        for pin, pfr, pva, what in cases:
            r = _lib.SfftRequest(n, k, 1, 1, pin, pfr, pva, cap, 1, 2)
            assert L.sfft_server_submit(srv._h, C.byref(r)) == -1
            assert what in _lib.last_error()
        st = srv.stats()
    assert st["rejected"] == 2 and st["submitted"] == 0 and st["completed"] == 0


def _cfg5_draw(count, seed):
    """A seeded draw of the serving workload: n uniform on {2^18, 2^20, 2^22,
    2^24}, k = floor(10 * 200^U), v2 (SNR 30 dB) or v3 (noiseless) 50/50;
    v2 draws the reference cannot run (W_Comb < 2k+1) are redrawn."""
    rng = np.random.default_rng(seed)
    sizes = [1 << 18, 1 << 20, 1 << 22, 1 << 24]
    reqs = []
    while len(reqs) < count:
        n = int(sizes[rng.integers(4)])
        k = int(np.floor(10 * 200 ** rng.random()))
        v = 2 if rng.random() < 0.5 else 3
        try:
            output_capacity(n, k, v)
        except sfft_amd.SfftError:
            continue
        reqs.append((v, n, k))
    return reqs


def test_cfg5_mix_equals_alone(torch_cuda):
    """The serving workload, Poisson arrivals at 2000/s into 3 workers:
    v3 up to k = 2000 (W_Man = 16384, above the LDS transform size), v2 at
    the feasibility edge (k = 2^j - 1 with W_Comb = 2k + 2), both
    precisions.  Every result equals the same request run alone (exact
    equality of located set and values), batches overlap on the device, and
    four small requests are checked against the oracle."""
    import oracle
    torch = torch_cuda
    reqs = [(v, n, k, "fp32_refine" if i % 2 else "fp64")
            for i, (v, n, k) in enumerate(_cfg5_draw(26, 555))]
    # the edges: v3 with k > 1000, v2 with B = W_Comb = 2k + 2 (k = 511 at 2^22)
    reqs += [(3, 1 << 24, 1800, "fp64"), (3, 1 << 22, 1200, "fp64"),
             (2, 1 << 22, 511, "fp32_refine"), (2, 1 << 18, 255, "fp64"),
             (2, 1 << 18, 30, "fp64"), (3, 1 << 18, 40, "fp64")]
    xs, seeds = [], []
    for i, (v, n, k, prec) in enumerate(reqs):
        x, _, _ = signals_device(n, k, [7000 + i], snr_db=30.0 if v == 2 else None)
        xs.append(x[0].contiguous())
        seeds.append((2000 + 11 * i, 300 + 13 * i))
    arrivals = np.cumsum(np.random.default_rng(9).exponential(1 / 2000.0, len(reqs)))
    import time
    with Server(workers=3, max_batch=8) as srv:
        for key in sorted({(n, k, v, p) for (v, n, k, p) in reqs}):
            srv.prepare(*key)
        t0 = time.perf_counter()
        tickets = []
        for i, (v, n, k, prec) in enumerate(reqs):
            dt = arrivals[i] - (time.perf_counter() - t0)
            if dt > 0:
                time.sleep(dt)
            tickets.append(srv.submit(xs[i], k, v, prec, seeds[i]))
        got = [srv.wait(t) for t in tickets]
        st = srv.stats()
    assert st["completed"] == len(reqs) and st["failed"] == 0
    assert st["max_concurrent_batches"] >= 2
    for i, (v, n, k, prec) in enumerate(reqs):
        print(f"case cfg5 v{v} n={n} k={k} {prec}", flush=True)
        f1, v1 = _alone(torch, v, n, k, prec, xs[i], seeds[i])
        assert np.array_equal(got[i][0].cpu().numpy(), f1)
        assert np.array_equal(got[i][1].cpu().numpy(), v1)
    # oracle checks of the last four (small n): same filters, same draws
    for i in range(len(reqs) - 4, len(reqs)):
        v, n, k, prec = reqs[i]
        if prec != "fp64":
            continue
        plan = sfft_amd.sfft(n, k, v, strict=False, seed=seeds[i])
        t1, r1 = plan.filter(0)
        t2, r2 = plan.filter(1)
        x = xs[i].cpu().numpy()
        oracle.seed_glibc(*seeds[i])
        fr = got[i][0].cpu().numpy()
        va = got[i][1].cpu().numpy()
        if v == 3:
            P = oracle.V3(n, k)
            P.set_filters(t1, r1, t2, r2)
            ref = P.exec(x, P.draw())
            if ref["early"]:
                assert len(fr) == 0
                continue
            assert np.array_equal(fr, ref["freqs"])
            assert np.abs(va - ref["vals"]).max() <= 1e-9 * max(1.0, np.abs(ref["vals"]).max())
        else:
            P = oracle.V12(n, k, True)
            P.set_filters(t1, r1, t2, r2)
            a, ai, off = P.draw()
            tr = P.exec(x, a, ai, off, trace=True)
            assert np.array_equal(np.unique(fr), np.unique(tr["hits"][:tr["hits_prefill"]]))
            assert np.array_equal(va, tr["out"][fr])
