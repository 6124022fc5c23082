"""GPU parity of the sFFT v1/v2 path against the CPU oracle, through the C ABI.

Two levels:
  * bit-exact: the oracle runs with the library's own filter taps, so every
    later step (comb, fold, FFT, selection, votes, median estimation) must
    agree bit for bit with the GPU FP64 form;
  * independent: the oracle designs its own filters; located sets must be
    identical and values within 1e-9 relative (filter designs agree to
    ~1e-12).
The FP32+refine form must give the identical located set and values within
1e-6 of the FP64 form.
"""
import numpy as np
import pytest

import oracle
import sfft_amd
from sfft_amd.signals import signal_np

pytestmark = pytest.mark.gpu

# (version, n, k, snr_db)
CASES = [
    (1, 1 << 14, 50, None),
    (2, 1 << 14, 50, None),
    (1, 1 << 16, 50, None),
    (2, 1 << 16, 50, None),
    (1, 1 << 20, 50, None),     # v1 at the noiseless n=2^20, k=50 workload
    (2, 1 << 22, 64, 40.0),     # one signal of the batched n=2^22, k=64 workload
    (2, 1 << 22, 100, 30.0),    # v2 n=2^22, k=100, SNR 30 dB
    (1, 1 << 18, 300, 20.0),    # v1 on the default parameters (k > 50)
]

# Transforms and selections above one workgroup's LDS: bucket FFTs as LDS
# radix-2 runs plus global passes, comb FFTs likewise, v1 selection over an
# HBM scratch array.
LARGE = [
    (1, 1 << 22, 1500, 30.0),   # B = 16384: one global pass (FP64)
    (1, 1 << 24, 2000, 30.0),   # B = 32768: two passes, selection in HBM
    (1, 1 << 24, 50, None),     # B_loc = 16384, B_est = 1024
    (2, 1 << 24, 10, 30.0),     # W_comb = 2^18: comb FFT with five passes
    (2, 1 << 24, 1000, 30.0),   # v2 with B = 16384 (the comb prefill
                                # locates 2000 * n/W = every frequency)
]


def _gpu_plan(version, n, k, precision="fp64"):
    return sfft_amd.sfft(n, k, version, strict=False, precision=precision,
                         seed=(17, 4242))


def _oracle_run(version, n, k, x, filters=None):
    oracle.seed_glibc(17, 4242)
    P = oracle.V12(n, k, version == 2)
    if filters is not None:
        P.set_filters(*filters)
    a, ai, off = P.draw()
    return P.exec(x, a, ai, off, trace=True)


def _plan_filters(plan):
    """The taps and responses the plan itself uses (for the oracle)."""
    lt, lf = plan.filter(0)
    et, ef = plan.filter(1)
    return lt, lf, et, ef


def _located(ref, version):
    if version == 2:
        return np.unique(ref["hits"][:ref["hits_prefill"]])
    return np.unique(ref["hits"])


def _run_gpu_dense(torch, plan, x):
    xd = torch.from_numpy(x).cuda()
    od = torch.empty_like(xd)
    plan.execute_device(xd, od)
    torch.cuda.synchronize()
    return od.cpu().numpy()


@pytest.mark.parametrize("version,n,k,snr", CASES + LARGE)
def test_fp64_bit_exact_given_filters(torch_cuda, version, n, k, snr):
    print(f"case v{version} n={n} k={k} snr={snr}", flush=True)
    x, f, v = signal_np(n, k, seed=1000 + n % 97 + k, snr_db=snr)
    plan = _gpu_plan(version, n, k)
    ref = _oracle_run(version, n, k, x, _plan_filters(plan))
    out = _run_gpu_dense(torch_cuda, plan, x)
    loc_ref = _located(ref, version)
    loc_gpu = np.nonzero(out)[0]
    assert np.array_equal(loc_gpu, loc_ref), (len(loc_gpu), len(loc_ref))
    diff = np.abs(out - ref["out"]).max()
    assert diff == 0.0, diff


@pytest.mark.parametrize("version,n,k,snr", CASES + LARGE[:4])
def test_fp64_buckets_bit_exact_given_filters(torch_cuda, version, n, k, snr):
    """Fold + bucket FFT output of every loop equals the oracle's x_samp."""
    print(f"case v{version} n={n} k={k} snr={snr}", flush=True)
    x, f, v = signal_np(n, k, seed=4000 + k, snr_db=snr)
    plan = _gpu_plan(version, n, k)
    ref = _oracle_run(version, n, k, x, _plan_filters(plan))
    plan.execute_sparse_host(x)
    got = plan.debug_buckets(0)
    assert len(got) == len(ref["x_samp"])
    assert np.abs(got - ref["x_samp"]).max() == 0.0


@pytest.mark.parametrize("version,n,k,snr", CASES)
def test_fp64_parity_independent_oracle(torch_cuda, version, n, k, snr):
    print(f"case v{version} n={n} k={k} snr={snr}", flush=True)
    x, f, v = signal_np(n, k, seed=2000 + k, snr_db=snr)
    ref = _oracle_run(version, n, k, x)
    plan = _gpu_plan(version, n, k)
    fr, va = plan.execute_sparse_host(x)
    loc_ref = _located(ref, version)
    assert np.array_equal(np.unique(fr), loc_ref)
    # values: norm-wise within 1e-9 of the largest coefficient (leakage-level
    # values carry the ~1e-12 difference of the two filter designs), and
    # every coefficient above 0.1 within 1e-9 relative
    refv = ref["out"][fr]
    diff = np.abs(va - refv)
    assert diff.max() <= 1e-9 * np.abs(refv).max(), diff.max()
    big = np.abs(refv) > 0.1
    assert (diff[big] / np.abs(refv[big])).max(initial=0.0) <= 1e-9
    # every true tone is located and equals conj(X) (reference semantics)
    if snr is None:
        assert np.all(np.isin(f, fr))
        got = dict(zip(fr.tolist(), va.tolist()))
        err = max(abs(got[int(ff)] - np.conj(vv)) for ff, vv in zip(f, v))
        assert err < 1e-6, err


@pytest.mark.parametrize("version,n,k,snr",
                         [c for c in CASES if c[1] >= 1 << 16] + LARGE[:4])
def test_fp32_refine_matches_fp64(torch_cuda, version, n, k, snr):
    print(f"case v{version} n={n} k={k} snr={snr}", flush=True)
    x, f, v = signal_np(n, k, seed=3000 + k, snr_db=snr)
    p64 = _gpu_plan(version, n, k)
    p32 = _gpu_plan(version, n, k, precision="fp32_refine")
    f64, v64 = p64.execute_sparse_host(x)
    f32, v32 = p32.execute_sparse_host(x)
    # v2 locates through the FP64 comb and v1 keeps its location loops in
    # FP64, so the located sets are identical in both forms
    assert np.array_equal(f32, f64)
    err = np.abs(v32 - v64)
    assert err.max() <= 1e-6, err.max()


@pytest.mark.parametrize("version,n,k,snr", [
    (2, 1 << 16, 50, None),      # 128 rows: one chunk (plain ascending tap order)
    (2, 1 << 22, 64, 40.0),      # 2048 rows: eight chunks of 256
    (1, 1 << 20, 50, None),      # v1: FP32 estimation loops only (B_est = 512)
    (2, 1 << 20, 64, 30.0),
    (1, 1 << 24, 50, None),      # B_est = 1024: 64 chunks
])
def test_fp32_buckets_bit_exact_model(torch_cuda, version, n, k, snr):
    """The FP32 form's buckets after the bucket FFTs equal, bit for bit, the
    model of its summation order (oracle/fp32_model.c) on the plan's taps
    and draws."""
    print(f"case fp32 buckets v{version} n={n} k={k} snr={snr}", flush=True)
    x, f, v = signal_np(n, k, seed=4300 + k, snr_db=snr)
    plan = _gpu_plan(version, n, k, precision="fp32_refine")
    plan.execute_sparse_host(x)
    got = plan.debug_buckets(0)
    p = dict(zip(["B_loc", "B_est", "B_thresh", "W_comb", "comb_loops", "loc_loops",
                  "est_loops", "thr_loops", "w_loc", "w_est"], plan.params()))
    L, Lloc = p["loc_loops"] + p["est_loops"], p["loc_loops"]
    oracle.seed_glibc(17, 4242)
    a, ai, off = oracle.v12_draws(n, L, p["comb_loops"] if version == 2 else 0, p["W_comb"])
    nmax = max(p["B_loc"], p["B_est"], p["W_comb"] if version == 2 else 0)
    lt, _ = plan.filter(0)
    et, _ = plan.filter(1)
    first32 = 0 if version == 2 else Lloc  # v1 keeps its location loops in FP64
    off_l = 0
    for l in range(L):
        B = p["B_loc"] if l < Lloc else p["B_est"]
        if l >= first32:
            want = oracle.fp32_buckets(x, n, B, lt if l < Lloc else et, ai[l], 256, nmax)
            assert np.array_equal(got[off_l:off_l + B], want), (l, np.abs(got[off_l:off_l + B] - want).max())
        off_l += B


@pytest.mark.parametrize("version", [1, 2])
def test_batch_equals_single(torch_cuda, version):
    """exec_many of a batch gives each signal the same result as alone,
    when the RNG draws are replayed in signal order."""
    torch = torch_cuda
    n, k = 1 << 18, 64
    xs = [signal_np(n, k, seed=s, snr_db=30.0)[0] for s in range(6)]
    plan = _gpu_plan(version, n, k)
    xd = [torch.from_numpy(x).cuda() for x in xs]
    fr, va = plan.make_sparse_outputs(len(xs))
    counts = plan.execute_many_sparse(xd, fr, va)
    torch.cuda.synchronize()
    single = _gpu_plan(version, n, k)
    for i, x in enumerate(xs):
        f1, v1 = single.execute_sparse_host(x)
        c = counts[i]
        assert c == len(f1)
        assert np.array_equal(fr[i, :c].cpu().numpy(), f1)
        assert np.abs(va[i, :c].cpu().numpy() - v1).max() == 0.0


@pytest.mark.parametrize("version", [1, 2])
def test_host_pointers_dense(torch_cuda, version):
    """sfft_exec with host in/out (the reference's own calling convention):
    the dense, zero-filled output equals the oracle's bit for bit."""
    n, k = 1 << 16, 50
    x, f, v = signal_np(n, k, seed=5)
    plan = sfft_amd.sfft(n, k, version, seed=(17, 4242))
    out = plan.execute(x)
    ref = _oracle_run(version, n, k, x, _plan_filters(plan))
    assert np.array_equal(out, ref["out"])


def test_host_pointers_dense_v3(torch_cuda):
    """v3 through sfft_exec with host in/out: the oracle's located
    frequencies carry its values, every other entry is zero."""
    n, k = 1 << 18, 100
    x, f, v = signal_np(n, k, seed=6)
    plan = sfft_amd.sfft(n, k, 3, seed=(17, 4242))
    t1, f1 = plan.filter(0)
    t2, f2 = plan.filter(1)
    oracle.seed_glibc(17, 4242)
    P = oracle.V3(n, k)
    P.set_filters(t1, f1, t2, f2)
    ref = P.exec(x, P.draw())
    out = plan.execute(x)
    if ref["early"]:
        assert not np.any(out)
        return
    assert np.array_equal(np.nonzero(out)[0], np.sort(ref["freqs"][ref["vals"] != 0]))
    assert np.abs(out[ref["freqs"]] - ref["vals"]).max() <= 1e-9 * np.abs(ref["vals"]).max()


@pytest.mark.parametrize("precision", ["fp64", "fp32_refine"])
@pytest.mark.parametrize("version,n,k,snr", [(2, 1 << 20, 64, 30.0), (1, 1 << 18, 100, 30.0)])
def test_schedules_give_identical_results(torch_cuda, version, n, k, snr, precision):
    """Fused and streamed folds (any resident-block count) sum every bucket
    in the same tap order, and the alias-row bucket layout (schedule 1, v2
    FP32) feeds estimation the same products: identical bits."""
    print(f"case v{version} n={n} k={k} {precision}", flush=True)
    x, f, v = signal_np(n, k, seed=6000 + k, snr_db=snr)
    res = []
    for mode, blocks in [(0, 0), (1, 0), (1, 96), (1, 1024), (2, 0)]:
        plan = _gpu_plan(version, n, k, precision)
        plan.set_schedule(mode, blocks)
        assert plan.schedule() == (mode, blocks)
        res.append(plan.execute_sparse_host(x))
    for fr, va in res[1:]:
        assert np.array_equal(fr, res[0][0])
        assert np.array_equal(va, res[0][1])


@pytest.mark.parametrize("precision", ["fp64", "fp32_refine"])
def test_unconjugated_values(torch_cuda, precision):
    """With the reference's conjugation undone, every schedule (the alias-row
    estimation of schedule 1 and the natural-order one of schedule 2) gives
    the same bits, the located set and the real parts are the conjugated
    form's, and each tone's value is the conjugate of the default one up to
    the choice between the two middle estimates of its imaginary part."""
    n, k = 1 << 20, 64
    print(f"case v2 n={n} k={k} {precision} unconjugated", flush=True)
    x, f, v = signal_np(n, k, seed=6100, snr_db=30.0)
    base = _gpu_plan(2, n, k, precision).execute_sparse_host(x)
    res = []
    for mode in (1, 2):
        plan = _gpu_plan(2, n, k, precision)
        plan.set_reference_conjugate(False)
        plan.set_schedule(mode, 0)
        res.append(plan.execute_sparse_host(x))
    assert np.array_equal(res[0][0], res[1][0]) and np.array_equal(res[0][1], res[1][1])
    fr, va = res[0]
    assert np.array_equal(fr, base[0])
    assert np.array_equal(va.real, base[1].real)
    tone = np.isin(fr, f)
    assert tone.sum() == k
    assert np.abs(va[tone] - np.conj(base[1][tone])).max() < 0.05


def test_autotune_keeps_draws_and_results(torch_cuda):
    n, k = 1 << 20, 64
    x, f, v = signal_np(n, k, seed=77, snr_db=30.0)
    tuned = _gpu_plan(2, n, k, "fp32_refine")
    mode, blocks = tuned.autotune(batch=8)
    assert mode in (0, 1, 2) and blocks >= 0
    fresh = _gpu_plan(2, n, k, "fp32_refine")
    a = tuned.execute_sparse_host(x)
    b = fresh.execute_sparse_host(x)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


@pytest.mark.parametrize("version,n,k,snr,precision", [
    (2, 1 << 20, 64, 30.0, "fp32_refine"), (1, 1 << 18, 100, 30.0, "fp64")])
def test_graph_replays_match_eager(torch_cuda, version, n, k, snr, precision):
    """Replays of the captured batch (new signals and new draws each time,
    eager torch work in between) equal the eager execs exactly."""
    torch = torch_cuda
    print(f"case graphs v{version} n={n} k={k} {precision}", flush=True)
    g = _gpu_plan(version, n, k, precision)
    g.enable_graphs(True)
    e = _gpu_plan(version, n, k, precision)
    for rep in range(4):
        xs = [torch.from_numpy(signal_np(n, k, seed=700 + 3 * rep + i, snr_db=snr)[0]).cuda()
              for i in range(3)]
        fg, vg = g.make_sparse_outputs(3)
        fe, ve = e.make_sparse_outputs(3)
        cg = g.execute_many_sparse(xs, fg, vg)
        _ = torch.stack(xs).abs().sum().item()   # eager work between replays
        ce = e.execute_many_sparse(xs, fe, ve)
        torch.cuda.synchronize()
        assert np.array_equal(cg, ce)
        for i in range(3):
            c = int(ce[i])
            assert torch.equal(fg[i, :c], fe[i, :c])
            assert torch.equal(vg[i, :c], ve[i, :c])


@pytest.mark.parametrize("version,n,k,snr", [(1, 1 << 18, 100, 30.0), (2, 1 << 20, 64, 30.0)])
def test_stage_timing_keeps_results(torch_cuda, version, n, k, snr):
    """Per-stage HIP-event timing: the stages that ran report time, and the
    timed execs give the untimed results."""
    print(f"case timing v{version} n={n} k={k}", flush=True)
    x, f, v = signal_np(n, k, seed=4100 + k, snr_db=snr)
    plain = _gpu_plan(version, n, k)
    timed = _gpu_plan(version, n, k)
    timed.enable_timing(True)
    for _ in range(3):
        a = plain.execute_sparse_host(x)
        b = timed.execute_sparse_host(x)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    st = timed.stage_times()
    timed.enable_timing(False)
    ran = ["fold", "estimate"] + (["select_vote"] if version == 1 else ["comb"])
    for name in ran:
        assert st[name][1] > 0 and st[name][0] > 0, (name, st)


@pytest.mark.parametrize("precision", ["fp64", "fp32_refine"])
def test_large_batch_tiles_equal_single(torch_cuda, precision):
    """A batch large enough for the wide estimation tiles (16 signals of
    n=2^22, k=64) gives every signal the bits it gets alone, where the
    narrow tiles run."""
    torch = torch_cuda
    from sfft_amd.signals import signals_device
    n, k, S = 1 << 22, 64, 16
    print(f"case wide tiles n={n} k={k} S={S} {precision}", flush=True)
    x, _, _ = signals_device(n, k, list(range(500, 500 + S)), snr_db=40.0)
    batch = _gpu_plan(2, n, k, precision)
    fr, va = batch.make_sparse_outputs(S)
    counts = batch.execute_many_sparse(x, fr, va)
    single = _gpu_plan(2, n, k, precision)
    f1, v1 = single.make_sparse_outputs(1)
    for i in range(S):
        c1 = single.execute_many_sparse(x[i:i + 1], f1, v1)
        torch.cuda.synchronize()
        c = int(counts[i])
        assert int(c1[0]) == c
        assert torch.equal(fr[i, :c], f1[0, :c])
        assert torch.equal(va[i, :c], v1[0, :c])
    del x
    torch.cuda.empty_cache()


def test_concurrent_calls_on_one_plan(torch_cuda):
    """Two host threads calling one plan take turns: each call gets the
    draws of one of the two sequential calls, and together they get both."""
    import threading
    torch = torch_cuda
    n, k = 1 << 18, 64
    x, f, v = signal_np(n, k, seed=2500, snr_db=30.0)
    seq = _gpu_plan(2, n, k)
    ref = [seq.execute_sparse_host(x) for _ in range(2)]
    shared = _gpu_plan(2, n, k)
    got = [None, None]

    def call(i):
        got[i] = shared.execute_sparse_host(x)
    ths = [threading.Thread(target=call, args=(i,)) for i in range(2)]
    [t.start() for t in ths]
    [t.join() for t in ths]
    torch.cuda.synchronize()

    def same(a, b):
        return np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    match = [[same(g, r) for r in ref] for g in got]
    assert (match[0][0] and match[1][1]) or (match[0][1] and match[1][0]), match


@pytest.mark.parametrize("n,k,snr", [(1 << 14, 50, None), (1 << 16, 50, 20.0),
                                     (1 << 20, 50, 30.0), (1 << 22, 64, 40.0)])
def test_v2_comb_votes_match_oracle(torch_cuda, n, k, snr):
    """The v2 location votes (skipped by the exec: they never change the
    output), counted on the GPU on demand, append the oracle's hits in the
    oracle's order (hits after the comb prefill; src/computefourier-1.0-2.0.cc
    125-184), and the output is unchanged by counting them."""
    print(f"case comb votes n={n} k={k} snr={snr}", flush=True)
    x, f, v = signal_np(n, k, seed=8800 + k, snr_db=snr)
    plan = _gpu_plan(2, n, k)
    plain = _gpu_plan(2, n, k)
    plan.debug_comb_votes(True)
    fr, va = plan.execute_sparse_host(x)
    f0, v0 = plain.execute_sparse_host(x)
    assert np.array_equal(fr, f0) and np.array_equal(va, v0)
    got = plan.comb_votes(0)
    ref = _oracle_run(2, n, k, x, _plan_filters(plan))
    appended = ref["hits"][ref["hits_prefill"]:ref["hits_found"]]
    assert ref["hits_found"] - ref["hits_prefill"] == len(got)
    assert np.array_equal(got, appended)
    # every appended frequency duplicates a prefill entry
    assert np.all(np.isin(appended, ref["hits"][:ref["hits_prefill"]]))


@pytest.mark.parametrize("precision", ["fp64", "fp32_refine"])
def test_many_plans_v12_mix_equals_alone(torch_cuda, precision):
    """sfft_exec_many_plans over v1/v2 plans of different (n, k), plus a v3
    plan, interleaved: the v1/v2 groups are enqueued back to back and
    waited for together, and every signal equals its plan's alone run
    (each plan draws for its own signals in order)."""
    torch = torch_cuda
    from sfft_amd.sfft import execute_many_plans
    keys = [(2, 1 << 18, 50), (1, 1 << 18, 20), (2, 1 << 20, 100), (3, 1 << 18, 100),
            (2, 1 << 16, 10)]
    order = [0, 1, 2, 3, 4, 2, 1, 0, 4, 3, 2]
    xs = [torch.from_numpy(signal_np(keys[g][1], keys[g][2], seed=700 + i,
                                     snr_db=None if keys[g][0] == 3 else 30.0)[0]).cuda()
          for i, g in enumerate(order)]

    def plan(g):
        v, n, k = keys[g]
        return sfft_amd.sfft(n, k, v, strict=False, precision=precision,
                             seed=(61 + g, 4242 + g))
    plans = [plan(g) for g in range(len(keys))]
    res = execute_many_plans([plans[g] for g in order], xs)
    torch.cuda.synchronize()
    for g, (v, n, k) in enumerate(keys):
        alone = plan(g)
        idx = [i for i, gg in enumerate(order) if gg == g]
        fr, va = alone.make_sparse_outputs(len(idx))
        counts = alone.execute_many_sparse([xs[i] for i in idx], fr, va)
        torch.cuda.synchronize()
        for j, i in enumerate(idx):
            print(f"case many-plans {precision} v{v} n={n} k={k} signal {i}", flush=True)
            c = int(counts[j])
            assert c > 0 and int(res[i][0].numel()) == c
            assert torch.equal(res[i][0], fr[j, :c])
            assert torch.equal(res[i][1], va[j, :c])


@pytest.mark.parametrize("graphs", [False, True])
def test_parted_batch_equals_single(torch_cuda, graphs):
    """A v2 FP32 batch of 97 signals of n=2^22, k=64 runs in two uneven
    parts (48 + 49: the fold and bucket FFTs of one part beside the
    estimation of the previous one, on two streams, eagerly or replayed from
    a HIP graph); every signal gets the bits it gets alone (a single signal
    runs unparted)."""
    torch = torch_cuda
    from sfft_amd.signals import signals_device
    n, k, S = 1 << 22, 64, 97
    print(f"case parted batch n={n} k={k} S={S} graphs={graphs}", flush=True)
    x, _, _ = signals_device(n, k, list(range(700, 700 + S)), snr_db=40.0)
    batch = _gpu_plan(2, n, k, "fp32_refine")
    batch.enable_graphs(graphs)
    fr, va = batch.make_sparse_outputs(S)
    for rep in range(2 if graphs else 1):   # capture, then a replay
        batch.seed(17, 4242)
        counts = batch.execute_many_sparse(x, fr, va)
        torch.cuda.synchronize()
    single = _gpu_plan(2, n, k, "fp32_refine")
    f1, v1 = single.make_sparse_outputs(1)
    for i in range(S):
        c1 = single.execute_many_sparse(x[i:i + 1], f1, v1)
        torch.cuda.synchronize()
        c = int(counts[i])
        assert int(c1[0]) == c
        assert torch.equal(fr[i, :c], f1[0, :c])
        assert torch.equal(va[i, :c], v1[0, :c])
    del x
    torch.cuda.empty_cache()
