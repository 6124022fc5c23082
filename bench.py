#!/usr/bin/env python3
"""Benchmark of the batched sFFT v2 hot path on one MI355X.

Workload: 256 signals, each n = 2^22 complex double, k = 64 unit-magnitude
random-phase tones plus complex Gaussian noise at SNR 40 dB (synthetic, built
on the GPU).  One step = one sfft_exec_many_sparse over the whole batch with
the inputs already resident in HBM.

Reported beside the headline (FP32 bucketization + FP64 estimation, batched):
  * the plain FP64 form: sfft_exec per signal into a dense, zero-filled
    device output -- the reference's calling convention on the GPU;
  * the FP64 batched form;
  * p50/p99 latency of single-signal execs;
  * accuracy: located-set diff and value diff FP32+refine vs FP64, and the
    per-coefficient L1 error against an FP64 full FFT of the signal;
  * roofline of the dominant kernel (fold + bucket FFT) from the library's
    own HIP events on its stream;
  * the serving loop under Poisson arrivals of mixed (n, k, version)
    requests: p50/p99 latency and throughput in both forms;
  * the CPU oracle timed on this host (OpenMP over signals) on a bounded
    sample.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("sFFT transforms/s + p50/p99 latency, n=2^22 k=64: "
          "FP32+refine vs FP64 GPU vs CPU")
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--n", type=int, default=1 << 22)
    ap.add_argument("--k", type=int, default=64)
    ap.add_argument("--snr", type=float, default=40.0)
    ap.add_argument("--lat-samples", type=int, default=200)
    ap.add_argument("--plain-signals", type=int, default=64)
    ap.add_argument("--cpu-signals", type=int, default=2048)
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving-loop run")
    ap.add_argument("--schedule", default=None,
                    help="fold schedule MODE,BLOCKS instead of autotuning (profiling runs)")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the v1 / v2 single-signal and batched v3 configurations")
    ap.add_argument("--serve-requests", type=int, default=160)
    ap.add_argument("--serve-rates", default="200,2000,0",
                    help="Poisson arrival rates of the serving runs, requests/s "
                         "(0: every request at once, the saturated throughput)")
    ap.add_argument("--serve-workers", type=int, default=4)
    ap.add_argument("--serve-batch", type=int, default=16)
    ap.add_argument("--only-configs", action="store_true",
                    help="run only the v1 / v2 single-signal and batched v3 configurations "
                         "(for profilers)")
    ap.add_argument("--only-uneven", action="store_true",
                    help="run only the v1/v2 uneven-(n, k) mix in one call (for profilers)")
    ap.add_argument("--only-fp32", action="store_true",
                    help="time only the headline form (for profilers)")
    ap.add_argument("--calib-copy", action="store_true",
                    help="first copy 1 GiB device-to-device (a kernel of known traffic "
                         "that scales the HBM byte counters)")
    return ap.parse_args()


TRAFFIC_PROFILE = "profiles/traffic_fold_chunks_v2_n22_k64_b256.md"


def measured_traffic(kernel, S, n, k):
    """HBM bytes per launch of the fold from the committed rocprofv3 counter
    runs of this kernel on this workload (FETCH_SIZE and WRITE_SIZE, scaled
    by a 1 GiB copy), and the file they come from; (None, None) for another
    kernel or workload."""
    if (S, n, k) != (256, 1 << 22, 64):
        return None, None
    key = "traffic_bytes_per_launch %s 256x2^22 k64:" % kernel
    try:
        with open(os.path.join(ROOT, TRAFFIC_PROFILE)) as fh:
            for line in fh:
                if line.startswith(key):
                    return float(line.split(":")[1]), TRAFFIC_PROFILE
    except OSError:
        pass
    return None, None


def copy_rate(torch, dev):
    """Device-to-device copy of 2 GiB (read + written bytes / time): the
    achievable HBM rate beside the vendor peak."""
    src = torch.empty(1 << 29, dtype=torch.float32, device=dev).fill_(1.0)
    dst = torch.empty_like(src)
    dst.copy_(src)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(5):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    gbs = 5 * 2 * src.numel() * 4 / (e0.elapsed_time(e1) * 1e-3) / 1e9
    del src, dst
    torch.cuda.empty_cache()
    return gbs


def log(msg):
    """Progress of the run on stderr (the JSON result is the only stdout line)."""
    print("bench: %s" % msg, file=sys.stderr, flush=True)


def pct(a, q):
    return float(np.percentile(np.asarray(a), q))


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; --gpus %d is not supported "
              "(batches shard by signal across GPUs with no exchange)"
              % args.gpus, file=sys.stderr)
        sys.exit(2)
    for v in ("RANK", "WORLD_SIZE"):
        if int(os.environ.get(v, "0") or 0) > (0 if v == "RANK" else 1):
            print("bench.py: multi-rank launch is not supported", file=sys.stderr)
            sys.exit(2)
    import torch
    import sfft_amd
    from sfft_amd.signals import signals_device

    n, k, S = args.n, args.k, args.batch
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    if args.only_configs:
        print(json.dumps({"configs": other_configs(torch, dev)}), flush=True)
        return
    if args.only_uneven:
        print(json.dumps({"v12_uneven_mix": uneven_v12_mix(torch, dev)}), flush=True)
        return
    if args.calib_copy:
        src = torch.ones(1 << 28, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)  # reads and writes exactly 1 GiB
        torch.cuda.synchronize()
        del src, dst
    seeds = list(range(1000, 1000 + S))
    x, fs, vs = signals_device(n, k, seeds, snr_db=args.snr, device=dev)
    xs = [x[i] for i in range(S)]
    torch.cuda.synchronize()

    # ---------------------------------------------------------- headline
    log("signals ready; headline plan")
    p32 = sfft_amd.sfft(n, k, 2, strict=False, precision="fp32_refine", seed=(17, 4242))
    if args.schedule:  # fold schedule only; same results
        p32.set_schedule(*[int(v) for v in args.schedule.split(",")])
        schedule32 = p32.schedule()
    else:
        schedule32 = p32.autotune(batch=min(S, 64))
    cap = p32.capacity
    fr32, va32 = p32.make_sparse_outputs(S)
    # headline: the batch replayed from a captured HIP graph (uploads, comb
    # beside the fold on a second stream, fold, bucket FFTs, estimation,
    # count read-back); the first exec of the shape captures it
    p32.enable_graphs(True)
    for _ in range(args.warmup):
        p32.execute_many_sparse(x, fr32, va32)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        counts32 = p32.execute_many_sparse(x, fr32, va32)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    p32.enable_graphs(False)
    ms_step = (t1 - t0) * 1e3 / args.steps
    tps32 = S * args.steps / (t1 - t0)
    # the same batch launched eagerly, with HIP events on the library's
    # streams around each stage (the roofline's kernel times)
    p32.execute_many_sparse(x, fr32, va32)
    torch.cuda.synchronize()
    p32.enable_timing(True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        counts32 = p32.execute_many_sparse(x, fr32, va32)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    stages32 = p32.stage_times()
    p32.enable_timing(False)
    tps32_eager = S * args.steps / (t1 - t0)

    result = {
        "metric": METRIC, "value": tps32, "unit": "transforms/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32+fp64", "data": "synthetic",
        "eager": {"value": tps32_eager, "unit": "transforms/s",
                  "ms_per_step": S / tps32_eager * 1e3,
                  "form": "the same batch launched eagerly (stage events on)"},
        "config": {"workload": "sFFT v2 batched exec_many_sparse replayed from a HIP graph, "
                               "FP32 buckets + FP64 estimation, %d signals, n=2^%d, k=%d, "
                               "SNR %g dB" % (S, int(np.log2(n)), k, args.snr),
                   "n": n, "k": k, "batch": S, "version": 2, "snr_db": args.snr,
                   "params": dict(zip(["B_loc", "B_est", "B_thresh", "W_comb",
                                       "comb_loops", "loc_loops", "est_loops",
                                       "thr_loops", "w_loc", "w_est"], p32.params())),
                   "fold_schedule": {"mode": schedule32[0], "resident_blocks": schedule32[1]}},
    }

    # ---------------------------------------------------------- roofline
    prm = result["config"]["params"]
    L = prm["loc_loops"] + prm["est_loops"]
    w_sum = prm["loc_loops"] * prm["w_loc"] + prm["est_loops"] * prm["w_est"]
    xs_size = prm["loc_loops"] * prm["B_loc"] + prm["est_loops"] * prm["B_est"]
    # per launch (one batch): gathered samples 16 B each, FP32 taps once,
    # FP32 bucket writes
    fold_bytes = S * (16 * w_sum + 8 * xs_size) + 8 * (prm["w_loc"] + prm["w_est"])
    fold_ms, fold_launches = stages32["fold"]
    fold_avg_ms = fold_ms / max(1, fold_launches)
    achieved = fold_bytes / (fold_avg_ms * 1e-3) / 1e9
    kname = "k_fold_chunks"
    traffic, traffic_src = measured_traffic(kname, S, n, k)
    result["roofline"] = {"kernel": kname, "bound": "hbm", "achieved": achieved,
                          "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                          "traffic": traffic, "traffic_source": traffic_src,
                          "alg_bytes_per_launch": fold_bytes,
                          "avg_launch_ms": fold_avg_ms, "launches": fold_launches}
    if traffic:  # line traffic the counters show, over this run's launch time
        result["roofline"]["traffic_gbs"] = traffic / (fold_avg_ms * 1e-3) / 1e9
        result["roofline"]["traffic_frac"] = result["roofline"]["traffic_gbs"] / HBM_PEAK_GBS
    result["roofline"]["copy_gbs"] = copy_rate(torch, dev)
    log("headline %.0f transforms/s (graph), %.0f eager" % (tps32, tps32_eager))
    result["stages_ms_per_step"] = {s: v[0] / args.steps for s, v in stages32.items() if v[1]}

    if not args.only_fp32:
        # ------------------------------------------------------ latency
        one_f, one_v = fr32[:1], va32[:1]

        def latency():
            lat = []
            p32.execute_many_sparse([xs[0]], one_f, one_v)  # warm (captures a graph)
            for i in range(args.lat_samples):
                torch.cuda.synchronize()
                ts = time.perf_counter()
                p32.execute_many_sparse([xs[i % S]], one_f, one_v)
                lat.append((time.perf_counter() - ts) * 1e3)
            return lat
        lat_eager = latency()
        p32.enable_graphs(True)
        lat = latency()
        p32.enable_graphs(False)
        result["latency_ms"] = {"p50": pct(lat, 50), "p99": pct(lat, 99),
                                "samples": len(lat),
                                "form": "fp32+refine, 1 signal, replayed HIP graph",
                                "eager_p50": pct(lat_eager, 50), "eager_p99": pct(lat_eager, 99)}

        log("latency done")
        # ------------------------------------------------------ FP64 forms
        p64 = sfft_amd.sfft(n, k, 2, strict=False, precision="fp64", seed=(17, 4242))
        p64.autotune(batch=min(S, 64))
        fr64, va64 = p64.make_sparse_outputs(S)
        p64.execute_many_sparse(xs, fr64, va64)  # warm-up
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(max(1, args.steps // 2)):
            counts64 = p64.execute_many_sparse(xs, fr64, va64)
        torch.cuda.synchronize()
        tps64b = S * max(1, args.steps // 2) / (time.perf_counter() - t0)

        # same draws for both forms: replay from the same seed
        p64.seed(17, 4242)
        p32.seed(17, 4242)
        counts64 = p64.execute_many_sparse(xs, fr64, va64)
        counts32 = p32.execute_many_sparse(xs, fr32, va32)
        torch.cuda.synchronize()
        set_diff, vmax, vrel, l1_sum, l1_cnt = 0, 0.0, 0.0, 0.0, 0
        for i in range(S):
            c32, c64 = int(counts32[i]), int(counts64[i])
            a, b = fr32[i, :c32], fr64[i, :c64]
            if c32 != c64 or not torch.equal(a, b):
                set_diff += 1
                continue
            dv = (va32[i, :c32] - va64[i, :c64]).abs()
            vmax = max(vmax, float(dv.max()))
            vrel = max(vrel, float(dv.max() / va64[i, :c64].abs().max().clamp_min(1e-300)))
            l1_sum += float(dv.sum())
            l1_cnt += c32
        # signal 0 with the first draws of seed (17, 4242), for the CPU
        # oracle's located-set check in the CPU leg
        c0 = int(counts64[0])
        gpu_check = {"x": xs[0].cpu().numpy(), "filters": (p64.filter(0), p64.filter(1)),
                     "freqs": fr64[0, :c0].cpu().numpy(), "vals": va64[0, :c0].cpu().numpy(),
                     "freqs32": fr32[0, :int(counts32[0])].cpu().numpy()}

        # plain FP64 form: one sfft_exec per signal into a dense device output
        plain = sfft_amd.sfft(n, k, 2, strict=False, precision="fp64", seed=(17, 4242))
        dense = torch.empty(n, dtype=torch.complex128, device=dev)
        plain.execute_device(xs[0], dense)
        torch.cuda.synchronize()
        lat64 = []
        P = min(args.plain_signals, S)
        t0 = time.perf_counter()
        for i in range(P):
            ts = time.perf_counter()
            plain.execute_device(xs[i], dense)
            lat64.append((time.perf_counter() - ts) * 1e3)
        torch.cuda.synchronize()
        tps_plain = P / (time.perf_counter() - t0)
        result["fp64_plain"] = {"value": tps_plain, "unit": "transforms/s",
                                "p50_ms": pct(lat64, 50), "p99_ms": pct(lat64, 99),
                                "signals": P, "form": "sfft_exec per signal, dense zero-filled "
                                                      "device output (reference convention)"}
        result["fp64_batched"] = {"value": tps64b, "unit": "transforms/s"}
        result["speedup_fp32_vs_fp64_plain"] = tps32 / tps_plain
        result["speedup_fp32_vs_fp64_batched"] = tps32 / tps64b

        log("fp64 forms done")
        # ------------------------------------------------------ accuracy
        # FP64 full FFT of the signal as the oracle; the reference's v1/v2
        # values are conj(X) (src/computefourier-1.0-2.0.cc:383-398), so
        # compare with conj.
        l1_loc, l1_true, off_energy = [], [], []
        for i in range(min(4, S)):
            Xf = torch.fft.fft(xs[i]) / n
            c = int(counts32[i])
            f_idx = fr32[i, :c].long()
            l1_loc.append(float((va32[i, :c] - Xf[f_idx].conj()).abs().mean()))
            # energy the output puts off the true support, relative to the
            # support's energy
            on = torch.isin(f_idx, torch.as_tensor(np.asarray(fs[i]), device=dev).long())
            e_off = float((va32[i, :c][~on].abs() ** 2).sum())
            e_sup = float((Xf[torch.as_tensor(np.asarray(fs[i]), device=dev).long()].abs() ** 2).sum())
            off_energy.append(e_off / max(e_sup, 1e-300))
            pos = {int(ff): j for j, ff in enumerate(fr32[i, :c].cpu().numpy())}
            tv = [abs(complex(va32[i, pos[int(ff)]].item()) - np.conj(complex(Xf[int(ff)].item())))
                  for ff in fs[i] if int(ff) in pos]
            l1_true.append(float(np.mean(tv)) if tv else float("nan"))
        result["accuracy"] = {
            "located_set_diff_fp32_vs_fp64": set_diff,
            "value_maxabs_fp32_vs_fp64": vmax,
            "value_max_rel_to_peak_fp32_vs_fp64": vrel,
            "off_support_energy_ratio": float(np.mean(off_energy)),
            "l1_fp32_vs_fp64_per_coefficient": l1_sum / max(1, l1_cnt),
            "l1_vs_fp64_fft_located_set": float(np.mean(l1_loc)),
            "l1_vs_fp64_fft_true_support": float(np.mean(l1_true)),
            "located_per_signal": int(counts32[0]),
        }

    # ---------------------------------------------------------- other configs
    if not args.only_fp32 and not args.no_configs:
        del fr32, va32
        if "fr64" in locals():
            del fr64, va64
        torch.cuda.empty_cache()
        log("other configurations")
        result["configs"] = other_configs(torch, dev)

    # ---------------------------------------------------------- serving
    if not args.no_serve and not args.only_fp32:
        torch.cuda.empty_cache()
        log("serving")
        result["serving"] = serving(args, torch, dev)

    # ---------------------------------------------------------- CPU
    if not args.no_cpu and not args.only_fp32:
        log("cpu baseline")
        result["cpu_baseline"] = cpu_baseline(args, locals().get("gpu_check"))
    print(json.dumps(result), flush=True)


def uneven_v12_mix(torch, dev, reps=10):
    """v1/v2 plans of uneven (n, k) in one sfft_exec_many_plans call: 12
    keys (n = 2^18 ... 2^22, k = 10 ... 300), 4 signals each (SNR 30 dB,
    FP32+refine); the call runs each plan's group of kernels in turn on one
    stream.  Timed beside the sum of each group run alone
    (execute_many_sparse of its 4 signals), and checked equal to those
    alone-runs signal for signal."""
    import sfft_amd
    from sfft_amd.sfft import ManyPlansCall, execute_many_plans
    from sfft_amd.signals import signals_device
    keys = [(2, 1 << 18, 10), (2, 1 << 18, 50), (2, 1 << 20, 20), (2, 1 << 20, 100),
            (2, 1 << 20, 300), (2, 1 << 22, 10), (2, 1 << 22, 64), (2, 1 << 22, 200),
            (1, 1 << 18, 20), (1, 1 << 20, 10), (1, 1 << 20, 50), (1, 1 << 22, 30)]
    per = 4
    plans = [sfft_amd.sfft(n, k, v, strict=False, precision="fp32_refine", seed=(31 + i, 4242))
             for i, (v, n, k) in enumerate(keys)]
    xs, order, groups = [], [], []
    for i, (v, n, k) in enumerate(keys):
        x, _, _ = signals_device(n, k, [5000 + 10 * i + j for j in range(per)], snr_db=30.0,
                                 device=dev)
        groups.append(x)
        for j in range(per):
            xs.append(x[j])
            order.append(plans[i])
    outs = [(torch.empty(p.capacity, dtype=torch.int32, device=dev),
             torch.empty(p.capacity, dtype=torch.complex128, device=dev)) for p in order]
    call = ManyPlansCall(order, xs, outs)  # pointer arrays built once
    call.call()  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        call.call()
    torch.cuda.synchronize()
    ms_mix = (time.perf_counter() - t0) * 1e3 / reps
    # the same groups alone, and the results of one seeded call of each
    gouts = [p.make_sparse_outputs(per) for p in plans]
    for p, x, (fr, va) in zip(plans, groups, gouts):
        p.execute_many_sparse(x, fr, va)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        for p, x, (fr, va) in zip(plans, groups, gouts):
            p.execute_many_sparse(x, fr, va)
    torch.cuda.synchronize()
    ms_alone = (time.perf_counter() - t0) * 1e3 / reps
    for i, p in enumerate(plans):
        p.seed(31 + i, 4242)
    counts = execute_many_plans(order, xs, outs, trim=False)
    for i, p in enumerate(plans):
        p.seed(31 + i, 4242)
    same = True
    for i, (p, x, (fr, va)) in enumerate(zip(plans, groups, gouts)):
        c = p.execute_many_sparse(x, fr, va)
        for j in range(per):
            m = i * per + j
            cm = int(counts[m])
            same &= cm == int(c[j]) and torch.equal(outs[m][0][:cm], fr[j, :cm]) and \
                torch.equal(outs[m][1][:cm], va[j, :cm])
    torch.cuda.synchronize()
    return {"plans": len(keys), "signals": len(xs), "keys": [list(k_) for k_ in keys],
            "ms_per_call": ms_mix, "transforms_per_s": len(xs) / ms_mix * 1e3,
            "sum_of_groups_alone_ms": ms_alone, "identical_to_alone": bool(same),
            "form": "sfft_exec_many_plans, one group of kernels per plan on one stream "
                    "(FP32+refine, SNR 30 dB)"}


def other_configs(torch, dev):
    """The reference's other workloads, measured beside the headline:
    plain v1 (n=2^20, k=50, noiseless) and v2 (n=2^22, k=100, SNR 30 dB)
    single signals in the FP64 form, with per-coefficient L1 error against
    the FP64 FFT of the signal; and batched v3, 64 signals of n=2^24 with k
    cycling over {200, 500, 1000} (noiseless) in one multi-plan launch."""
    import sfft_amd
    from sfft_amd.sfft import ManyPlansCall, execute_many_plans
    from sfft_amd.signals import signals_device
    out = {}
    for name, (ver, n, k, snr) in {"v1_n20_k50": (1, 1 << 20, 50, None),
                                   "v2_n22_k100_snr30": (2, 1 << 22, 100, 30.0)}.items():
        x, fs, vs = signals_device(n, k, [77], snr_db=snr, device=dev)
        p64 = sfft_amd.sfft(n, k, ver, strict=False, precision="fp64", seed=(17, 4242))
        p32 = sfft_amd.sfft(n, k, ver, strict=False, precision="fp32_refine", seed=(17, 4242))
        d = torch.empty(n, dtype=torch.complex128, device=dev)
        p64.execute_device(x[0], d)
        torch.cuda.synchronize()
        reps = 20
        t0 = time.perf_counter()
        for _ in range(reps):
            p64.execute_device(x[0], d)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / reps
        f64, v64 = p64.make_sparse_outputs(1)
        f32, v32 = p32.make_sparse_outputs(1)
        p64.seed(17, 4242)
        p32.seed(17, 4242)
        c64 = int(p64.execute_many_sparse(x, f64, v64)[0])
        c32 = int(p32.execute_many_sparse(x, f32, v32)[0])
        Xf = torch.fft.fft(x[0]) / n
        loc = f64[0, :c64].long()
        l1_loc = float((v64[0, :c64] - Xf[loc].conj()).abs().mean())
        tf = torch.from_numpy(fs[0]).to(dev)
        got = torch.zeros(n, dtype=torch.complex128, device=dev)
        got[loc] = v64[0, :c64]
        l1_true = float((got[tf] - Xf[tf].conj()).abs().mean())
        same = c32 == c64 and torch.equal(f32[0, :c32], f64[0, :c64])
        out[name] = {"fp64_ms_per_transform": ms, "located": c64,
                     "l1_vs_fp64_fft_located": l1_loc, "l1_vs_fp64_fft_true_support": l1_true,
                     "located_set_diff_fp32_vs_fp64": 0 if same else 1,
                     "value_maxabs_fp32_vs_fp64": float((v32[0, :c32] - v64[0, :c64]).abs().max())
                     if same else None}
        del x, d, Xf, got
        torch.cuda.empty_cache()
    out["v12_uneven_mix"] = uneven_v12_mix(torch, dev)
    torch.cuda.empty_cache()
    # batched v3, uneven k in one launch
    n, S = 1 << 24, 64
    ks = [200, 500, 1000]
    plans = {kk: sfft_amd.sfft(n, kk, 3, seed=(17 + kk, 4242)) for kk in ks}
    xs, truth = [], []
    for i in range(S):
        kk = ks[i % 3]
        x, fs, vs = signals_device(n, kk, [9000 + i], device=dev)
        xs.append(x[0])
        truth.append((fs[0], vs[0]))
    order = [plans[ks[i % 3]] for i in range(S)]
    outs = [(torch.empty(p.capacity, dtype=torch.int32, device=dev),
             torch.empty(p.capacity, dtype=torch.complex128, device=dev)) for p in order]
    call = ManyPlansCall(order, xs, outs)  # pointer arrays built once
    call.call()  # warm-up
    torch.cuda.synchronize()
    reps = 20
    t0 = time.perf_counter()
    for _ in range(reps):
        call.call()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    # oracle parity of the first entry of each plan: the plans' first draws
    # again (re-seeded), the oracle on the plans' filters
    import oracle
    for kk in ks:
        plans[kk].seed(17 + kk, 4242)
    counts0 = execute_many_plans(order, xs, outs, trim=False)
    torch.cuda.synchronize()
    parity = {}
    for j, kk in enumerate(ks):
        t1, r1 = plans[kk].filter(0)
        t2, r2 = plans[kk].filter(1)
        P = oracle.V3(n, kk, bare=True)
        P.set_filters(t1, r1, t2, r2)
        oracle.seed_glibc(17 + kk, 4242)
        ref = P.exec(xs[j].cpu().numpy(), P.draw())
        c = int(counts0[j])
        fr, va = outs[j][0][:c].cpu().numpy(), outs[j][1][:c].cpu().numpy()
        same = (not ref["early"] and np.array_equal(fr, ref["freqs"])) or (ref["early"] and c == 0)
        parity[str(kk)] = {"located_set_equal": bool(same),
                           "value_maxabs_vs_oracle": float(np.abs(va - ref["vals"]).max())
                           if same and c else 0.0, "located": c}
        del P, t1, r1, t2, r2
    res = [(fr[:int(c)], va[:int(c)]) for (fr, va), c in zip(outs, counts0)]
    found, err = [], 0.0
    for (fr, va), (fs, vs) in zip(res, truth):
        got = dict(zip(fr.cpu().numpy().tolist(), va.cpu().numpy().tolist()))
        hit = [int(f) in got for f in fs]
        found.append(float(np.mean(hit)))
        err = max([err] + [abs(got[int(f)] - v) for f, v in zip(fs, vs) if int(f) in got])
    out["v3_n24_b64_k200_500_1000"] = {
        "transforms_per_s": S / dt, "ms_per_batch": dt * 1e3, "form": "fp64, one launch",
        "recovered_fraction_by_k": {str(kk): float(np.mean(found[j::3])) for j, kk in enumerate(ks)},
        "max_value_error_recovered": err,
        "oracle_parity_first_entry_by_k": parity}
    # the refine form on the same batch: same draws, same located sets, the
    # values re-estimated with integer-exact phases
    plans_r = {kk: sfft_amd.sfft(n, kk, 3, precision="fp32_refine", seed=(17 + kk, 4242))
               for kk in ks}
    order_r = [plans_r[ks[i % 3]] for i in range(S)]
    outs_r = [(torch.empty(p.capacity, dtype=torch.int32, device=dev),
               torch.empty(p.capacity, dtype=torch.complex128, device=dev)) for p in order_r]
    call_r = ManyPlansCall(order_r, xs, outs_r)
    call_r.call()  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        call_r.call()
    torch.cuda.synchronize()
    dt_r = (time.perf_counter() - t0) / reps
    for kk in ks:
        plans_r[kk].seed(17 + kk, 4242)
    counts_r = execute_many_plans(order_r, xs, outs_r, trim=False)
    torch.cuda.synchronize()
    same_sets, err_r, err_all, kept = True, 0.0, 0.0, 0
    flags = {kk: plans_r[kk].last_status() for kk in ks}
    for i in range(S):
        c, cf = int(counts_r[i]), int(counts0[i])
        same_sets &= c == cf and torch.equal(outs_r[i][0][:c], outs[i][0][:cf])
        fs, vs = truth[i]
        got = dict(zip(outs_r[i][0][:c].cpu().numpy().tolist(),
                       outs_r[i][1][:c].cpu().numpy().tolist()))
        err_r = max([err_r] + [abs(got[int(f)] - v) for f, v in zip(fs, vs) if int(f) in got])
        tv = dict(zip(np.asarray(fs).tolist(), np.asarray(vs).tolist()))
        err_all = max([err_all] + [abs(g - tv.get(f, 0)) for f, g in got.items()])
    for kk in ks:
        kept += sum(1 for fl in flags[kk] if fl & 8)
    out["v3_n24_b64_k200_500_1000_refined"] = {
        "transforms_per_s": S / dt_r, "ms_per_batch": dt_r * 1e3,
        "form": "fp64 location + values re-estimated by least squares over fresh Mansour "
                "buckets at 8 strided offsets, integer-exact phases, Prony step for up to "
                "2 unlocated tones per bucket (precision fp32_refine)",
        "located_sets_equal_to_faithful": bool(same_sets),
        "max_value_error_recovered": err_r,
        "max_value_error_all_located": err_all,
        "signals_keeping_some_reference_values": kept}
    return out


def serving(args, torch, dev):
    """Poisson arrivals of single-signal requests: n uniform on
    {2^18, 2^20, 2^22, 2^24}, k = floor(10 * 200^U), v2 (SNR 30 dB) or v3
    (noiseless) 50/50; v2 draws the plan cannot run (W_Comb < 2k+1) are
    rejected and counted.  Inputs and plans are device-resident before the
    arrivals start.  The same trace runs at several arrival rates (rate 0:
    every request submitted at once, the saturated throughput) in the
    FP32+refine and the FP64 forms, each form on one server whose plans stay
    resident across the loads.  Latency = submit to the end of the request's
    batch.  A sample of the results is compared with the same requests run
    alone (exact equality)."""
    from concurrent.futures import ThreadPoolExecutor

    import sfft_amd
    from sfft_amd import SfftError
    from sfft_amd.server import Server, output_capacity
    from sfft_amd.signals import signals_device

    rng = np.random.default_rng(2024)
    N = args.serve_requests
    sizes = [1 << 18, 1 << 20, 1 << 22, 1 << 24]
    reqs = []
    for i in range(N):
        n = int(sizes[rng.integers(4)])
        k = int(np.floor(10 * 200 ** rng.random()))
        v = 2 if rng.random() < 0.5 else 3
        reqs.append((n, k, v))
    unit = np.cumsum(rng.exponential(1.0, N))  # arrival times at rate 1
    caps = []
    for (n, k, v) in reqs:
        try:
            caps.append(output_capacity(n, k, v))
        except SfftError:
            caps.append(None)
    dummy = {}
    xs, outs = [], []
    for i, (n, k, v) in enumerate(reqs):
        if caps[i] is None:
            if n not in dummy:
                dummy[n] = torch.zeros(n, dtype=torch.complex128, device=dev)
            xs.append(dummy[n])
            outs.append(None)
            continue
        x, _, _ = signals_device(n, k, [5000 + i], snr_db=30.0 if v == 2 else None, device=dev)
        xs.append(x[0])
        outs.append((torch.empty(caps[i], dtype=torch.int32, device=dev),
                     torch.empty(caps[i], dtype=torch.complex128, device=dev)))
    torch.cuda.synchronize()
    keys = sorted({reqs[i] for i in range(N) if caps[i] is not None})
    feasible = [i for i in range(N) if caps[i] is not None]
    checked = feasible[::max(1, len(feasible) // 8)][:8]
    out = {"requests": N, "workers": args.serve_workers, "max_batch": args.serve_batch,
           "distinct_plans": len(keys), "rejected_infeasible": sum(c is None for c in caps),
           "mix": "n uniform {2^18,2^20,2^22,2^24}, k=floor(10*200^U), v2 SNR 30 dB / v3 "
                  "noiseless 50/50"}
    loads = [float(r) for r in args.serve_rates.split(",")]
    for prec in ("fp32_refine", "fp64"):
        res = {}
        with Server(workers=args.serve_workers, max_batch=args.serve_batch) as srv:
            t0 = time.perf_counter()
            with ThreadPoolExecutor(8) as ex:
                list(ex.map(lambda key: srv.prepare(key[0], key[1], key[2], prec), keys))
            prep_s = time.perf_counter() - t0
            for rate in loads:
                s0 = srv.stats()
                at = unit / rate if rate > 0 else np.zeros(N)
                tickets, rejected = [], 0
                start = time.perf_counter()
                for i, (n, k, v) in enumerate(reqs):
                    dt = at[i] - (time.perf_counter() - start)
                    if dt > 0:
                        time.sleep(dt)
                    try:
                        tickets.append((i, srv.submit(xs[i], k, v, prec, (i + 1, 7 * i + 3),
                                                      outputs=outs[i])))
                    except SfftError:
                        rejected += 1
                got = {i: srv.wait(t) for i, t in tickets}
                span = time.perf_counter() - start
                s1 = srv.stats()
                lat = [g[2] for g in got.values()]
                done = s1["completed"] - s0["completed"]
                nb = s1["batches"] - s0["batches"]
                res["saturated" if rate <= 0 else "%g_per_s" % rate] = {
                    "offered_rate_per_s": rate if rate > 0 else None,
                    "p50_ms": pct(lat, 50), "p99_ms": pct(lat, 99),
                    "mean_ms": float(np.mean(lat)), "completed": done,
                    "rejected": rejected, "failed": s1["failed"] - s0["failed"],
                    "transforms_per_s": done / span, "batches": nb,
                    "mean_batch": (s1["batched_requests"] - s0["batched_requests"]) / max(1, nb),
                    "max_concurrent_batches": s1["max_concurrent_batches"]}
            st = srv.stats()
        # the last load's results of a sample of requests against each run alone
        mism = 0
        for i in checked:
            n, k, v = reqs[i]
            fr, va, _ = got[i]
            plan = sfft_amd.sfft(n, k, v, strict=False, precision=prec, seed=(i + 1, 7 * i + 3))
            f1, v1 = plan.make_sparse_outputs(1)
            c1 = int(plan.execute_many_sparse([xs[i]], f1, v1)[0])
            torch.cuda.synchronize()
            mism += not (c1 == len(fr) and torch.equal(f1[0, :c1], fr) and
                         torch.equal(v1[0, :c1], va))
            del plan, f1, v1
        out[prec] = {"loads": res, "plan_prepare_s": prep_s,
                     "hbm_peak_gb": st["hbm_peak"] / 1e9,
                     "checked_vs_alone": len(checked), "mismatches_vs_alone": mism}
    return out


def oracle_check(check, n, k):
    """Located set and values of the FP64 GPU form on one signal against the
    oracle run on the plan's filters with the same draws (bit-exact by
    construction: same arithmetic, same order)."""
    import oracle
    (lt, lf), (et, ef) = check["filters"]
    oracle.seed_glibc(17, 4242)
    P = oracle.V12(n, k, True)
    P.set_filters(lt, lf, et, ef)
    a, ai, off = P.draw()
    tr = P.exec(check["x"], a, ai, off, trace=True)
    located = np.unique(tr["hits"][:tr["hits_prefill"]])
    fr = check["freqs"]
    return {"signal": "batch entry 0, first draws of seed (17, 4242)",
            "located": int(len(located)),
            "located_set_diff_fp64_vs_oracle": int(len(np.setxor1d(np.unique(fr), located))),
            "located_set_diff_fp32_vs_oracle":
                int(len(np.setxor1d(np.unique(check["freqs32"]), located))),
            "value_maxabs_fp64_vs_oracle": float(np.abs(check["vals"] - tr["out"][fr]).max())
            if len(fr) else 0.0}


def cpu_baseline(args, check=None):
    """The CPU oracle (restatement of the reference) on this host's cores."""
    import ctypes as C
    import oracle
    from sfft_amd.signals import signal_np

    n, k = args.n, args.k
    distinct = min(16, args.cpu_signals)
    threads = max(1, min(args.cpu_threads, os.cpu_count() or 1))
    P = oracle.V12(n, k, True)
    xs = [signal_np(n, k, 1000 + i, snr_db=args.snr)[0] for i in range(distinct)]
    outs = [np.empty(n, np.complex128) for _ in range(threads)]  # one per worker
    num = args.cpu_signals
    L = P.loops
    oracle.seed_glibc(17, 4242)
    a = np.zeros((num, L), np.int32)
    ai = np.zeros((num, L), np.int32)
    off = np.zeros((num, max(1, P.p.comb_loops)), np.int32)
    for i in range(num):
        a[i], ai[i], off[i] = P.draw()
    xp = np.array([xs[i % distinct].ctypes.data for i in range(num)], dtype=np.uint64)
    op = np.array([o.ctypes.data for o in outs], dtype=np.uint64)
    # the timing build (the reference's -O3 -ffast-math flags) is the
    # baseline; the parity build (-O2, no contraction) is timed beside it
    t0 = time.perf_counter()
    P.exec_many(xp, a, ai, off, op, -threads, fast=True)
    dt = time.perf_counter() - t0
    t0 = time.perf_counter()
    P.exec_many(xp, a, ai, off, op, -threads)
    dt_strict = time.perf_counter() - t0
    # 1-thread latency: sequential single-signal execs (first 20 draws)
    lat1 = []
    for i in range(min(20, num)):
        ts = time.perf_counter()
        P.exec_many(xp[i:i + 1], a[i:i + 1], ai[i:i + 1], off[i:i + 1], op[:1], -1, fast=True)
        lat1.append((time.perf_counter() - ts) * 1e3)
    # FFTW (the reference's FFT library) is not in this image: the
    # restatement uses its own FFT either way; say whether the host has it
    try:
        import ctypes.util
        fftw = ctypes.util.find_library("fftw3") or ""
        C.CDLL("libfftw3.so.3")
        fftw = fftw or "libfftw3.so.3"
    except OSError:
        fftw = ""
    cpu_model = ""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    cpu_model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    out = {"value": num / dt, "unit": "transforms/s", "cores": threads, "kind": "port",
            "sample": "%d execs (%d distinct signals) of v2 n=2^%d k=%d SNR %g dB; "
                      "reference semantics incl. the dense zero-filled output; "
                      "OpenMP over signals; %.1f s wall"
                      % (num, distinct, int(np.log2(n)), k, args.snr, dt),
            "build": "oracle/_build/libsfft_oracle_fast.so: -O3 -ffast-math -march=x86-64-v3 "
                     "-fopenmp (the reference's wscript:10 flags, x86-64-v3 for native)",
            "parity_build_value": num / dt_strict,
            "parity_build": "-O2 -ffp-contract=off (the bit-exact checker)",
            "fft": "own radix-2/Bluestein FP64 FFT" + (
                "; %s present on this host but not used" % fftw if fftw else
                "; libfftw3.so.3 not found on this host"),
            "host_cpus": os.cpu_count(), "cpu_model": cpu_model,
            "latency_1thread_ms": {"p50": pct(lat1, 50), "p99": pct(lat1, 99),
                                   "samples": len(lat1)}}
    if check is not None:
        out["parity"] = oracle_check(check, n, k)
    return out


if __name__ == "__main__":
    main()
