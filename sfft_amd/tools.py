"""Command-line tools with the reference harnesses' flags.

    python -m sfft_amd.tools timing        [-n N] [-k K] [-r R] [-o] [-v V]
    python -m sfft_amd.tools timing_many   [-n N] [-k K] [-i NUM] [-s] [-o] [-v V] [--check]
    python -m sfft_amd.tools verification  [-n N] [-k K] [-r R] [-o] [-v V]

timing (src/timing.cc with src/simulation.cc:32-92) times R sfft_exec
calls on one input and prints "TIME t" (seconds).  timing_many
(src/timing_many.cc) times sfft_exec_many over NUM inputs (-s: one plan per
input, each run with sfft_exec).  verification (src/verification.cc) checks
every input frequency above 0.1 is recovered within 0.1 and prints "OK"
(exit 0) or the first error (exit 1 missing, 2 wrong value).

Inputs come from the reference's generator (sfft_generate_input: X[f] = 1
at k frequencies drawn with replacement); the reference seeds it with
time ^ (pid * 171717), as here unless --seed is given.  Like the reference,
inputs and outputs are host arrays (the PCIe copies are timed); -d keeps
them in device memory instead, and -p fp32_refine selects the FP32 form.
"""
import argparse
import os
import sys
import time

import numpy as np

from . import _lib
from .sfft import FFTW_ESTIMATE, FFTW_MEASURE, PRECISIONS, SfftError, sfft
from .signals import reference_input


def _parser(prog, n, k, extra):
    ap = argparse.ArgumentParser(prog=prog, add_help=False)
    ap.add_argument("-n", type=int, default=n, help="problem size")
    ap.add_argument("-k", type=int, default=k, help="number of frequencies")
    ap.add_argument("-o", action="store_true", help="FFTW_MEASURE (autotune the plan)")
    ap.add_argument("-v", type=int, default=1, choices=[1, 2, 3], help="sFFT version")
    ap.add_argument("-h", action="help")
    ap.add_argument("-d", action="store_true", help="device-resident input and output")
    ap.add_argument("-p", default="fp64", choices=sorted(PRECISIONS), help="precision form")
    ap.add_argument("--seed", type=int, default=None, help="drand48 seed of the input")
    for a, kw in extra:
        ap.add_argument(a, **kw)
    return ap


def _seed(args):
    if args.seed is not None:
        return args.seed
    return (int(time.time()) ^ (os.getpid() * 171717)) & 0x7FFFFFFF


def _plan(args):
    opt = FFTW_MEASURE if args.o else FFTW_ESTIMATE
    try:
        return sfft(args.n, args.k, args.v, optimization=opt, strict=False, precision=args.p)
    except SfftError as e:
        print("error:", e, file=sys.stderr)
        sys.exit(1)


def _run(plan, x, d, torch):
    if d:
        out = torch.empty_like(x)
        plan.execute_device(x, out)
        torch.cuda.synchronize()
        return out
    return plan.execute(x)


def timing(argv):
    args = _parser("timing", 16384, 50, [("-r", dict(type=int, default=1))]).parse_args(argv)
    plan = _plan(args)
    x, _, _ = reference_input(args.n, args.k, _seed(args))
    torch = None
    if args.d:
        import torch
        x = torch.from_numpy(x).cuda()
    ts = time.perf_counter()
    for _ in range(args.r):
        _run(plan, x, args.d, torch)
    print("TIME", time.perf_counter() - ts)
    return 0


def timing_many(argv):
    args = _parser("timing_many", 1 << 18, 1 << 14, [
        ("-i", dict(type=int, default=100)), ("-s", dict(action="store_true")),
        ("--check", dict(action="store_true",
                         help="after timing, check every output as verification does"))
    ]).parse_args(argv)
    plans = [_plan(args) for _ in range(args.i if args.s else 1)]
    seed = _seed(args)
    ins = [reference_input(args.n, args.k, seed + t) for t in range(args.i)]
    xs = [x for x, _, _ in ins]
    L = _lib.load()
    import ctypes as C
    if args.d:
        import torch
        xd = [torch.from_numpy(x).cuda() for x in xs]
        outs = [torch.empty_like(x) for x in xd]
        ip = [t.data_ptr() for t in xd]
        op = [t.data_ptr() for t in outs]
    else:
        outs = [np.zeros(args.n, np.complex128) for _ in xs]
        ip = [x.ctypes.data for x in xs]
        op = [o.ctypes.data for o in outs]
    ts = time.perf_counter()
    if args.s:
        for i in range(args.i):
            L.sfft_exec(plans[i].sfft_plan, C.c_void_p(ip[i]), C.c_void_p(op[i]))
    else:
        L.sfft_exec_many(plans[0].sfft_plan, args.i, (C.c_void_p * args.i)(*ip),
                         (C.c_void_p * args.i)(*op))
    if args.d:
        torch.cuda.synchronize()
    print("TIME", time.perf_counter() - ts)
    if not args.check:
        return 0
    bad = 0
    for t, (_, X, _) in enumerate(ins):
        ans = outs[t].cpu().numpy() if args.d else outs[t]
        bad += _first_error(X, ans) is not None
    print("CHECK %d of %d outputs OK" % (args.i - bad, args.i))
    return 0 if bad == 0 else 3


def _first_error(X, ans):
    """verification's criterion (src/verification.cc:35-50): every input
    frequency above 0.1 recovered within 0.1; None, or (i, code)."""
    for i in np.nonzero(np.abs(X) > 0.1)[0]:
        if ans[i] == 0:
            return i, 1
        if abs(ans[i] - X[i]) > 0.1:
            return i, 2
    return None


def verification(argv):
    args = _parser("verification", 16384, 50, [("-r", dict(type=int, default=1))]).parse_args(argv)
    plan = _plan(args)
    x, X, _ = reference_input(args.n, args.k, _seed(args))
    ans = None
    for _ in range(args.r):
        ans = plan.execute(x)
    err = _first_error(X, ans)
    if err is not None:
        i, code = err
        if code == 1:
            print("ERROR: Frequency %d was not recovered!" % i)
        else:
            print("ERROR: Error of frequency %d is too big:\n  Expected: %s\n  Actual  : %s"
                  % (i, X[i], ans[i]))
        return code
    print("OK")
    return 0


TOOLS = {"timing": timing, "timing_many": timing_many, "verification": verification}


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    if not argv or argv[0] not in TOOLS:
        print(__doc__)
        return 1
    return TOOLS[argv[0]](argv[1:])


if __name__ == "__main__":
    sys.exit(main())
