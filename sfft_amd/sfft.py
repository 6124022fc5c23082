"""Python host interface over libsfft_amd, mirroring the reference's ctypes
class (python/sfft/sfft.py:33-85): same constructor arguments, validation
and ``execute(a)`` returning the dense spectrum estimate.

GPU-resident use goes through ``execute_device`` / ``execute_many_sparse``
with torch tensors (torch only provides device memory and the stream).
"""
import ctypes as C

import numpy as np

from . import _lib

# flags copied from fftw3.h (python/sfft/sfft.py:6-8)
FFTW_MEASURE = 0
FFTW_ESTIMATE = 1 << 6

# python/sfft/sfft.py:10-31: the (n, k) pairs the reference's wrapper allows
# for v1/v2.
V1_V2_INPUT_PARAMETERS = [
    {"size": 8192, "sparsity": 50}, {"size": 16384, "sparsity": 50},
    {"size": 32768, "sparsity": 50}, {"size": 65536, "sparsity": 50},
    {"size": 131072, "sparsity": 50}, {"size": 262144, "sparsity": 50},
    {"size": 524288, "sparsity": 50}, {"size": 1048576, "sparsity": 50},
    {"size": 2097152, "sparsity": 50}, {"size": 4194304, "sparsity": 50},
    {"size": 8388608, "sparsity": 50}, {"size": 16777216, "sparsity": 50},
    {"size": 4194304, "sparsity": 50}, {"size": 4194304, "sparsity": 100},
    {"size": 4194304, "sparsity": 200}, {"size": 4194304, "sparsity": 500},
    {"size": 4194304, "sparsity": 1000}, {"size": 4194304, "sparsity": 2000},
    {"size": 4194304, "sparsity": 2500}, {"size": 4194304, "sparsity": 4000},
]

PRECISIONS = {"fp64": _lib.SFFT_PRECISION_FP64,
              "fp32_refine": _lib.SFFT_PRECISION_FP32_REFINE}


class SfftError(RuntimeError):
    pass


def derive_params(n, k, version):
    """Parameters libsfft_amd derives for (n, k, version 1..3), host only.

    Returns (ok, dict); ok is False when sfft_make_plan would reject it.
    """
    p = (C.c_int * 10)()
    r = _lib.load().sfft_derive_params(n, k, version - 1, p)
    if r < 0:
        raise ValueError(_lib.last_error())
    if version == 3:
        keys = ["W_Man", "B_g1", "B_g2", "w_g1", "w_g2"]
    else:
        keys = ["B_loc", "B_est", "B_thresh", "W_comb", "comb_loops",
                "loc_loops", "est_loops", "thr_loops", "w_loc", "w_est"]
    d = {kk: p[i] for i, kk in enumerate(keys)}
    return r == 0, d


def design_filter(lobefrac, tolerance, n, b):
    """The library's host filter design: (time taps, n-point response)."""
    w_cap = int((1 / np.pi) * (1 / lobefrac) * np.arccosh(1. / tolerance)) + 2
    t = np.zeros(w_cap, np.complex128)
    f = np.zeros(n, np.complex128)
    w = _lib.load().sfft_design_filter(lobefrac, tolerance, n, b,
                                       t.ctypes.data_as(C.c_void_p), w_cap,
                                       f.ctypes.data_as(C.c_void_p))
    if w < 0:
        raise SfftError(_lib.last_error())
    return t[:w].copy(), f


class sfft:
    """sFFT plan (python/sfft/sfft.py:33-85).

    ``strict`` keeps the reference wrapper's (n, k) whitelist for v1/v2;
    with ``strict=False`` every (n, k) the library accepts is allowed.
    """

    def __init__(self, length=16384, sparsity=50, version=1,
                 optimization=FFTW_ESTIMATE, strict=True, precision="fp64",
                 seed=None):
        for name, v in (("length", length), ("sparsity", sparsity),
                        ("version", version), ("optimization", optimization)):
            if not isinstance(v, (int, np.integer)):
                raise TypeError("%s is not an integer" % name)
        if version not in [1, 2, 3]:
            raise ValueError("sFFT version %d is not valid.  Try 1, 2, or 3." % (version))
        if strict and version in [1, 2] and \
                {"size": length, "sparsity": sparsity} not in V1_V2_INPUT_PARAMETERS:
            raise ValueError("n = %d and k = %d is not a valid input parameter combination "
                             "for sFFT version %d." % (length, sparsity, version))
        if optimization not in [FFTW_MEASURE, FFTW_ESTIMATE]:
            raise ValueError("FFTW optimization %d is not valid." % (optimization))
        self.length = int(length)
        self.sparsity = int(sparsity)
        self.version = int(version)
        self.optimization = optimization
        self._L = _lib.load()
        self.sfft_plan = self._L.sfft_make_plan(self.length, self.sparsity,
                                                self.version - 1, FFTW_ESTIMATE)
        if not self.sfft_plan:
            self.sfft_plan = None
            raise SfftError("sfft_make_plan(%d, %d, v%d) failed: %s"
                            % (length, sparsity, version, _lib.last_error()))
        self.set_precision(precision)
        if seed is not None:
            self.seed(*seed)
        self._stream = None
        if optimization == FFTW_MEASURE:  # tune the schedule at this precision
            self.autotune(64)

    # ------------------------------------------------------------ reference
    def execute(self, a):
        """Dense spectrum estimate of one host signal (numpy in, numpy out)."""
        a = np.require(np.asanyarray(a), np.complex128, ['C', 'ALIGNED'])
        if a.shape != (self.length,):
            raise ValueError("input must have shape (%d,)" % self.length)
        b = np.empty_like(a)
        self._L.sfft_exec(self.sfft_plan, a.ctypes.data_as(C.c_void_p),
                          b.ctypes.data_as(C.c_void_p))
        self._check()
        return b

    def __del__(self):
        if getattr(self, "sfft_plan", None):
            self._L.sfft_free_plan(self.sfft_plan)
            self.sfft_plan = None

    # ------------------------------------------------------------ extensions
    def _check(self):
        e = _lib.last_error()
        if e:
            raise SfftError(e)

    def set_precision(self, precision):
        code = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        if self._L.sfft_plan_set_precision(self.sfft_plan, code) != 0:
            raise SfftError(_lib.last_error())
        self.precision = precision

    def set_schedule(self, fold_mode, stream_blocks=0):
        """Fold schedule (0 fused, 1 streamed with stream_blocks resident
        workgroups); never changes results."""
        if self._L.sfft_plan_set_schedule(self.sfft_plan, fold_mode, stream_blocks) != 0:
            raise SfftError(_lib.last_error())

    def schedule(self):
        m, b = C.c_int(0), C.c_int(0)
        self._L.sfft_plan_get_schedule(self.sfft_plan, C.byref(m), C.byref(b))
        return m.value, b.value

    def enable_graphs(self, enabled=True):
        """Replay v1/v2 execs from captured HIP graphs (identical results)."""
        self._L.sfft_plan_enable_graphs(self.sfft_plan, 1 if enabled else 0)

    def autotune(self, batch=64):
        """Times the fold schedules on synthetic signals, keeps the fastest."""
        if self._L.sfft_plan_autotune(self.sfft_plan, batch) != 0:
            raise SfftError(_lib.last_error())
        return self.schedule()

    def seed(self, random_seed=17, drand48_seed=4242):
        """Private RNG streams equal to srandom(random_seed)/srand48(..)."""
        self._L.sfft_plan_seed(self.sfft_plan, random_seed, drand48_seed)

    def set_reference_conjugate(self, enabled):
        self._L.sfft_plan_set_reference_conjugate(self.sfft_plan, int(bool(enabled)))

    def params(self):
        p = (C.c_int * 10)()
        self._L.sfft_plan_params(self.sfft_plan, p)
        return list(p)

    @property
    def capacity(self):
        return int(self._L.sfft_plan_output_capacity(self.sfft_plan))

    def memory(self):
        a, b = C.c_int64(), C.c_int64()
        self._L.sfft_plan_memory(self.sfft_plan, C.byref(a), C.byref(b))
        return a.value, b.value

    def set_stream(self, stream):
        """Run this plan's execs on a torch.cuda.Stream (sfft_plan_set_stream)."""
        import torch
        self._bind_stream(torch, stream.cuda_stream)

    def _bind_stream(self, torch, s=None):
        if s is None:
            s = torch.cuda.current_stream().cuda_stream
        if s != self._stream:
            self._L.sfft_plan_set_stream(self.sfft_plan, C.c_void_p(s))
            self._stream = s

    def execute_device(self, x, out):
        """sfft_exec on device tensors: x, out complex128 of length n."""
        import torch
        assert x.is_cuda and out.is_cuda and x.dtype == torch.complex128
        assert out.dtype == torch.complex128 and x.numel() == out.numel() == self.length
        self._bind_stream(torch)
        self._L.sfft_exec(self.sfft_plan, C.c_void_p(x.data_ptr()),
                          C.c_void_p(out.data_ptr()))
        self._check()
        return out

    def execute_many_device(self, xs, outs):
        """sfft_exec_many on lists of device tensors (dense outputs)."""
        import torch
        self._bind_stream(torch)
        num = len(xs)
        ins = (C.c_void_p * num)(*[t.data_ptr() for t in xs])
        ous = (C.c_void_p * num)(*[t.data_ptr() for t in outs])
        self._L.sfft_exec_many(self.sfft_plan, num, ins, ous)
        self._check()
        return outs

    def make_sparse_outputs(self, num, device="cuda"):
        """Device buffers for execute_many_sparse: (freq[num,cap], val[num,cap])."""
        import torch
        cap = self.capacity
        freq = torch.empty((num, cap), dtype=torch.int32, device=device)
        val = torch.empty((num, cap), dtype=torch.complex128, device=device)
        return freq, val

    def execute_many_sparse(self, xs, freq, val):
        """sfft_exec_many_sparse: xs list of device signals (or a 2-D tensor),
        results into freq/val [num, cap]; returns the located counts."""
        import torch
        self._bind_stream(torch)
        num = len(xs)
        rows = np.arange(num, dtype=np.uint64)
        if isinstance(xs, (list, tuple)):
            ins = np.array([t.data_ptr() for t in xs], dtype=np.uint64)
        else:
            ins = np.uint64(xs.data_ptr()) + rows * np.uint64(xs.stride(0) * xs.element_size())
        # sfft_output_c records {count, capacity, freq, val} built as one
        # array: per-row ctypes field writes cost about a microsecond each
        outs = np.zeros((num, 4), dtype=np.uint64)
        outs[:, 1] = freq.shape[1]
        outs[:, 2] = np.uint64(freq.data_ptr()) + rows * np.uint64(freq.stride(0) * freq.element_size())
        outs[:, 3] = np.uint64(val.data_ptr()) + rows * np.uint64(val.stride(0) * val.element_size())
        r = self._L.sfft_exec_many_sparse(self.sfft_plan, num, ins.ctypes.data_as(C.c_void_p),
                                          outs.ctypes.data_as(C.POINTER(_lib.SfftOutput)))
        if r != 0:
            raise SfftError(_lib.last_error())
        return outs[:, 0].astype(np.int64)

    def filter(self, which):
        """The plan's filter `which` (0 location / Gauss-1, 1 estimation /
        Gauss-2): (taps, full n-point response with zeros outside the
        offsets the exec reads)."""
        half = C.c_int(0)
        w = self._L.sfft_plan_filter(self.sfft_plan, which, None, 0, None, 0, C.byref(half))
        if w < 0:
            raise SfftError(_lib.last_error())
        t = np.zeros(w, np.complex128)
        r = np.zeros(2 * half.value + 1, np.complex128)
        self._L.sfft_plan_filter(self.sfft_plan, which, t.ctypes.data_as(C.c_void_p), w,
                                 r.ctypes.data_as(C.c_void_p), len(r), C.byref(half))
        full = np.zeros(self.length, np.complex128)
        d = np.arange(-half.value, half.value + 1)
        full[d % self.length] = r
        return t, full

    def last_status(self):
        """v3 flags per signal of the last exec call: 1 early return,
        2 map overflow, 4 loop stopped at the pass cap, 8 the refine form
        kept the reference's values in some bucket, 16 a wait for a bucket
        task gave up (a defect; the result is not valid)."""
        m = self._L.sfft_plan_last_status(self.sfft_plan, None, 0)
        if m < 0:
            raise SfftError(_lib.last_error())
        flags = (C.c_int * max(m, 1))()
        self._L.sfft_plan_last_status(self.sfft_plan, flags, m)
        return [flags[i] for i in range(m)]

    STAGES = ["comb", "fold", "bucket_fft", "select_vote", "estimate", "v3"]

    def enable_timing(self, enabled=True):
        self._L.sfft_plan_enable_timing(self.sfft_plan, int(bool(enabled)))

    def stage_times(self):
        """{stage: (total_ms, launches)} from the library's HIP events."""
        ms = (C.c_double * 6)()
        ln = (C.c_int64 * 6)()
        self._L.sfft_plan_stage_times(self.sfft_plan, ms, ln)
        return {s: (ms[i], ln[i]) for i, s in enumerate(self.STAGES)}

    def debug_buckets(self, signal=0, count=1 << 30):
        """Bucket values (after the bucket FFTs) of the last v1/v2 batch."""
        cnt = min(count, 1 << 26)
        buf = np.zeros(cnt, np.complex128)
        m = self._L.sfft_debug_buckets(self.sfft_plan, signal,
                                       buf.ctypes.data_as(C.c_void_p), cnt)
        if m < 0:
            raise SfftError(_lib.last_error())
        return buf[:m].copy()

    def debug_v3(self, serial_subtractions=False, map_limit=0):
        """v3 plans (diagnostics): every subtraction pass serial, and/or the
        map of located frequencies capped at map_limit entries."""
        if self._L.sfft_plan_debug_v3(self.sfft_plan, int(bool(serial_subtractions)),
                                      int(map_limit)) != 0:
            raise SfftError(_lib.last_error())

    def debug_v3_inline_pre(self, enabled=True):
        """v3 plans (diagnostic): each transform runs its own pre tasks."""
        if self._L.sfft_plan_debug_v3_inline_pre(self.sfft_plan, int(bool(enabled))) != 0:
            raise SfftError(_lib.last_error())

    def debug_comb_votes(self, enabled=True):
        """v2 FP64 plans: count the location votes in later execs."""
        if self._L.sfft_plan_debug_comb_votes(self.sfft_plan, int(bool(enabled))) != 0:
            raise SfftError(_lib.last_error())

    def comb_votes(self, signal=0):
        """Frequencies the reference's v2 location votes append to its hit
        list, in its order (src/computefourier-1.0-2.0.cc:125-184), for
        signal `signal` of the last batch."""
        m = self._L.sfft_debug_comb_votes(self.sfft_plan, signal, None, 0)
        if m < 0:
            raise SfftError(_lib.last_error())
        buf = np.zeros(max(1, m), np.int32)
        self._L.sfft_debug_comb_votes(self.sfft_plan, signal, buf.ctypes.data_as(C.c_void_p), m)
        return buf[:m].copy()

    def execute_sparse_host(self, x):
        """One host signal, sparse result on the host: (freqs, values)."""
        x = np.require(np.asanyarray(x), np.complex128, ['C', 'ALIGNED'])
        cap = self.capacity
        fr = np.zeros(cap, np.int32)
        va = np.zeros(cap, np.complex128)
        o = _lib.SfftOutput(0, cap, fr.ctypes.data, va.ctypes.data)
        r = self._L.sfft_exec_sparse(self.sfft_plan, x.ctypes.data_as(C.c_void_p),
                                     C.byref(o))
        if r != 0:
            raise SfftError(_lib.last_error())
        m = min(o.count, cap)
        return fr[:m].copy(), va[:m].copy()


class ManyPlansCall:
    """A prepared sfft_exec_many_plans over fixed buffers: the plan, input
    and output pointer arrays are built once, so each call() is the C call
    alone (for loops that re-run one batch shape, e.g. a serving step or a
    benchmark).  The tensors must stay allocated while the object is used.
    call() returns the counts (int64 array) the call wrote."""

    def __init__(self, plans, xs, outputs=None, bind_streams=True):
        import torch
        num = len(xs)
        assert len(plans) == num
        if bind_streams:
            stream = torch.cuda.current_stream().cuda_stream
            for p in {id(p): p for p in plans}.values():
                p._bind_stream(torch, stream)
        if outputs is None:
            outputs = [(torch.empty(p.capacity, dtype=torch.int32, device=x.device),
                        torch.empty(p.capacity, dtype=torch.complex128, device=x.device))
                       for p, x in zip(plans, xs)]
        self.plans, self.xs, self.outputs, self.num = list(plans), list(xs), outputs, num
        self._pl = np.array([C.cast(p.sfft_plan, C.c_void_p).value for p in plans],
                            dtype=np.uint64)
        self._ins = np.array([x.data_ptr() for x in xs], dtype=np.uint64)
        # sfft_output_c records {count, capacity, freq, val}
        self._outs = np.zeros((num, 4), dtype=np.uint64)
        self._outs[:, 1] = [fr.numel() for fr, _ in outputs]
        self._outs[:, 2] = [fr.data_ptr() for fr, _ in outputs]
        self._outs[:, 3] = [va.data_ptr() for _, va in outputs]
        self._fn = _lib.load().sfft_exec_many_plans
        self._args = (num, self._pl.ctypes.data_as(C.c_void_p),
                      self._ins.ctypes.data_as(C.c_void_p),
                      self._outs.ctypes.data_as(C.POINTER(_lib.SfftOutput)))

    def call(self):
        if self._fn(*self._args) != 0:
            raise SfftError(_lib.last_error())
        return self._outs[:, 0].astype(np.int64)


def execute_many_plans(plans, xs, outputs=None, trim=True, bind_streams=True):
    """sfft_exec_many_plans: signal i of xs (device tensors) runs plans[i];
    the v3 signals of all plans run as one launch.  outputs: list of
    (freq, val) device tensors (allocated at each plan's capacity when
    None).  Every plan runs on torch's current stream, or with
    bind_streams=False on the stream it was given (set_stream).  Returns
    [(freq[:count], val[:count])], or with trim=False the counts (int64
    array) the call wrote, leaving the views to the caller."""
    m = ManyPlansCall(plans, xs, outputs, bind_streams)
    counts = m.call()
    if not trim:
        return counts
    return [(fr[:min(int(counts[i]), fr.numel())], va[:min(int(counts[i]), va.numel())])
            for i, (fr, va) in enumerate(m.outputs)]
