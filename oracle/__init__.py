"""CPU oracle for the sFFT hot path -- TEST INFRASTRUCTURE ONLY.

ctypes binding of ``oracle/sfft_oracle.c``, a plain-C restatement of the
reference (ETH/MIT sFFT 0.1).  Only ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` import this package, and only as
the checker.  The product library never routes through it.

``ref_helpers()`` loads ``oracle/_ref/libsfft_ref_utils.so`` when it was built
from the reference's own ``src/utils.cc`` and ``src/parameters.cc``.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libsfft_oracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libsfft_ref_utils.so")

_lib = None
_ref = None


def build(with_ref=None):
    """Compile the oracle (and oracle/_ref when the reference tree exists)."""
    targets = ["all"]
    if with_ref is None:
        with_ref = os.path.isdir("/root/reference/src")
    if with_ref:
        targets.append("ref")
    subprocess.check_call(["make", "-s", "-C", HERE] + targets)


class V12Params(C.Structure):
    _fields_ = [("n", C.c_int), ("k", C.c_int), ("with_comb", C.c_int),
                ("Bcst_loc", C.c_double), ("Bcst_est", C.c_double),
                ("Comb_cst", C.c_double), ("tol_loc", C.c_double),
                ("tol_est", C.c_double),
                ("loc_loops", C.c_int), ("est_loops", C.c_int),
                ("thr_loops", C.c_int), ("comb_loops", C.c_int),
                ("BB_loc", C.c_double), ("BB_est", C.c_double),
                ("lobefrac_loc", C.c_double), ("lobefrac_est", C.c_double),
                ("b_loc", C.c_int), ("b_est", C.c_int),
                ("B_loc", C.c_int), ("B_est", C.c_int), ("B_thresh", C.c_int),
                ("W_comb", C.c_int), ("w_loc", C.c_int), ("w_est", C.c_int),
                ("x_samp_size", C.c_long)]


class V12Plan(C.Structure):
    _fields_ = [("p", V12Params),
                ("loc_time", C.c_void_p), ("loc_freq", C.c_void_p),
                ("est_time", C.c_void_p), ("est_freq", C.c_void_p)]


class V12Trace(C.Structure):
    _fields_ = [("x_sampt", C.c_void_p), ("x_samp", C.c_void_p),
                ("J", C.c_void_p), ("comb_approved", C.c_void_p),
                ("num_comb", C.c_int), ("hits", C.c_void_p),
                ("hits_cap", C.c_long), ("hits_found", C.c_long),
                ("hits_prefill", C.c_long)]


class V3Params(C.Structure):
    _fields_ = [("n", C.c_int), ("k", C.c_int), ("W_Man", C.c_int),
                ("B_g1", C.c_int), ("B_g2", C.c_int), ("w_g1", C.c_int),
                ("w_g2", C.c_int), ("b_g1", C.c_int), ("b_g2", C.c_int),
                ("BB1", C.c_double), ("BB2", C.c_double),
                ("lobefrac_g1", C.c_double), ("lobefrac_g2", C.c_double),
                ("tol_g1", C.c_double), ("tol_g2", C.c_double)]


class V3Plan(C.Structure):
    _fields_ = [("p", V3Params),
                ("time1", C.c_void_p), ("freq1", C.c_void_p),
                ("time2", C.c_void_p), ("freq2", C.c_void_p)]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build(with_ref=False)
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        L.or_gcd.argtypes = [C.c_int, C.c_int]
        L.or_mod_inverse.argtypes = [C.c_int, C.c_int]
        L.or_floor_to_pow2.argtypes = [C.c_double]
        L.or_find_largest_indices.argtypes = [P, C.c_int, P, C.c_int]
        L.or_sort_ints.argtypes = [P, C.c_int]
        L.or_fft.argtypes = [P, P, C.c_int, C.c_int]
        L.or_dolph_width.argtypes = [C.c_double, C.c_double]
        L.or_make_dolphchebyshev_t.argtypes = [C.c_double, C.c_double, C.c_int, P]
        L.or_make_multiple_t.argtypes = [P, C.c_int, C.c_int, C.c_int, P]
        L.or_v12_params_get.argtypes = [C.c_int, C.c_int, C.c_int, C.POINTER(V12Params)]
        L.or_v12_plan_make.argtypes = [C.c_int, C.c_int, C.c_int]
        L.or_v12_plan_make.restype = C.POINTER(V12Plan)
        L.or_v12_plan_free.argtypes = [C.POINTER(V12Plan)]
        L.or_v12_plan_set_filters.argtypes = [C.POINTER(V12Plan), P, P, P, P]
        L.or_v12_draw.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, P, P, P]
        L.or_v12_exec.argtypes = [C.POINTER(V12Plan), P, P, P, P, P, C.POINTER(V12Trace)]
        L.or_v3_params_get.argtypes = [C.c_int, C.c_int, C.POINTER(V3Params)]
        L.or_v3_plan_make.argtypes = [C.c_int, C.c_int]
        L.or_v3_plan_make.restype = C.POINTER(V3Plan)
        L.or_v3_plan_make_bare.argtypes = [C.c_int, C.c_int]
        L.or_v3_plan_make_bare.restype = C.POINTER(V3Plan)
        L.or_v3_plan_free.argtypes = [C.POINTER(V3Plan)]
        L.or_v3_plan_set_filters.argtypes = [C.POINTER(V3Plan), P, P, P, P]
        L.or_v3_draw.argtypes = [C.c_int, P]
        L.or_v3_exec.argtypes = [C.POINTER(V3Plan), P, P, P, P, P, P, C.c_long]
        L.or_v3_exec.restype = C.c_long
        L.or_v12_exec_many.argtypes = [C.POINTER(V12Plan), C.c_int, P, P, P, P, P, C.c_int]
        L.fm_fold32.argtypes = [P, C.c_int, C.c_int, C.c_int, P, C.c_uint, C.c_int, P]
        L.fm_fft32.argtypes = [P, C.c_int, C.c_int]
        _lib = L
    return _lib


FAST_PATH = os.path.join(HERE, "_build", "libsfft_oracle_fast.so")
_lib_fast = None


def lib_fast():
    """The timing build of the oracle (the reference's -O3 -ffast-math
    flags; not for parity): or_v12_exec_many only.  Plans made by lib()
    run on it (same source, same structures)."""
    global _lib_fast
    if _lib_fast is None:
        if not os.path.exists(FAST_PATH):
            build(with_ref=False)
        L = C.CDLL(FAST_PATH)
        P = C.c_void_p
        L.or_v12_exec_many.argtypes = [C.POINTER(V12Plan), C.c_int, P, P, P, P, P, C.c_int]
        _lib_fast = L
    return _lib_fast


def ref_helpers():
    """The reference's own utils/parameters (oracle/_ref), or None."""
    global _ref
    if _ref is None and os.path.exists(REF_PATH):
        R = C.CDLL(REF_PATH)
        P = C.c_void_p
        R.ref_gcd.argtypes = [C.c_int, C.c_int]
        R.ref_mod_inverse.argtypes = [C.c_int, C.c_int]
        R.ref_floor_to_pow2.argtypes = [C.c_double]
        R.ref_find_largest_indices.argtypes = [P, C.c_int, P, C.c_int, P]
        R.ref_radix_sort.argtypes = [P, C.c_int]
        R.ref_get_params.argtypes = [C.c_int, C.c_int, C.c_int, P, P]
        _ref = R
    return _ref


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


# ---------------------------------------------------------------- helpers

def fft(x, backward=False):
    x = np.ascontiguousarray(x, dtype=np.complex128)
    out = np.empty_like(x)
    lib().or_fft(_p(out), _p(x), len(x), int(backward))
    return out


def find_largest_indices(samples, num):
    s = np.ascontiguousarray(samples, dtype=np.float64)
    out = np.empty(num, dtype=np.int32)
    lib().or_find_largest_indices(_p(out), num, _p(s), len(s))
    return out


def dolph_chebyshev(lobefrac, tol):
    w = lib().or_dolph_width(lobefrac, tol)
    x = np.empty(w, dtype=np.complex128)
    lib().or_make_dolphchebyshev_t(lobefrac, tol, w, _p(x))
    return x


def make_multiple(x, n, b):
    x = np.ascontiguousarray(x, dtype=np.complex128).copy()
    h = np.empty(n, dtype=np.complex128)
    lib().or_make_multiple_t(_p(x), len(x), n, b, _p(h))
    return x, h


def v12_params(n, k, with_comb):
    p = V12Params()
    lib().or_v12_params_get(n, k, int(with_comb), C.byref(p))
    return p


def v3_params(n, k):
    p = V3Params()
    lib().or_v3_params_get(n, k, C.byref(p))
    return p


def params_dict(p):
    return {f: getattr(p, f) for f, _ in p._fields_}


class V12:
    """Oracle plan for sFFT v1 (with_comb=False) or v2 (with_comb=True)."""

    def __init__(self, n, k, with_comb):
        self._h = lib().or_v12_plan_make(n, k, int(with_comb))
        self.p = V12Params.from_buffer_copy(self._h.contents.p)  # outlives the plan
        self.n = self.p.n
        self.loops = self.p.loc_loops + self.p.est_loops

    def __del__(self):
        if getattr(self, "_h", None) is not None:
            lib().or_v12_plan_free(self._h)
            self._h = None

    def filters(self):
        p, h = self.p, self._h.contents
        def arr(ptr, cnt):
            return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(C.c_double)),
                                         shape=(2 * cnt,)).view(np.complex128).copy()
        return {"loc_time": arr(h.loc_time, p.w_loc),
                "loc_freq": arr(h.loc_freq, self.n),
                "est_time": arr(h.est_time, p.w_est),
                "est_freq": arr(h.est_freq, self.n)}

    def set_filters(self, loc_time, loc_freq, est_time, est_freq):
        arrs = [np.ascontiguousarray(a, dtype=np.complex128)
                for a in (loc_time, loc_freq, est_time, est_freq)]
        assert len(arrs[0]) == self.p.w_loc and len(arrs[2]) == self.p.w_est
        assert len(arrs[1]) == self.n and len(arrs[3]) == self.n
        lib().or_v12_plan_set_filters(self._h, *[_p(a) for a in arrs])

    def draw(self):
        """RNG draws of one exec from the process-global glibc streams."""
        a = np.zeros(self.loops, dtype=np.int32)
        ai = np.zeros(self.loops, dtype=np.int32)
        off = np.zeros(max(1, self.p.comb_loops), dtype=np.int32)
        lib().or_v12_draw(self.n, self.loops, self.p.comb_loops if self.p.with_comb else 0,
                          self.p.W_comb, _p(a), _p(ai), _p(off))
        return a, ai, off

    def exec(self, x, a, ai, comb_off, trace=False):
        x = np.ascontiguousarray(x, dtype=np.complex128)
        assert len(x) == self.n
        out = np.empty(self.n, dtype=np.complex128)
        a = np.ascontiguousarray(a, dtype=np.int32)
        ai = np.ascontiguousarray(ai, dtype=np.int32)
        comb_off = np.ascontiguousarray(comb_off, dtype=np.int32)
        if not trace:
            lib().or_v12_exec(self._h, _p(x), _p(a), _p(ai), _p(comb_off), _p(out), None)
            return out
        p = self.p
        xs = p.x_samp_size
        t = V12Trace()
        bufs = {"x_sampt": np.zeros(xs, np.complex128), "x_samp": np.zeros(xs, np.complex128),
                "J": np.zeros(self.loops * p.B_thresh, np.int32),
                "comb_approved": np.zeros(max(1, p.comb_loops * p.B_thresh), np.int32),
                "hits": np.zeros(2 * self.n + 16, np.int32)}
        for kname, arr in bufs.items():
            setattr(t, kname, arr.ctypes.data)
        t.hits_cap = len(bufs["hits"])
        lib().or_v12_exec(self._h, _p(x), _p(a), _p(ai), _p(comb_off), _p(out), C.byref(t))
        res = {"out": out, "x_sampt": bufs["x_sampt"], "x_samp": bufs["x_samp"],
               "J": bufs["J"].reshape(self.loops, p.B_thresh),
               "comb_approved": bufs["comb_approved"][:t.num_comb],
               "hits": bufs["hits"][:min(t.hits_found, t.hits_cap)],
               "hits_found": t.hits_found, "hits_prefill": t.hits_prefill}
        return res

    def exec_many(self, xs_ptrs, draws_a, draws_ai, draws_off, outs_ptrs, threads, fast=False):
        """Timed CPU path: OpenMP over signals with pre-drawn RNG values;
        fast=True runs the timing build (the reference's compiler flags)."""
        L = lib_fast() if fast else lib()
        L.or_v12_exec_many(self._h, len(xs_ptrs), _p(xs_ptrs), _p(draws_a),
                           _p(draws_ai), _p(draws_off), _p(outs_ptrs), threads)


class V3:
    """Oracle plan for sFFT v3 (bare=True: no filter design, for
    set_filters)."""

    def __init__(self, n, k, bare=False):
        self._h = (lib().or_v3_plan_make_bare if bare else lib().or_v3_plan_make)(n, k)
        self.p = V3Params.from_buffer_copy(self._h.contents.p)  # outlives the plan
        self.n = self.p.n

    def __del__(self):
        if getattr(self, "_h", None) is not None:
            lib().or_v3_plan_free(self._h)
            self._h = None

    def draw(self):
        d = np.zeros(6, dtype=np.int32)
        lib().or_v3_draw(self.n, _p(d))
        return d

    def set_filters(self, time1, freq1, time2, freq2):
        arrs = [np.ascontiguousarray(a, dtype=np.complex128)
                for a in (time1, freq1, time2, freq2)]
        assert len(arrs[0]) == self.p.w_g1 and len(arrs[2]) == self.p.w_g2
        assert len(arrs[1]) == self.n and len(arrs[3]) == self.n
        lib().or_v3_plan_set_filters(self._h, *[_p(a) for a in arrs])

    def exec(self, x, draws):
        x = np.ascontiguousarray(x, dtype=np.complex128)
        out = np.empty(self.n, dtype=np.complex128)
        d = np.ascontiguousarray(draws, dtype=np.int32)
        cap = 64 * self.p.k + 4 * self.p.W_Man + 1024
        fr = np.zeros(cap, np.int32)
        vals = np.zeros(cap, np.complex128)
        early = C.c_int(0)
        cnt = lib().or_v3_exec(self._h, _p(x), _p(d), _p(out), C.cast(C.pointer(early), C.c_void_p),
                               _p(fr), _p(vals), cap)
        m = min(cnt, cap) if not early.value else 0
        return {"out": out, "count": cnt, "early": bool(early.value),
                "freqs": fr[:m], "vals": vals[:m]}


def seed_glibc(random_seed=17, drand48_seed=4242):
    """srandom()/srand48() of the process-global glibc streams."""
    libc = C.CDLL(None)
    libc.srandom.argtypes = [C.c_uint]
    libc.srand48.argtypes = [C.c_long]
    libc.srandom(random_seed)
    libc.srand48(drand48_seed)


def fp32_buckets(x, n, B, taps, ai, seg_rows, nmax):
    """The FP32 form's buckets of one loop after its B-point FFT (fp32_model.c:
    the library's FP32 summation order, not the reference's): x the signal,
    taps the filter's w taps (rounded to FP32 here as the plan does), ai the
    loop's gather multiplier, nmax the plan's twiddle table size."""
    L = lib()
    xr = np.ascontiguousarray(x, dtype=np.complex128)
    t32 = np.ascontiguousarray(np.stack([np.asarray(taps).real, np.asarray(taps).imag], 1)
                               .astype(np.float32))
    out = np.zeros(2 * B, np.float32)
    L.fm_fold32(xr.ctypes.data, n, B, len(taps), t32.ctypes.data, int(ai) & 0xffffffff,
                seg_rows, out.ctypes.data)
    L.fm_fft32(out.ctypes.data, B, nmax)
    return out[0::2].astype(np.float64) + 1j * out[1::2].astype(np.float64)


def v12_draws(n, loops, comb_loops, W_comb):
    """One exec's RNG draws (a, ai, comb offsets) from the process-global
    glibc streams, without building an oracle plan."""
    a = np.zeros(loops, dtype=np.int32)
    ai = np.zeros(loops, dtype=np.int32)
    off = np.zeros(max(1, comb_loops), dtype=np.int32)
    lib().or_v12_draw(n, loops, comb_loops, W_comb, _p(a), _p(ai), _p(off))
    return a, ai, off
