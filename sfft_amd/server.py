"""Python interface of the serving loop (include/sfft_server.h).

Requests are device-resident torch tensors; results land in device
tensors the wrapper allocates (or the caller passes).  The C++ server does
the batching, the HBM admission and the timing; this module only keeps the
tensors alive while their request is in flight.
"""
import ctypes as C

from . import _lib
from .sfft import PRECISIONS, SfftError


def output_capacity(n, k, version):
    """Located-set capacity a request of (n, k, version 1..3) needs; raises
    when the plan cannot be built for it (e.g. W_Comb < 2k+1 for v2)."""
    c = _lib.load().sfft_output_capacity(n, k, version - 1)
    if c < 0:
        raise SfftError(_lib.last_error())
    return int(c)


class Server:
    def __init__(self, hbm_budget=0, workers=4, max_batch=64):
        self._L = _lib.load()
        self._h = self._L.sfft_server_create(int(hbm_budget), int(workers), int(max_batch))
        if not self._h:
            raise SfftError(_lib.last_error())
        self._live = {}

    def prepare(self, n, k, version, precision="fp32_refine"):
        """Builds the plan of (n, k, version, precision) ahead of its requests."""
        if self._L.sfft_server_prepare(self._h, n, k, version - 1, PRECISIONS[precision]) != 0:
            raise SfftError(_lib.last_error())

    def submit(self, x, k, version, precision="fp32_refine", seed=(1, 1), outputs=None):
        """Queues the transform of device tensor x (complex128, n values).

        seed = (srandom seed, srand48 seed) of this request's private RNG
        streams.  Returns a ticket; raises when the request is rejected.
        """
        import torch
        n = int(x.numel())
        assert x.dtype == torch.complex128 and x.is_cuda and x.is_contiguous()
        if outputs is None:
            try:
                cap = output_capacity(n, k, version)
            except SfftError:
                cap = 1  # let the server reject it (and count it)
            fr = torch.empty(cap, dtype=torch.int32, device=x.device)
            va = torch.empty(cap, dtype=torch.complex128, device=x.device)
        else:
            fr, va = outputs
            cap = int(fr.numel())
        r = _lib.SfftRequest(n, int(k), version - 1, PRECISIONS[precision], x.data_ptr(),
                             fr.data_ptr(), va.data_ptr(), cap, int(seed[0]), int(seed[1]))
        t = self._L.sfft_server_submit(self._h, C.byref(r))
        if t < 0:
            raise SfftError(_lib.last_error())
        self._live[t] = (x, fr, va)
        return int(t)

    def poll(self, ticket):
        return int(self._L.sfft_server_poll(self._h, ticket))

    def wait(self, ticket):
        """(freq, val, latency_ms) of a finished request (device tensors)."""
        cnt = C.c_int64(0)
        lat = C.c_double(0)
        rc = self._L.sfft_server_wait(self._h, ticket, C.byref(cnt), C.byref(lat))
        x, fr, va = self._live.pop(ticket)
        if rc != 0:
            raise SfftError(_lib.last_error())
        c = min(int(cnt.value), int(fr.numel()))
        return fr[:c], va[:c], float(lat.value)

    def stats(self):
        s = _lib.SfftServerStats()
        self._L.sfft_server_get_stats(self._h, C.byref(s))
        return {name: getattr(s, name) for name, _ in s._fields_}

    def close(self):
        if self._h:
            self._L.sfft_server_destroy(self._h)
            self._h = None
            self._live.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
