"""The C-ABI library without a GPU: it loads, exports every symbol
include/sfft.h declares, rejects what the reference cannot run, and its
host-side plan logic (parameters, filter design) matches the oracle."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import oracle
import sfft_amd
from sfft_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _header_functions(name="sfft.h"):
    txt = open(os.path.join(ROOT, "include", name)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(sfft_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = _lib.load()
    declared = _header_functions()
    assert len(declared) >= 20
    for name in declared:
        assert hasattr(L, name), name
    assert sorted(_lib.EXPORTS) == declared


def test_library_exports_every_server_symbol():
    L = _lib.load()
    declared = _header_functions("sfft_server.h")
    for name in declared:
        assert hasattr(L, name), name
    assert sorted(_lib.SERVER_EXPORTS) == declared


def test_output_capacity_without_a_plan():
    """Feasibility and located-set capacity from the parameters alone."""
    L = _lib.load()
    # v2: comb_loops * 2k * n/W_Comb (n=2^22, k=64: W_Comb = 4096)
    assert L.sfft_output_capacity(1 << 22, 64, 1) == 128 * 1024
    # v2 with W_Comb < 2k+1: the reference's comb selection cannot run
    assert L.sfft_output_capacity(1 << 22, 1000, 1) == -1
    assert "W_Comb" in _lib.last_error()
    assert L.sfft_output_capacity(1000, 10, 0) == -1
    # v3: the map holds 8k + W_Man + 2 B_g1 + 2 B_g2 + 64, rounded up to 2^p
    assert L.sfft_output_capacity(1 << 20, 100, 2) == 2048


def test_enum_values_match_reference():
    # src/sfft.h:37-42
    assert (_lib.SFFT_VERSION_1, _lib.SFFT_VERSION_2, _lib.SFFT_VERSION_3) == (0, 1, 2)


def test_malloc_is_16_byte_aligned():
    L = _lib.load()
    for s in [1, 16, 1000, 1 << 20]:
        p = L.sfft_malloc(s)
        assert p and p % 16 == 0
        L.sfft_free(p)


def test_bad_inputs_return_null_with_reason():
    L = _lib.load()
    assert not L.sfft_make_plan(1 << 16, 50, 7, 64)
    assert "version" in _lib.last_error()
    assert not L.sfft_make_plan(1000, 50, 0, 64)
    assert "power of two" in _lib.last_error()
    # v2 at n=2^22 needs W_Comb >= 2k+1: k=600 is infeasible in the reference
    assert not L.sfft_make_plan(1 << 22, 600, 1, 64)
    assert "2k+1" in _lib.last_error()


def test_make_plan_without_gpu_fails_loudly():
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    L = _lib.load()
    assert not L.sfft_make_plan(1 << 14, 50, 0, 64)
    assert _lib.last_error() != ""


@pytest.mark.parametrize("version", [1, 2])
def test_derived_params_match_oracle(version):
    for p in range(13, 25):
        n = 1 << p
        for k in [10, 50, 64, 100, 300, 1000]:
            ok, d = sfft_amd.derive_params(n, k, version)
            o = oracle.params_dict(oracle.v12_params(n, k, version == 2))
            for key, val in d.items():
                if key == "W_comb" and version == 1:
                    continue
                assert val == o[key], (n, k, key)


def test_v3_params_match_oracle():
    for p in range(10, 25):
        n = 1 << p
        for k in [8, 10, 50, 200, 500, 1000, 2000]:
            ok, d = sfft_amd.derive_params(n, k, 3)
            o = oracle.params_dict(oracle.v3_params(n, k))
            for key in ["W_Man", "B_g1", "B_g2", "w_g1", "w_g2"]:
                assert d[key] == o[key], (n, k, key)


def test_feasibility_matches_reference_limits():
    """v2 needs B, W_Comb >= 2k+1 (src/utils.cc:134-138): max k per n."""
    for n, kmax in [(1 << 18, 255), (1 << 20, 319), (1 << 22, 511), (1 << 24, 1023)]:
        ok_at, d = sfft_amd.derive_params(n, kmax, 2)
        bad_at, _ = sfft_amd.derive_params(n, kmax + 1, 2)
        w_lim = d["W_comb"] >= 2 * kmax + 1 and d["B_loc"] >= 2 * kmax + 1
        assert w_lim
        assert not bad_at


@pytest.mark.parametrize("n,bb,tol,b", [(1 << 16, 452, 1e-8, 191), (1 << 20, 3238, 1e-6, 427),
                                        (1 << 14, 64, 1e-8, 256)])
def test_library_filter_design_matches_oracle(n, bb, tol, b):
    t, f = sfft_amd.design_filter(0.5 / bb, tol, n, b)
    x = oracle.dolph_chebyshev(0.5 / bb, tol)
    ot, of = oracle.make_multiple(x, n, b)
    assert len(t) == len(ot)
    assert np.abs(t - ot).max() <= 1e-11 * np.abs(ot).max()
    assert np.abs(f - of).max() <= 1e-11


def test_python_mirror_validation():
    with pytest.raises(ValueError):
        sfft_amd.sfft(4194304, 64, 2)         # not in the reference's whitelist
    with pytest.raises(ValueError):
        sfft_amd.sfft(16384, 50, 4)
    with pytest.raises(TypeError):
        sfft_amd.sfft(16384.0, 50, 1)
    with pytest.raises(ValueError):
        sfft_amd.sfft(16384, 50, 1, optimization=3)


def test_reference_input_generator():
    """src/simulation.cc:95-112 + src/utils.cc:250-280 from a private
    drand48 stream: frequencies equal glibc's drand48 draws after
    srand48(seed), x is the unnormalised inverse DFT, the noise follows the
    Box-Muller form with the stated deviation."""
    from sfft_amd.signals import reference_input
    n, k, seed = 1 << 12, 20, 4242
    x, X, snr = reference_input(n, k, seed)
    libc = C.CDLL("libc.so.6")
    libc.drand48.restype = C.c_double
    libc.srand48(C.c_long(seed))
    want = sorted({int(np.floor(libc.drand48() * n)) for _ in range(k)})
    assert np.nonzero(X)[0].tolist() == want
    assert np.all(X[want] == 1.0)
    assert np.abs(x - np.fft.ifft(X) * n).max() < 1e-9
    assert snr == 1e9
    # noisy: the first noise draws continue the same stream
    xn, Xn, snr_n = reference_input(n, k, seed, std_noise=0.5)
    assert np.array_equal(Xn, X)
    u, v = libc.drand48(), libc.drand48()
    g0 = 0.5 * np.sqrt(-2 * np.log(u)) * np.exp(2j * np.pi * v)
    assert abs((xn[0] - x[0]) - g0) < 1e-12
    noise = xn - x
    assert abs(np.mean(np.abs(noise) ** 2) / (2 * 0.25) - 1) < 0.1
    assert abs(snr_n / (np.sum(np.abs(x) ** 2) / np.sum(np.abs(noise) ** 2)) - 1) < 1e-9


def test_tools_usage_without_gpu(capsys):
    from sfft_amd import tools
    assert tools.main([]) == 1
    assert "timing_many" in capsys.readouterr().out
    with pytest.raises(SystemExit):
        tools.main(["timing", "-h"])


def _build_c_caller(out_dir):
    """gcc build of tests/c/sfft_caller.c (the reference's test/test.c call
    sequence) against include/sfft.h, linked with libsfft_amd.so."""
    import subprocess
    exe = os.path.join(str(out_dir), "sfft_caller")
    cmd = ["gcc", "-std=c11", "-Wall", "-Wextra", "-Werror", "-O2",
           "-I" + os.path.join(ROOT, "include"), "-o", exe,
           os.path.join(ROOT, "tests", "c", "sfft_caller.c"),
           "-L" + os.path.dirname(_lib.LIB_PATH), "-lsfft_amd",
           "-Wl,-rpath," + os.path.dirname(_lib.LIB_PATH), "-lm"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    return exe


def test_c_caller_compiles_and_links(tmp_path):
    """The header's C branch (complex_t = double complex) compiles warning-free
    and a C program links against the library's exports."""
    exe = _build_c_caller(tmp_path)
    assert os.path.getsize(exe) > 0


def test_cpp_includes_header(tmp_path):
    """The C++ branch (complex_t = std::complex<double>) compiles too."""
    import subprocess
    src = tmp_path / "t.cc"
    src.write_text('#include "sfft.h"\n#include "sfft_server.h"\n'
                   'int main() { complex_t z(1, 2); sfft_output_c o{0, 0, nullptr, nullptr};\n'
                   '  return (int)(z.real() + (double)o.count) - 1; }\n')
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-fsyntax-only",
                        "-I" + os.path.join(ROOT, "include"), str(src)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
