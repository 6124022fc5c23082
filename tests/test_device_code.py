"""Checks on the generated gfx950 device code (CPU only: hipcc cross-compiles).

A called device function (not a kernel) whose far branches -- past the
+-128 KB reach of a direct branch -- the compiler expands through
s[30:31], the register pair that holds its return address, returns into
its own body: a GPU fault at run time.  One such function (an 860 KB
refine solve) faulted on its first run.  Every called function of the
library's device code is checked here, from the assembly of each source.
"""
import os
import re
import shutil
import subprocess
import tempfile
from concurrent.futures import ThreadPoolExecutor

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)

pytestmark = pytest.mark.skipif(not (os.path.exists("/opt/rocm/bin/hipcc") or shutil.which("hipcc")),
                                reason="hipcc not available")


def _assembly(src, out):
    import importlib.util
    spec = importlib.util.spec_from_file_location("sfft_build", os.path.join(ROOT, "sfft_amd", "build.py"))
    b = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(b)
    cmd = [b._hipcc()] + b.COMMON + ["-x", "hip"] + b.DEVICE + [
        "--cuda-device-only", "-S", os.path.join(b.CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    return open(out).read()


def _functions(asm):
    """(name, is_kernel, body) of every function in an assembly file."""
    kernels = set(re.findall(r"^\s*\.amdhsa_kernel (\S+)", asm, re.M))
    for m in re.finditer(r"^(\S+):\s*; @\S+\n(.*?)^\.Lfunc_end\d+:", asm, re.S | re.M):
        yield m.group(1), m.group(1) in kernels, m.group(2)


def test_called_functions_keep_their_return_address():
    srcs = ["kernels_v12.hip", "kernels_v3.hip", "kernels_fft.hip"]
    with tempfile.TemporaryDirectory() as d:
        with ThreadPoolExecutor(max_workers=3) as ex:
            asms = list(ex.map(lambda s: _assembly(s, os.path.join(d, s + ".s")), srcs))
    seen = 0
    for src, asm in zip(srcs, asms):
        for name, is_kernel, body in _functions(asm):
            seen += 1
            if is_kernel:
                continue
            assert "s_getpc_b64 s[30:31]" not in body, (src, name)
    assert seen > 10
