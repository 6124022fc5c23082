"""The reference-harness tools (sfft_amd.tools) on the GPU."""
import pytest

from sfft_amd import tools

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("version", [1, 2, 3])
def test_verification_ok(torch_cuda, version, capsys):
    rc = tools.main(["verification", "-n", "65536", "-k", "50", "-v", str(version),
                     "--seed", "123"])
    out = capsys.readouterr().out
    assert rc == 0 and out.strip().endswith("OK"), out


def test_timing_prints_time(torch_cuda, capsys):
    assert tools.main(["timing", "-n", "65536", "-k", "50", "-r", "3", "-v", "2",
                       "--seed", "1"]) == 0
    assert capsys.readouterr().out.startswith("TIME ")
    assert tools.main(["timing_many", "-n", "65536", "-k", "50", "-i", "8", "-v", "1",
                       "-d", "-p", "fp32_refine", "--seed", "1"]) == 0
    assert capsys.readouterr().out.startswith("TIME ")
    assert tools.main(["timing_many", "-n", "65536", "-k", "50", "-i", "4", "-s",
                       "--seed", "1"]) == 0
    assert capsys.readouterr().out.startswith("TIME ")


@pytest.mark.parametrize("flags", [["-v", "1"], ["-v", "2", "-d"], ["-v", "3"],
                                   ["-v", "2", "-s"], ["-v", "2", "-p", "fp32_refine"]])
def test_timing_many_outputs_verified(torch_cuda, capsys, flags):
    """timing_many's dense sfft_exec_many outputs (host arrays, device with
    -d, one plan per input with -s), each checked with verification's
    criterion against the generator's spectrum (src/verification.cc:35-50)."""
    rc = tools.main(["timing_many", "-n", "65536", "-k", "50", "-i", "6", "--seed", "7",
                     "--check"] + flags)
    out = capsys.readouterr().out
    assert rc == 0 and "CHECK 6 of 6 outputs OK" in out, out
