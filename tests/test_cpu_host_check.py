"""The host paths of the C ABI (MPS reader, host scaling, validation errors,
SpMV plan rule) built with AddressSanitizer and UBSan and run over every
Netlib file and some malformed MPS text (saddlepoint_amd/csrc/host_check.cpp);
no GPU is used."""
import shutil
import subprocess
from pathlib import Path

import pytest

CSRC = Path(__file__).resolve().parents[1] / "saddlepoint_amd" / "csrc"


def test_host_paths_under_sanitizers():
    if shutil.which("g++") is None or shutil.which("make") is None:
        pytest.skip("needs g++ and make")
    r = subprocess.run(["make", "-C", str(CSRC), "host-check"], capture_output=True, text=True, timeout=1200)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "host check ok" in r.stdout, out[-4000:]
    assert "ERROR: AddressSanitizer" not in out and "runtime error" not in out, out[-4000:]
