import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
NETLIB = ROOT / "tests" / "golden" / "netlib"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C ABI on cuda:0)")
    config.addinivalue_line("markers", "slow: long-running")


def _build_if_missing():
    from saddlepoint_amd import _lib
    from oracle import oracle as O

    if not O.LIB_PATH.exists():
        O.build()
    if not (_lib.LIB_PATH.exists() and _lib.CODE_OBJECT.exists()):
        import subprocess

        subprocess.run(["make", "-C", str(ROOT / "saddlepoint_amd" / "csrc")], check=True)


_build_if_missing()


@pytest.fixture(scope="session")
def gpu():
    """The HIP backend on device 0. GPU tests fail (not skip) when it cannot open."""
    from saddlepoint_amd.backend import HipBackend

    b = HipBackend(0)
    yield b
    b.close()


def netlib(name):
    from saddlepoint_amd.mps import read_mps_file

    return read_mps_file(NETLIB / f"{name}.mps")


def reference_csv():
    import csv

    with open(NETLIB / "reference.csv") as f:
        return {r["name"]: (r["status"], float(r["objective"]) if r["objective"] else None)
                for r in csv.DictReader(f)}
