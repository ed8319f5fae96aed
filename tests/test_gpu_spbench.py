"""``python -m saddlepoint_amd.spbench`` writes the reference's sp-bench CSV
schema (crates/saddlepoint-cli/src/bin/sp-bench.rs:22-23, 150-172): one row per
(instance, solver), residuals from the KKT checker, appended to --out with one
header."""
import csv

import pytest

from conftest import NETLIB
from saddlepoint_amd import spbench

pytestmark = pytest.mark.gpu


def test_spbench_rows(tmp_path, capsys):
    out = tmp_path / "bench.csv"
    files = [str(NETLIB / f"{n}.mps") for n in ("afiro", "israel", "woodinfe")]
    assert spbench.main(["--solvers", "gpu,gpu-mixed", "--tol", "1e-4", "--out", str(out)] + files) == 0
    assert spbench.main(["--solvers", "gpu", "--termination", "cupdlpx", "--out", str(out), files[0]]) == 0
    rows = list(csv.DictReader(open(out)))
    assert list(rows[0].keys()) == spbench.HEADER.split(",")
    assert len(rows) == 7  # one header despite two runs
    by = {(r["instance"].lower(), r["solver"]): r for r in rows}  # the instance is the MPS NAME
    assert by[("afiro", "gpu")]["status"] == "Optimal" and by[("afiro", "gpu+step")]["status"] == "Optimal"
    assert by[("woodinfe", "gpu")]["status"] == "PrimalInfeasible"
    for key in (("afiro", "gpu"), ("israel", "gpu"), ("israel", "gpu-mixed")):
        r = by[key]
        assert int(r["rows"]) > 0 and int(r["nnz"]) > 0 and int(r["iterations"]) % 64 == 0
        assert max(float(r["rel_primal"]), float(r["rel_dual"]), float(r["rel_gap"])) <= 2e-4
    assert capsys.readouterr().out.startswith(spbench.HEADER)
