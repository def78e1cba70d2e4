"""The self-consistency command line (python -m mbsymmetry_amd.scloop) on
config 1: FP64 against the mixed mode from the same G0, and the FP32 run with
its FP64 cleanup.  Checks the summary it prints, not only that it runs."""
import json

import pytest

from mbsymmetry_amd import scloop

pytestmark = pytest.mark.gpu


def _run(capsys, *args):
    print("[case] scloop", *args, flush=True)
    assert scloop.main(["--system", "config1", *args]) == 0
    out = capsys.readouterr().out.strip().splitlines()
    lines = [l for l in out if l.startswith("[scloop]")]
    return json.loads(out[-1]), lines


def test_cli_fp64_against_mixed(capsys, tmp_path):
    path = tmp_path / "sc.npz"
    summary, lines = _run(capsys, "--precision", "fp64", "--compare", "mixed", "--iters", "4", "--out", str(path))
    runs = summary["runs"]
    assert set(runs) == {"fp64", "mixed"}
    for r in runs.values():
        assert r["iterations"] == len(r["max_abs_dG"]) <= 4
        assert r["max_abs_dG"][-1] < r["max_abs_dG"][0]      # the loop contracts
        assert r["max_abs_G"] <= 1.0 + 1e-12                 # causal bound, S = I
    # one progress line per iteration of each run
    assert len(lines) == runs["fp64"]["iterations"] + runs["mixed"]["iterations"]
    assert summary["max_abs_G_difference"] <= 1e-8
    saved = scloop.load_results(str(path))
    assert int(saved["iterations"]) == runs["fp64"]["iterations"]


def test_cli_fp32_cleanup_reaches_fp64(capsys):
    summary, _ = _run(capsys, "--precision", "fp32", "--cleanup", "--compare", "fp64", "--iters", "30",
                      "--tol", "1e-9")
    clean, ref = summary["runs"]["fp32"], summary["runs"]["fp64"]
    assert clean["converged"] and ref["converged"]
    assert clean["precisions"][0] == "fp32" and clean["precisions"][-1] == "fp64"
    assert summary["max_abs_G_difference"] <= 1e-8
