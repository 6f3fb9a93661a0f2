"""The reference's commands.txt recipe, verbatim in shape, through hbj_cli:
self-tests, text-to-bin, run-pcg-case (rel 1e-4, abs 1e-7, max 2000),
compare-x (rtol 1e-3, atol 1e-2 / 1e0 / 1e1)."""
import json
import subprocess

import numpy as np
import pytest

from test_caseio import write_text_dataset

pytestmark = pytest.mark.gpu


def run(args):
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    return r.returncode, r.stdout, r.stderr


@pytest.mark.parametrize("cmd,expect", [
    ("vec-test", "VecTest OK: all y[i] == 7"),
    ("dot-test", "DotTest OK: got 32"),
    ("spmv-test", "SpmvTest OK: y == [16, 21, 38]"),
    ("block-jacobi-test", "BlockJacobiTest OK: z == r (identity blocks)"),
    ("pcg-update-scalars-test", "PcgUpdateScalarsTest OK: alpha=5, minus_alpha=-5, beta=0.5"),
])
def test_self_tests(hbj, cmd, expect):
    rc, out, err = run([hbj.CLI_PATH, cmd])
    assert rc == 0, err
    assert expect in out


def test_info(hbj):
    rc, out, err = run([hbj.CLI_PATH, "--backend", "auto", "info"])
    assert rc == 0, err
    assert '"command": "info"' in out and "gfx950" in out


@pytest.mark.parametrize("k,atol,its", [(1, "1e-2", 42), (2, "1e0", 370), (3, "1e1", 1273)])
def test_commands_txt_recipe(hbj, oracle, tmp_path, k, atol, its):
    print(f"recipe case {k}", flush=True)
    d = oracle.load_dataset(k)
    write_text_dataset(tmp_path / f"ds{k}.txt", d, extra_section=False)
    case = tmp_path / "input" / f"case_{k}"
    outd = tmp_path / "output" / f"case_{k}"
    rc, out, err = run([hbj.CLI_PATH, "text-to-bin", "--input", str(tmp_path / f"ds{k}.txt"),
                        "--out-dir", str(case), "--force"])
    assert rc == 0, err
    rc, out, err = run([hbj.CLI_PATH, "--backend", "auto", "run-pcg-case", "--case-dir", str(case),
                        "--max-iters", "2000", "--rel-tol", "1e-4", "--abs-tol", "1e-7",
                        "--out-x", str(outd / "x.bin"), "--out-metrics", str(outd / "metrics.json")])
    assert rc == 0, err
    m = json.load(open(outd / "metrics.json"))
    assert m["converged"] is True and m["iterations"] == its and m["error"] is None
    assert m["n"] == int(d["n"]) and m["nnz"] == int(d["nnz"])
    rc, out, err = run([hbj.CLI_PATH, "compare-x", "--x-ref", str(case / "x_ref.bin"), "--x",
                        str(outd / "x.bin"), "--rel-tol", "1e-3", "--abs-tol", atol, "--top-k", "10"])
    assert rc == 0, out + err
    assert "pass               : true" in out
    # the CLI's x is the oracle's x bit for bit
    st, lu = oracle.build_lu_blocks_6(int(d["n"]), d["row_ptr"], d["col_idx"], d["values"], d["block_starts"])
    st, it, xo, _ = oracle.pcg(d["row_ptr"], d["col_idx"], d["values"], lu, d["block_starts"], d["b"],
                               d["x0"], 2000, 1e-4, 1e-7)
    assert np.array_equal(hbj.read_x_bin(outd / "x.bin").view(np.uint32), xo.view(np.uint32))


def test_run_pcg_case_failure_writes_metrics_only(hbj, tmp_path):
    L = hbj.lattice(5)
    d = {"n": L["n"], "nnz": L["nnz"], "row_ptr": L["row_ptr"], "col_idx": L["col_idx"],
         "values": L["values"], "b": L["rhs"], "x0": np.zeros(L["n"], np.float32),
         "block_starts": L["block_starts"], "x_ref": np.zeros(L["n"], np.float32)}
    write_text_dataset(tmp_path / "l.txt", d, extra_section=False)
    rc, _, err = run([hbj.CLI_PATH, "text-to-bin", "--input", str(tmp_path / "l.txt"), "--out-dir",
                      str(tmp_path / "c")])
    assert rc == 0, err
    rc, out, err = run([hbj.CLI_PATH, "run-pcg-case", "--case-dir", str(tmp_path / "c"), "--max-iters", "3",
                        "--out-x", str(tmp_path / "o" / "x.bin"), "--out-metrics", str(tmp_path / "o" / "m.json")])
    assert rc == 0
    assert "Solve failed: PCG(BlockJacobiGpu): did not converge in 3 iterations" in err
    m = json.load(open(tmp_path / "o" / "m.json"))
    assert m["converged"] is False and m["iterations"] is None
    assert not (tmp_path / "o" / "x.bin").exists()


def test_run_pcg_case_one_reduction(hbj, oracle, tmp_path):
    """--algorithm one-reduction: the optional PCG through the CLI; x is the
    oracle's one-reduction x bit for bit and passes compare-x."""
    d = oracle.load_dataset(2)
    write_text_dataset(tmp_path / "ds2.txt", d, extra_section=False)
    case, outd = tmp_path / "case", tmp_path / "out"
    rc, out, err = run([hbj.CLI_PATH, "text-to-bin", "--input", str(tmp_path / "ds2.txt"),
                        "--out-dir", str(case), "--force"])
    assert rc == 0, err
    rc, out, err = run([hbj.CLI_PATH, "run-pcg-case", "--case-dir", str(case), "--max-iters", "2000",
                        "--rel-tol", "1e-4", "--abs-tol", "1e-7", "--algorithm", "one-reduction",
                        "--out-x", str(outd / "x.bin"), "--out-metrics", str(outd / "metrics.json")])
    assert rc == 0, err
    m = json.load(open(outd / "metrics.json"))
    st, lu = oracle.build_lu_blocks_6(int(d["n"]), d["row_ptr"], d["col_idx"], d["values"],
                                      d["block_starts"])
    sto, ito, xo = oracle.ccg(d["row_ptr"], d["col_idx"], d["values"], lu, d["block_starts"], d["b"],
                              d["x0"], 2000, 1e-4, 1e-7)
    assert m["converged"] is True and m["algorithm"] == "one-reduction" and m["iterations"] == ito
    x = hbj.read_x_bin(str(outd / "x.bin"))
    assert np.array_equal(x.view(np.uint32), xo.view(np.uint32))
    rc, out, err = run([hbj.CLI_PATH, "compare-x", "--x-ref", str(case / "x_ref.bin"), "--x",
                        str(outd / "x.bin"), "--rel-tol", "1e-3", "--abs-tol", "1e0"])
    assert rc == 0 and "pass               : true" in out
    rc, out, err = run([hbj.CLI_PATH, "run-pcg-case", "--case-dir", str(case), "--algorithm", "cg",
                        "--out-x", str(outd / "y.bin"), "--out-metrics", str(outd / "m2.json")])
    assert rc != 0 and "--algorithm must be pcg or one-reduction" in err
