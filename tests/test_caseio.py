"""Case-directory formats, text-to-bin and compare-x (host code, no GPU).

The text datasets are regenerated from the committed .npz fixtures in the
reference's text grammar (n / nnz headers, <name> + len k + tokens, comments),
then parsed back; the binary layout is checked byte by byte against the
format the reference documents (convert main.rs:64-81)."""
import os
import struct
import subprocess

import numpy as np
import pytest

F = np.float32


def write_text_dataset(path, d, extra_section=True, wrap=16):
    with open(path, "w") as f:
        f.write("# wgpu-solver-backend dataset (text)\n\n")
        f.write(f"n {int(d['n'])}\nnnz {int(d['nnz'])}\n\n")
        secs = [("row_ptr_u32", d["row_ptr"], "%d"), ("col_idx_u32", d["col_idx"], "%d"),
                ("values_f32", d["values"], "%.9e"), ("b_f32", d["b"], "%.9e"),
                ("x0_f32", d["x0"], "%.9e")]
        if extra_section:
            secs.append(("comment_vector_f32", np.arange(5, dtype=F), "%.3e"))
        secs += [("block_starts_u32", d["block_starts"], "%d"), ("x_ref_f32", d["x_ref"], "%.9e")]
        for name, arr, fmt in secs:
            f.write(f"{name}\nlen {arr.size}\n")
            for i in range(0, arr.size, wrap):
                f.write(" ".join(fmt % v for v in arr[i:i + wrap]) + " \n")
            f.write("\n")


@pytest.mark.parametrize("k", [1, 2, 3])
def test_text_to_bin_roundtrip(hbj, oracle, tmp_path, k):
    d = oracle.load_dataset(k)
    txt = tmp_path / f"ds{k}.txt"
    write_text_dataset(txt, d)
    out = tmp_path / f"case_{k}"
    hbj.text_to_bin(txt, out)
    c = hbj.load_case_dir(out)
    for key in ("row_ptr", "col_idx", "values", "b", "x0", "block_starts"):
        assert np.array_equal(c[key].view(np.uint32), np.ascontiguousarray(d[key]).view(np.uint32)), key
    assert np.array_equal(hbj.read_x_bin(out / "x_ref.bin"), d["x_ref"])
    raw = open(out / "matrix.csr.bin", "rb").read()
    n, nnz = int(d["n"]), int(d["nnz"])
    assert struct.unpack("<8I", raw[:32]) == (0x43535231, 1, n, n, nnz, 0, 0, 0)
    assert len(raw) == 32 + 4 * (n + 1) + 8 * nnz
    with pytest.raises(hbj.SolverError, match="already exists. Use --force to overwrite."):
        hbj.text_to_bin(txt, out)
    hbj.text_to_bin(txt, out, force=True)


def test_loader_error_texts(hbj, tmp_path):
    out = tmp_path / "c"
    d = {"n": 2, "nnz": 2, "row_ptr": np.array([0, 1, 2], np.uint32), "col_idx": np.array([0, 1], np.uint32),
         "values": np.ones(2, F), "b": np.ones(2, F), "x0": np.zeros(2, F),
         "block_starts": np.array([0, 2], np.uint32), "x_ref": np.ones(2, F)}
    write_text_dataset(tmp_path / "t.txt", d, extra_section=False)
    hbj.text_to_bin(tmp_path / "t.txt", out)
    m = bytearray(open(out / "matrix.csr.bin", "rb").read())
    good = bytes(m)
    m[0] = 0
    open(out / "matrix.csr.bin", "wb").write(m)
    with pytest.raises(hbj.SolverError, match=r"Bad matrix magic in .*: expected 0x43535231, got 0x43535200"):
        hbj.load_case_dir(out)
    open(out / "matrix.csr.bin", "wb").write(good + b"\0")
    with pytest.raises(hbj.SolverError, match="matrix.csr.bin has 1 trailing bytes after expected payload"):
        hbj.load_case_dir(out)
    open(out / "matrix.csr.bin", "wb").write(good[:-3])
    with pytest.raises(hbj.SolverError, match=r"Unexpected EOF while reading values\[1\]"):
        hbj.load_case_dir(out)
    open(out / "matrix.csr.bin", "wb").write(good)
    open(out / "block_starts.bin", "wb").write(struct.pack("<4I", 3, 0, 0, 2))
    with pytest.raises(hbj.SolverError, match="block_starts must be strictly increasing"):
        hbj.load_case_dir(out)
    open(out / "block_starts.bin", "wb").write(struct.pack("<3I", 2, 0, 2))
    open(out / "rhs.bin", "wb").write(struct.pack("<I3f", 3, 1, 1, 1))
    with pytest.raises(hbj.SolverError, match="rhs length mismatch: b.n=3 but matrix n_rows=2"):
        hbj.load_case_dir(out)


def test_text_parser_errors(hbj, tmp_path):
    p = tmp_path / "bad.txt"
    p.write_text("n 2\nnnz 1\nrow_ptr_u32\nlen 3\n0 1\ncol_idx_u32\nlen 1\n0\n")
    with pytest.raises(hbj.SolverError, match="Section row_ptr_u32: hit unexpected header 'col_idx_u32'"):
        hbj.text_to_bin(p, tmp_path / "o")
    p.write_text("n 2\nnnz 1\nrow_ptr_u32\nlen 3\n0 1 1\n")
    with pytest.raises(hbj.SolverError, match="Missing col_idx_u32"):
        hbj.text_to_bin(p, tmp_path / "o2")


def test_compare_x_metric(hbj):
    rng = np.random.default_rng(0)
    xr = rng.standard_normal(1000).astype(F)
    x = (xr + rng.standard_normal(1000).astype(F) * F(1e-3)).astype(F)
    s = hbj.compare_x(xr, x, top_k=5)
    err = np.abs(x - xr).astype(F).astype(np.float64)  # |b - a| in f32, summed in f64
    assert np.isclose(s["l2_err"], np.sqrt(np.sum(err ** 2)), rtol=1e-14)
    assert np.isclose(s["l2_ref"], np.linalg.norm(xr.astype(np.float64)), rtol=1e-14)
    assert np.isclose(s["rmse"], np.sqrt(np.mean(err ** 2)), rtol=1e-14)
    assert s["max_abs_err"] == err.max()
    assert s["worst"] == list(np.argsort(-err, kind="stable")[:5])
    z = hbj.compare_x(np.zeros(3, F), np.zeros(3, F))
    assert z["rel_l2"] == 0.0 and z["rmse"] == 0.0


def test_cli_text_to_bin_and_compare(hbj, oracle, tmp_path):
    d = oracle.load_dataset(1)
    write_text_dataset(tmp_path / "ds1.txt", d)
    cli = hbj.CLI_PATH
    r = subprocess.run([cli, "text-to-bin", "--input", str(tmp_path / "ds1.txt"), "--out-dir",
                        str(tmp_path / "case_1"), "--force"], capture_output=True, text=True)
    assert r.returncode == 0 and "Wrote binaries into" in r.stdout
    hbj.write_x_bin(tmp_path / "x.bin", d["x_ref"] * F(1.0005))
    r = subprocess.run([cli, "compare-x", "--x-ref", str(tmp_path / "case_1" / "x_ref.bin"), "--x",
                        str(tmp_path / "x.bin"), "--rel-tol", "1e-3", "--abs-tol", "1e-2"],
                       capture_output=True, text=True)
    assert r.returncode == 0 and "pass               : true" in r.stdout
    r = subprocess.run([cli, "compare-x", "--x-ref", str(tmp_path / "case_1" / "x_ref.bin"), "--x",
                        str(tmp_path / "x.bin"), "--rel-tol", "1e-6", "--abs-tol", "1e-9"],
                       capture_output=True, text=True)
    assert r.returncode == 2 and "CompareX failed tolerances (rel_l2 and rmse)" in r.stderr
