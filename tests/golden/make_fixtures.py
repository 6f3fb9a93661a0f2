"""Write the golden fixtures under tests/golden/ from the reference's data files.

Run in a container where the reference checkout is mounted:

    python tests/golden/make_fixtures.py --reference /root/reference

Outputs (data only, no reference source):

* ``ds{1,2,3}.npz`` -- the three FEA systems shipped as text datasets
  (``datasets/pcg_block_jacobi_dataset_{1,2,3}.txt``): CSR arrays, ``b``, ``x0``,
  ``block_starts`` and ``x_ref``.  Parsed with the section grammar of the
  reference's ``text-to-bin`` tool (``crates/wgpu_solver_backend_convert_cli/
  src/main.rs:104-401``: ``n``/``nnz`` header lines, then ``<name>`` +
  ``len <k>`` + k whitespace-separated tokens, ``#`` comments, unknown sections
  skipped).
* ``kats.json`` -- the five known-answer self-tests of the reference CLI
  (``crates/wgpu_solver_backend_cli/src/main.rs:171-443``), transcribed as
  input/expected vectors.

The datasets are MIT/Apache-2.0 licensed by their author (reference
``README.md:99-101``).
"""
import argparse
import json
import os

import numpy as np

SECTIONS = {
    "row_ptr_u32": np.uint32,
    "col_idx_u32": np.uint32,
    "values_f32": np.float32,
    "b_f32": np.float32,
    "x0_f32": np.float32,
    "block_starts_u32": np.uint32,
    "x_ref_f32": np.float32,
}


def parse_dataset(path):
    """Parse one text dataset into a dict of numpy arrays."""
    with open(path, "r") as f:
        lines = [ln.strip() for ln in f]
    lines = [ln for ln in lines if ln and not ln.startswith("#")]
    out = {}
    i = 0
    while i < len(lines):
        ln = lines[i]
        head = ln.split()
        if head[0] in ("n", "nnz") and len(head) == 2:
            out[head[0]] = int(head[1])
            i += 1
            continue
        name = head[0]
        assert lines[i + 1].startswith("len "), (path, ln)
        k = int(lines[i + 1].split()[1])
        i += 2
        toks = []
        while len(toks) < k:
            toks.extend(lines[i].split())
            i += 1
        assert len(toks) == k, (name, len(toks), k)
        if name in SECTIONS:
            dt = SECTIONS[name]
            if dt == np.float32:
                out[name] = np.array([float(t) for t in toks], dtype=np.float64).astype(np.float32)
            else:
                out[name] = np.array([int(t) for t in toks], dtype=np.uint32)
    return out


def kats():
    """The reference CLI self-tests (cli main.rs:171-443) as data."""
    return {
        "vec_test": {"cite": "cli main.rs:171-211", "n": 1024, "alpha": 3.0, "x": 2.0, "y": 1.0,
                     "expected": 7.0, "tol": 1e-6},
        "dot_test": {"cite": "cli main.rs:213-248", "a": [1.0, 2.0, 3.0], "b": [4.0, 5.0, 6.0],
                     "expected": 32.0, "tol": 1e-5},
        "spmv_test": {"cite": "cli main.rs:250-315", "row_ptr": [0, 2, 4, 6], "col_idx": [0, 2, 0, 1, 1, 2],
                      "values": [10.0, 2.0, 3.0, 9.0, 7.0, 8.0], "x": [1.0, 2.0, 3.0],
                      "expected": [16.0, 21.0, 38.0], "tol": 1e-6},
        "block_jacobi_test": {"cite": "cli main.rs:317-368", "n": 12, "block_starts": [0, 6, 12],
                              "lu": "identity", "r": list(range(1, 13)), "expected": list(range(1, 13)),
                              "tol": 1e-6},
        "pcg_update_scalars_test": {"cite": "cli main.rs:370-443", "slots": 8,
                                    "PAP": 0, "RZ_NEW": 1, "RZ_OLD": 2, "ALPHA": 3, "MINUS_ALPHA": 4,
                                    "BETA": 5, "pAp": 2.0, "rz_old": 10.0, "rz_new": 5.0,
                                    "alpha": 5.0, "minus_alpha": -5.0, "beta": 0.5, "tol": 1e-6},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--reference", default="/root/reference")
    ap.add_argument("--out", default=os.path.dirname(os.path.abspath(__file__)))
    args = ap.parse_args()
    for k in (1, 2, 3):
        src = os.path.join(args.reference, "datasets", f"pcg_block_jacobi_dataset_{k}.txt")
        d = parse_dataset(src)
        n, nnz = d["n"], d["nnz"]
        assert d["row_ptr_u32"].size == n + 1 and d["col_idx_u32"].size == nnz
        np.savez_compressed(
            os.path.join(args.out, f"ds{k}.npz"),
            n=np.uint32(n), nnz=np.uint32(nnz),
            row_ptr=d["row_ptr_u32"], col_idx=d["col_idx_u32"], values=d["values_f32"],
            b=d["b_f32"], x0=d["x0_f32"], block_starts=d["block_starts_u32"], x_ref=d["x_ref_f32"],
        )
        print(f"ds{k}: n={n} nnz={nnz} blocks={d['block_starts_u32'].size - 1}")
    with open(os.path.join(args.out, "kats.json"), "w") as f:
        json.dump(kats(), f, indent=1)


if __name__ == "__main__":
    main()
