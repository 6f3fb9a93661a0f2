"""Build recipe of the HIP library (libhbj.so, gfx950) -- in-tree, no JIT cache.

    python -m wgpu_solver_backend_amd._build
"""
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libhbj.so")
CLI = os.path.join(PKG, "hbj_cli")

SOURCES = ["kernels.hip", "runtime.hip", "server.hip", "noalloc.hip", "setup.cpp", "lattice.cpp", "caseio.cpp"]

# -ffp-contract=off: the FP32 path restates the reference's operation order
# (no fused multiply-adds), IEEE division is the hipcc default.
# -packed-fp32-ops: no v_pk_* FP32 instructions are generated.
HIPFLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if c and (os.path.isabs(c) and os.path.exists(c) or not os.path.isabs(c)):
            return c
    return "hipcc"


def build(verbose=False):
    srcs = [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    cmd = [hipcc()] + HIPFLAGS + ["-shared", "-I", os.path.join(ROOT, "include"), "-I", CSRC] + srcs + [
        "-o", LIB + ".tmp", "-lpthread"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(ln for ln in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in ln)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + out)
    if verbose and out.strip():
        print(out)
    os.replace(LIB + ".tmp", LIB)
    cli_src = os.path.join(CSRC, "cli.cpp")
    if os.path.exists(cli_src):
        cmd = ["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"), cli_src, "-o", CLI + ".tmp",
               "-L", PKG, "-lhbj", "-Wl,-rpath,$ORIGIN", "-lpthread"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("g++ (cli) failed:\n" + r.stdout + r.stderr)
        os.replace(CLI + ".tmp", CLI)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print("built", LIB)
