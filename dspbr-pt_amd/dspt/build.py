"""Builds the gfx950 engine library (libdspt.so) in-tree with hipcc.

The library is the product: every HIP kernel plus the C ABI declared in
include/dspt.h. It is built for gfx950 only.
"""
import os
import shutil
import subprocess
import sys

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG_DIR)  # .../dspbr-pt_amd
CSRC = os.path.join(ROOT, "csrc")
LIB_PATH = os.path.join(PKG_DIR, "libdspt.so")

HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # strict IEEE f32: no contraction, so traversal matches the oracle bit for bit
    "-ffp-contract=off",
    # keep the compiler from forming packed f32 VALU ops
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]
SOURCES = ["dspt_host.hip", "bvh_build.cpp"]


def hipcc_path():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (ROCm is required to build the engine)")


def build(verbose=False, out=None, defines=()):
    """Compiles the library to `out` (default: the in-tree libdspt.so);
    `defines` are extra -D tuning macros (for variant builds)."""
    out = os.path.abspath(out or LIB_PATH)
    cmd = ([hipcc_path()] + HIP_FLAGS + [f"-D{d}" for d in defines] + ["-o", out + ".tmp"] +
           [os.path.join(CSRC, s) for s in SOURCES])
    if verbose:
        print(" ".join(cmd), flush=True)
    proc = subprocess.run(cmd, cwd=CSRC, capture_output=True, text=True)
    noise = "packed-fp32-ops"
    err = "\n".join(l for l in proc.stderr.splitlines() if noise not in l and l.strip())
    if proc.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + err)
    if err and verbose:
        print(err, file=sys.stderr)
    os.replace(out + ".tmp", out)
    return out


if __name__ == "__main__":
    import argparse
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=None)
    ap.add_argument("-D", dest="defines", action="append", default=[])
    a = ap.parse_args()
    print(build(verbose=True, out=a.out, defines=a.defines))
