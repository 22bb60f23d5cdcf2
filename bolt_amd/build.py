"""Build the gfx950 HIP library ``bolt_amd/libboltx_hip.so`` in-tree with hipcc.

``python -m bolt_amd.build`` compiles every HIP source of the package.  The
library is loaded by ``bolt_amd._lib``; nothing is JIT-compiled at import.
"""
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
SRC = os.path.join(PKG, "csrc", "boltx.hip")
OUT = os.path.join(PKG, "libboltx_hip.so")

# -packed-fp32-ops: no v_pk_*_f32 in generated code (scalar FP32 only).
HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-ffp-contract=fast-honor-pragmas",
    "-I", os.path.join(ROOT, "include"),
]


def hipcc():
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the HIP library cannot be built")


def build(verbose=False):
    """Compile ``libboltx_hip.so`` (always recompiles)."""
    tmp = OUT + ".tmp"
    cmd = [hipcc()] + HIPCC_FLAGS + ["-o", tmp, SRC]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    noise = "is not a recognized feature for this target"
    err = "\n".join(l for l in res.stderr.splitlines() if noise not in l and l.strip())
    if res.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + err)
    if err and verbose:
        print(err, file=sys.stderr)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
