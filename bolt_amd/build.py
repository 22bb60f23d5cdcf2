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
CSRC = os.path.join(PKG, "csrc")
# translation units, compiled in parallel and linked into one library
SOURCES = [os.path.join(CSRC, n) for n in ("boltx.hip", "ql_f64.hip", "ql_f32.hip", "ens.hip")]
OUT = os.path.join(PKG, "libboltx_hip.so")

# -packed-fp32-ops: no v_pk_*_f32 in generated code (scalar FP32 only).
HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-ffp-contract=fast-honor-pragmas",
    "-I", os.path.join(ROOT, "include"),
]


def hipcc():
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the HIP library cannot be built")


def _filter(stderr):
    noise = "is not a recognized feature for this target"
    return "\n".join(l for l in stderr.splitlines() if noise not in l and l.strip())


def build(verbose=False):
    """Compile ``libboltx_hip.so`` (always recompiles every unit)."""
    objs, procs = [], []
    for src in SOURCES:
        obj = OUT + "." + os.path.splitext(os.path.basename(src))[0] + ".o"
        cmd = [hipcc()] + HIPCC_FLAGS + ["-c", "-o", obj, src]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stderr=subprocess.PIPE, text=True)))
        objs.append(obj)
    failed = []
    for src, p in procs:
        err = _filter(p.communicate()[1])
        if p.returncode != 0:
            failed.append(f"{src}:\n{err}")
        elif err and verbose:
            print(err, file=sys.stderr)
    if failed:
        raise RuntimeError("hipcc failed:\n" + "\n".join(failed))
    tmp = OUT + ".tmp"
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("link failed:\n" + _filter(res.stderr))
    for o in objs:
        os.remove(o)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
