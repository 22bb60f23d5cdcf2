"""Build the CPU harness tests/_build/libstage_host.so (host code only)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
OUT_DIR = os.path.join(os.path.dirname(HERE), "_build")
OUT = os.path.join(OUT_DIR, "libstage_host.so")


def build(contract="fast-honor-pragmas"):
    os.makedirs(OUT_DIR, exist_ok=True)
    out = OUT if contract == "fast-honor-pragmas" else OUT.replace(".so", "_%s.so" % contract)
    src = os.path.join(HERE, "stage_host.hip")
    cmd = ["/opt/rocm/bin/hipcc", "--offload-host-only", "-x", "hip", "-O2", "-std=c++17",
           "-fPIC", "-shared", "-mfma", "-DBX_NO_KERNELS", "-ffp-contract=%s" % contract, "-o", out, src]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return out


if __name__ == "__main__":
    print(build())
