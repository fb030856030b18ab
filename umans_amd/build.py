"""Build libumans_gpu.so in-tree for gfx950 with hipcc (no GPU needed to compile).

    python -m umans_amd.build          # or: python umans_amd/build.py

Flags: -ffp-contract=off keeps every float expression rounded as the reference's
scalar x86 code rounds it; packed FP32 ops are disabled (their results are not
trusted on these machines).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libumans_gpu.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("UMANS_OFFLOAD_ARCH", "gfx950")

SOURCES = ["api.hip", "registry.cpp", "scenario.cpp", "extapi.cpp"]
FLAGS = ["-O3", "-std=c++17", "-fPIC", "-shared", f"--offload-arch={ARCH}", "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-command-line-argument", "-I" + os.path.join(ROOT, "include"), "-I" + CSRC]


def build(verbose=False, out=LIB, defines=()):
    cmd = ([HIPCC] + FLAGS + ["-D" + d for d in defines] + [os.path.join(CSRC, s) for s in SOURCES] +
           ["-o", out + ".tmp"])
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libumans_gpu.so")
    os.replace(out + ".tmp", out)
    return out


if __name__ == "__main__":
    print(build(verbose=True))
