"""Build libumans_gpu.so in-tree for gfx950 with hipcc (no GPU needed to compile).

    python -m umans_amd.build          # or: python umans_amd/build.py

The translation units compile in parallel into umans_amd/_build/ and are linked into
one shared library: the host code and the grid, population and adjoint kernels
(api.hip), the step kernel's instances of each policy class (step_class.hip, once per
class and half of the list capacities), the cooperative-sampling kernel
(step_coop.hip) and the host-only scenario loader, registry and embedding API.

Flags: -ffp-contract=off keeps every float expression rounded as the reference's
scalar x86 code rounds it; packed FP32 ops are disabled (their results are not
trusted on these machines).
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libumans_gpu.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("UMANS_OFFLOAD_ARCH", "gfx950")

# (object name, source, extra defines)
UNITS = ([("api", "api.hip", ())] +
         [(f"step_class{c}_{h}", "step_class.hip", (f"UMANS_CLS={c}", f"UMANS_PART={h}"))
          for c in range(4) for h in range(2)] +
         [("step_coop", "step_coop.hip", ()), ("registry", "registry.cpp", ()),
          ("scenario", "scenario.cpp", ()), ("extapi", "extapi.cpp", ())])
FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-command-line-argument", "-I" + os.path.join(ROOT, "include"), "-I" + CSRC]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
    return r.returncode


def build(verbose=False, out=LIB, defines=(), jobs=None):
    os.makedirs(OBJ, exist_ok=True)
    tag = "_".join(d.replace("=", "") for d in defines)
    cmds, objs = [], []
    for name, src, defs in UNITS:
        obj = os.path.join(OBJ, f"{name}{'_' + tag if tag else ''}.o")
        objs.append(obj)
        cmds.append([HIPCC] + FLAGS + ["-D" + d for d in tuple(defines) + tuple(defs)] +
                    ["-c", os.path.join(CSRC, src), "-o", obj])
    jobs = jobs or min(len(cmds), os.cpu_count() or 1, 16)
    with ThreadPoolExecutor(jobs) as ex:
        rcs = list(ex.map(lambda c: _run(c, verbose), cmds))
    if any(rcs):
        raise RuntimeError("hipcc failed building libumans_gpu.so")
    if _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}"] + objs + ["-o", out + ".tmp"], verbose):
        raise RuntimeError("linking libumans_gpu.so failed")
    os.replace(out + ".tmp", out)
    return out


if __name__ == "__main__":
    print(build(verbose=True))
