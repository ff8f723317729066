"""Build the in-tree shared libraries.

  lib/libibhost.so   host mesh builder (g++, OpenMP)
  lib/libibgpu.so    HIP engine for gfx950 (hipcc)
  ../oracle/_build/liboracle.so   the CPU oracle (gcc) — test infrastructure

Run: python immsered-boundary_amd/build.py  (or __graft_entry__.build()).
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
INC = os.path.join(ROOT, "include")
LIB = os.path.join(HERE, "lib")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("IBG_ARCH", "gfx950")

# packed FP32 instructions are disabled for every HIP build of this project;
# MFMA accumulators are kept in ordinary VGPRs (no accvgpr moves between the
# network layers and the activation arithmetic)
HIP_FLAGS = ["-O3", "-std=c++17", f"--offload-arch={ARCH}", "-fPIC", "-shared",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-mllvm", "-amdgpu-mfma-vgpr-form"]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines()
                    if "packed-fp32-ops' is not a recognized feature" not in l)
    if out.strip():
        print(out)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}")


def build_host():
    os.makedirs(LIB, exist_ok=True)
    src = os.path.join(HERE, "csrc", "host", "ib_mesh.cpp")
    out = os.path.join(LIB, "libibhost.so")
    _run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fPIC", "-shared", "-fopenmp",
          f"-I{INC}", src, "-o", out + ".tmp"])
    os.replace(out + ".tmp", out)


def build_gpu():
    os.makedirs(LIB, exist_ok=True)
    src = os.path.join(HERE, "csrc", "device", "ibg_engine.hip")
    out = os.path.join(LIB, "libibgpu.so")
    _run([HIPCC] + HIP_FLAGS + [f"-I{INC}", f"-I{os.path.join(HERE, 'csrc', 'device')}",
                                src, "-o", out + ".tmp"])
    os.replace(out + ".tmp", out)


def build_solver_driver():
    """C++ host driver (NS_Solver-shaped class over the C ABI)."""
    src = os.path.join(HERE, "csrc", "host", "ns_solver.cpp")
    if not os.path.exists(src):
        return
    out = os.path.join(LIB, "ibg_run")
    _run(["g++", "-O2", "-std=c++17", f"-I{INC}", f"-I{os.path.join(HERE, 'csrc', 'host')}",
          src, "-o", out + ".tmp", f"-L{LIB}", "-libgpu", "-libhost",
          f"-Wl,-rpath,$ORIGIN"])
    os.replace(out + ".tmp", out)


def build_oracle():
    bdir = os.path.join(ROOT, "oracle", "_build")
    os.makedirs(bdir, exist_ok=True)
    out = os.path.join(bdir, "liboracle.so")
    _run(["gcc", "-O2", "-std=c99", "-ffp-contract=off", "-fPIC", "-shared", "-fopenmp",
          f"-I{INC}", os.path.join(ROOT, "oracle", "oracle.c"),
          os.path.join(ROOT, "oracle", "mlp.c"), os.path.join(ROOT, "oracle", "retag.c"),
          "-o", out + ".tmp", "-lm"])
    os.replace(out + ".tmp", out)


def build_all():
    build_host()
    build_gpu()
    build_oracle()
    build_solver_driver()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    for w in what:
        {"all": build_all, "host": build_host, "gpu": build_gpu,
         "oracle": build_oracle, "driver": build_solver_driver}[w]()
