"""Build the native libraries of the package in-tree.

    python gwfa-cuda-version_amd/build.py [--hip-only|--host-only]

Outputs (git-ignored, shipped with the tree):
    gwfa-cuda-version_amd/lib/libgwfa_host.so   host I/O + synthetic data (g++)
    gwfa-cuda-version_amd/lib/libgwfa_hip.so    HIP kernels + C-ABI (hipcc, gfx950)
    gwfa-cuda-version_amd/lib/gwf-test          gwf-test compatible CLI
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "lib")
INC = os.path.join(ROOT, "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# Packed FP32 instructions are disabled on every HIP translation unit.
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
]
HOST_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-result"]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def _atomic_link(cmd_without_out, out):
    tmp = out + ".tmp"
    _run(cmd_without_out + ["-o", tmp])
    os.replace(tmp, out)


def build_host():
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libgwfa_host.so")
    _atomic_link(["g++", *HOST_FLAGS, "-shared", f"-I{INC}",
                  os.path.join(CSRC, "gwf_host.cpp"), "-lz"], out)
    return out


HIP_SOURCES = ["gwf_kernels.hip", "gwf_runtime.hip", "gwf_serve.hip", "gwf_noalloc.hip"]


def build_hip():
    os.makedirs(LIB, exist_ok=True)
    srcs = [os.path.join(CSRC, s) for s in HIP_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    objs = []
    for s in srcs:
        o = os.path.join(LIB, os.path.basename(s) + ".o")
        _run([HIPCC, *HIP_FLAGS, f"-I{INC}", "-c", s, "-o", o])
        objs.append(o)
    out = os.path.join(LIB, "libgwfa_hip.so")
    _atomic_link([HIPCC, "-shared", f"--offload-arch={ARCH}", *objs, "-lpthread"], out)
    for o in objs:
        os.remove(o)
    return out


def build_cli():
    cli = os.path.join(CSRC, "gwf_test_cli.cpp")
    if not os.path.exists(cli):
        return None
    out = os.path.join(LIB, "gwf-test")
    _atomic_link([HIPCC, "-O2", "-std=c++17", f"-I{INC}", cli,
                  f"-L{LIB}", "-lgwfa_hip", "-lgwfa_host", f"-Wl,-rpath,{LIB}",
                  "-Wl,-rpath,$ORIGIN"], out)
    return out


def build_all():
    build_host()
    build_hip()
    build_cli()


if __name__ == "__main__":
    arg = sys.argv[1] if len(sys.argv) > 1 else ""
    if arg == "--host-only":
        build_host()
    elif arg == "--hip-only":
        build_hip()
    else:
        build_all()
