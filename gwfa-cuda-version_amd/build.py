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
# gwf_kernels.hip is compiled once per state-key layout (the split of a
# state's 64-bit key between vertex, diagonal and offset; gwf_dev.h); the
# layout-0 object also holds the layout-independent code.
KEY_LAYOUTS = 4


def _compile_jobs():
    jobs = []
    for name in HIP_SOURCES:
        src = os.path.join(CSRC, name)
        if name == "gwf_kernels.hip":
            for lay in range(KEY_LAYOUTS):
                jobs.append((src, os.path.join(LIB, f"{name}.l{lay}.o"), [f"-DGWF_KEY_LAYOUT={lay}"]))
        else:
            jobs.append((src, os.path.join(LIB, name + ".o"), []))
    return jobs


def build_hip():
    os.makedirs(LIB, exist_ok=True)
    jobs = _compile_jobs()
    width = max(1, min(len(jobs), int(os.environ.get("MAX_JOBS", os.cpu_count() or 4))))
    running, objs, failed = [], [], None
    for src, obj, extra in jobs:
        cmd = [HIPCC, *HIP_FLAGS, *extra, f"-I{INC}", "-c", src, "-o", obj]
        print("+", " ".join(cmd), flush=True)
        running.append((subprocess.Popen(cmd), cmd))
        objs.append(obj)
        while len(running) >= width:
            p, c = running.pop(0)
            if p.wait() != 0 and failed is None:
                failed = c
    for p, c in running:
        if p.wait() != 0 and failed is None:
            failed = c
    if failed is not None:
        raise subprocess.CalledProcessError(1, failed)
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
