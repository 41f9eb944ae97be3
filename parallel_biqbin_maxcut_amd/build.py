"""Build of libbiqbin_gpu.so (HIP kernels for gfx950 + the C host runtime).

hipcc compiles every .hip file for gfx950; gcc compiles the C host files;
hipcc links the shared library in place (parallel_biqbin_maxcut_amd/).
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libbiqbin_gpu.so")
CLI_SRC = os.path.join(HERE, "cli", "biqbin_gpu_main.c")
CLI = os.path.join(HERE, "biqbin_gpu")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
    # packed FP32 instructions are not used on this target
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + CSRC, "-I" + os.path.join(ROOT, "include"),
]
CC_FLAGS = [
    "-O2", "-fPIC", "-pthread", "-std=c11", "-D_POSIX_C_SOURCE=200809L", "-D__HIP_PLATFORM_AMD__",
    "-Wall", "-Wextra", "-Wno-unused-parameter",
    "-I" + os.path.join(ROCM, "include"), "-I" + CSRC, "-I" + os.path.join(ROOT, "include"),
]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError("build step failed: " + os.path.basename(cmd[-1]))
    return r


def sources():
    hip = sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))
    c = sorted(f for f in os.listdir(CSRC) if f.endswith(".c"))
    return hip, c


def build(jobs=8, verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    hip, c = sources()
    cmds = []
    objs = []
    for f in hip:
        o = os.path.join(BUILD, f[:-4] + ".o")
        cmds.append([os.path.join(ROCM, "bin", "hipcc")] + HIPCC_FLAGS + ["-c", os.path.join(CSRC, f), "-o", o])
        objs.append(o)
    for f in c:
        o = os.path.join(BUILD, f[:-2] + ".o")
        cmds.append(["gcc"] + CC_FLAGS + ["-c", os.path.join(CSRC, f), "-o", o])
        objs.append(o)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(_run, cmds))
    tmp = LIB + ".tmp"
    _run([os.path.join(ROCM, "bin", "hipcc"), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
         + ["-lpthread"])
    os.replace(tmp, LIB)
    # command-line solver, linked against the library next to it
    _run(["gcc"] + CC_FLAGS + [CLI_SRC, "-o", CLI + ".tmp", "-L" + HERE, "-lbiqbin_gpu",
                                "-Wl,-rpath,$ORIGIN"])
    os.replace(CLI + ".tmp", CLI)
    if verbose:
        print("built", LIB, CLI)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
