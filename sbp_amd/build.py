"""Builds the gfx950 shared library ``sbp_amd/libsbp_amd.so`` (HIP kernels + C-ABI + C++ drop-in API) and the
command-line driver ``sbp_amd/sbp_cli``.

Every translation unit is compiled with FP contraction off, so each FP64 operation in the delta-MDL and
MDL terms is a single IEEE rounding (the fixed-point sums then make them bit-identical to a CPU
evaluation), and with packed-FP32 instructions disabled.
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libsbp_amd.so")
SOURCES = ["csrc/build.hip", "csrc/mdl.hip", "csrc/finetune.hip", "csrc/merge.hip", "csrc/graph.hip", "csrc/split.hip", "csrc/update.hip",
           "csrc/capi.hip",
           "host/sbp_api.cpp", "host/reference_api.cpp", "host/args.cpp", "host/batch.cpp", "host/evaluate.cpp"]
INCLUDE_SBP = os.path.join(os.path.dirname(HERE), "include", "sbp")  # the reference's header names
CLI = os.path.join(HERE, "sbp_cli")
API_CHECK = os.path.join(os.path.dirname(HERE), "tests", "dropin", "api_check")
CARRY_LIB = os.path.join(HERE, "libsbp_carry.so")  # host-only carry-out, for CPU-side checks

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("SBP_OFFLOAD_ARCH", "gfx950")
FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-result",
]


def _compile(src):
    obj = os.path.join(BUILD, os.path.basename(os.path.splitext(src)[0]) + ".o")
    cmd = [HIPCC, *FLAGS, "-c", os.path.join(HERE, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(len(SOURCES), os.cpu_count() or 4)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    cmd = [HIPCC, *FLAGS, "-I" + INCLUDE_SBP, "-o", CLI + ".tmp", os.path.join(HERE, "host", "main.cpp"), "-L" + HERE,
           "-lsbp_amd", "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"CLI link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(CLI + ".tmp", CLI)
    # the drop-in API check program (tests/dropin/api_check.cpp), built against the reference's header names
    cmd = [HIPCC, *FLAGS, "-I" + INCLUDE_SBP, "-o", API_CHECK + ".tmp", API_CHECK + ".cpp", "-L" + HERE, "-lsbp_amd",
           "-Wl,-rpath,$ORIGIN/../../sbp_amd"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"api_check link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(API_CHECK + ".tmp", API_CHECK)
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-pthread", "-ffp-contract=off", "-o", CARRY_LIB + ".tmp",
           os.path.join(HERE, "host", "carry_capi.cpp")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"host carry-out build failed:\n{r.stdout}\n{r.stderr}")
    os.replace(CARRY_LIB + ".tmp", CARRY_LIB)
    if verbose:
        print(f"built {LIB} and {CLI}")
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
