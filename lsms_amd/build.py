"""Build the native library lsms_amd/liblsmsx.so for gfx950 with hipcc.

    python -m lsms_amd.build

Objects go to lsms_amd/_build/; the shared library is written in-tree next
to this file so it travels with the source tree.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "liblsmsx.so")
DRIVER = os.path.join(HERE, "_build", "host_mirror_driver")
SS_DRIVER = os.path.join(HERE, "_build", "single_site_driver")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_SOURCES = ["kkr_assembly.hip", "tau_lu.hip", "ir.hip", "inverse.hip", "lsmsx_capi.hip",
               "update_f16.hip", "single_site.hip", "relativistic.hip"]
# sources whose MFMA accumulators stay in AGPRs (no -amdgpu-mfma-vgpr-form)
AGPR_SOURCES = {"update_f16.hip"}
CXX_SOURCES = ["tables.cpp"]

HIP_FLAGS = [
    "-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}",
    # packed FP32 instructions are kept out of every device code object
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]
# MFMA accumulators in ordinary VGPRs (unified register file): no accumulator
# copies between AGPRs and VGPRs around the K loops (all sources but
# AGPR_SOURCES, whose accumulators are too many for the VGPRs)
VGPR_FORM = ["-mllvm", "-amdgpu-mfma-vgpr-form=1"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r


def build(verbose=False):
    os.makedirs(OUT, exist_ok=True)
    jobs = []
    for s in HIP_SOURCES:
        flags = HIP_FLAGS + ([] if s in AGPR_SOURCES else VGPR_FORM)
        jobs.append([HIPCC, *flags, "-c", os.path.join(SRC, s), "-o", os.path.join(OUT, s + ".o")])
    for s in CXX_SOURCES:
        jobs.append(["g++", "-O2", "-fPIC", "-std=c++17", "-c", os.path.join(SRC, s), "-o",
                     os.path.join(OUT, s + ".o")])
    with ThreadPoolExecutor(max_workers=4) as ex:
        for r in ex.map(_run, jobs):
            if verbose and (r.stdout or r.stderr):
                print(r.stdout, r.stderr, file=sys.stderr)
    objs = [os.path.join(OUT, s + ".o") for s in HIP_SOURCES + CXX_SOURCES]
    tmp = LIB + ".tmp"
    # rocSOLVER (with rocBLAS for its handle): the column-pivoted fallback
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", tmp, *objs, "-L/opt/rocm/lib",
          "-lrocsolver", "-lrocblas"])
    os.replace(tmp, LIB)
    # C++ driver of the LSMS-shaped host interface (run by the GPU tests)
    drv = os.path.join(os.path.dirname(HERE), "tests", "cpp", "host_mirror_driver.cpp")
    if os.path.exists(drv):
        _run(["g++", "-O2", "-std=c++17", "-I/opt/rocm/include", drv, "-o", DRIVER,
              f"-L{HERE}", "-llsmsx", "-Wl,-rpath,$ORIGIN/..",
              "-L/opt/rocm/lib", "-lamdhip64"])
    drv2 = os.path.join(os.path.dirname(HERE), "tests", "cpp", "single_site_driver.cpp")
    if os.path.exists(drv2):
        _run(["g++", "-O2", "-std=c++17", "-I/opt/rocm/include", drv2, "-o", SS_DRIVER,
              f"-L{HERE}", "-llsmsx", "-Wl,-rpath,$ORIGIN/..",
              "-L/opt/rocm/lib", "-lamdhip64"])
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
