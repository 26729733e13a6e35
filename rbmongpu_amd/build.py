"""In-tree build of the gfx950 library and its Python module.

  librbmhip.so       hipcc --offload-arch=gfx950: kernels + C ABI (include/rbmhip.h)
  _pyRBMonGPU*.so    g++: pybind11 module over the C ABI, rpath $ORIGIN

Run `python rbmongpu_amd/build.py` (or __graft_entry__.build(); the script is loaded by path so
the package itself, which needs the built module, is not imported).  Objects go to
rbmongpu_amd/_build/; the two shared libraries land in rbmongpu_amd/.
"""
from __future__ import annotations

import os
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "librbmhip.so")

HIP_SOURCES = ["rbm_kernels.hip", "deep_kernels.hip", "deep_mfma.hip", "made_kernels.hip", "linalg_kernels.hip",
               "gram_kernels.hip", "noalloc.hip", "capi.cpp"]
# Packed FP32 instructions are kept out of the generated code.
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-result"]


def _hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.isabs(cand) and os.path.exists(cand) or not os.path.isabs(cand)):
            return cand
    return "hipcc"


def _run_parallel(cmds):
    procs = [(c, subprocess.Popen(c, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)) for c in cmds]
    failed = []
    for c, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            failed.append((c, out))
    if failed:
        for c, out in failed:
            sys.stderr.write(" ".join(c) + "\n" + out + "\n")
        raise RuntimeError("build failed")


def ext_suffix():
    return sysconfig.get_config_var("EXT_SUFFIX") or ".so"


def module_path():
    return os.path.join(HERE, "_pyRBMonGPU" + ext_suffix())


def build(verbose: bool = False) -> None:
    os.makedirs(BUILD, exist_ok=True)
    hipcc = _hipcc()
    objs = []
    cmds = []
    for src in HIP_SOURCES:
        obj = os.path.join(BUILD, src.rsplit(".", 1)[0] + ".o")
        objs.append(obj)
        cmds.append([hipcc, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
    _run_parallel(cmds)
    tmp = LIB + ".tmp"
    subprocess.run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs,
                    "-L/opt/rocm/lib", "-lrocprofiler-sdk-roctx", "-Wl,-rpath,/opt/rocm/lib"], check=True)
    os.replace(tmp, LIB)

    import pybind11
    py_inc = sysconfig.get_paths()["include"]
    mod = module_path()
    tmp = mod + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-fvisibility=hidden",
           "-I", pybind11.get_include(), "-I", py_inc, "-I", os.path.join(ROOT, "include"),
           os.path.join(CSRC, "pymodule.cpp"), "-o", tmp,
           "-L", HERE, "-lrbmhip", "-Wl,-rpath,$ORIGIN"]
    subprocess.run(cmd, check=True)
    os.replace(tmp, mod)
    if verbose:
        print("built", LIB, "and", mod)


if __name__ == "__main__":
    build(verbose=True)
