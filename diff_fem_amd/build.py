"""Build libpdm.so (HIP for gfx950 + host C++) in-tree with hipcc."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libpdm.so")
SOURCES = ["pdm_setup.cpp", "pdm_engine.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("PDM_OFFLOAD_ARCH", "gfx950")
FLAGS = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def build(verbose=False):
    objs = []
    build_dir = os.path.join(HERE, "_build")
    os.makedirs(build_dir, exist_ok=True)
    for src in SOURCES:
        obj = os.path.join(build_dir, os.path.splitext(src)[0] + ".o")
        cmd = [HIPCC, "--offload-arch=" + ARCH, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
        if src.endswith(".hip"):
            cmd[1:1] = ["-x", "hip"]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(obj)
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", OUT + ".tmp", *objs]
    subprocess.run(cmd, check=True)
    os.replace(OUT + ".tmp", OUT)
    build_pybind(verbose)
    return OUT


def pybind_module_path():
    import sysconfig
    return os.path.join(HERE, "pybind", "pdm_pybind" + sysconfig.get_config_var("EXT_SUFFIX"))


def build_pybind(verbose=False):
    """The reference-side pybind11 binding over the C ABI (pybind/pdm_pybind.cc),
    linked against libpdm.so (host C++ only, no GPU code of its own)."""
    import sysconfig
    import pybind11
    out = pybind_module_path()
    cmd = [os.environ.get("CXX", "g++"), "-O2", "-std=c++17", "-shared", "-fPIC",
           os.path.join(HERE, "pybind", "pdm_pybind.cc"), "-I", os.path.join(os.path.dirname(HERE), "include"),
           "-I", sysconfig.get_paths()["include"], "-I", pybind11.get_include(), "-L", HERE, "-lpdm",
           "-Wl,-rpath,$ORIGIN/..", "-o", out + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(out + ".tmp", out)
    return out


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
