"""Builds libnca_hip.so (HIP kernels + hipRTC JIT + C ABI) in-tree for gfx950.

The HIP sources are compiled with hipcc, one object per source, then linked against
libhiprtc.  All device code is built with packed fp32 instructions disabled and without
floating-point contraction (generated expressions must evaluate a*b+c as two roundings).
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(ROOT, "_build")
LIB = os.path.join(HERE, "libnca_hip.so")

SOURCES = ["runtime.cpp", "jit.cpp", "tape.cpp", "expr.cpp", "stencil.cpp", "batch.cpp",
           "reduce.hip", "bin.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("PYTORCH_ROCM_ARCH", "gfx950")

FLAGS = [
    "-std=c++20", "-O3", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-ffp-contract=off", "-Wno-c23-extensions", "-Wno-unused-command-line-argument",
    f"-I{CSRC}", f"-I{os.path.join(HERE, 'include')}", f"-I{os.path.join(ROOT, 'include')}",
]


def _compile(src):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    path = os.path.join(CSRC, src)
    cmd = [HIPCC] + FLAGS + (["-x", "hip"] if src.endswith(".cpp") else []) + ["-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {src}\n{r.stderr}")
    return obj


HOST = os.path.join(HERE, "host")
HOST_LIB = os.path.join(HERE, "libnca_host.so")
CXX = os.environ.get("CXX", "g++")
HOST_FLAGS = ["-std=c++20", "-O2", "-fPIC", "-Wall", "-Wno-unused-result",
              f"-I{os.path.join(ROOT, 'include')}", f"-I{os.path.join(HERE, 'include')}", f"-I{HOST}"]


def _run(cmd, what):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"{what} failed\n{' '.join(cmd)}\n{r.stderr}")


def build_host(verbose=False):
    """The C++ host API (libnca_host.so) and its Python module (_ncarray), host code only,
    linked against libnca_hip.so in the same folder."""
    import sysconfig
    import pybind11
    tmp = HOST_LIB + ".tmp"
    _run([CXX] + HOST_FLAGS + ["-shared", os.path.join(HOST, "ncarray_amd.cpp"), "-o", tmp,
          f"-L{HERE}", "-lnca_hip", "-Wl,-rpath,$ORIGIN"], "libnca_host.so")
    os.replace(tmp, HOST_LIB)
    mod = os.path.join(HERE, "_ncarray" + sysconfig.get_config_var("EXT_SUFFIX"))
    tmp = mod + ".tmp"
    _run([CXX] + HOST_FLAGS + ["-shared", f"-I{pybind11.get_include()}", f"-I{sysconfig.get_paths()['include']}",
          "-fvisibility=hidden", os.path.join(HOST, "pyncarray.cpp"), "-o", tmp,
          f"-L{HERE}", "-lnca_host", "-lnca_hip", "-Wl,-rpath,$ORIGIN"], "_ncarray module")
    os.replace(tmp, mod)
    test_src = os.path.join(ROOT, "tests", "cpp", "test_host_api.cpp")
    if os.path.exists(test_src):
        exe = os.path.join(ROOT, "tests", "cpp", "test_host_api")
        # the test also creates HIP streams itself (HIP's host API, as hipcc compiles host code)
        _run([CXX] + HOST_FLAGS + ["-I/opt/rocm/include", "-D__HIP_PLATFORM_AMD__", test_src, "-o", exe + ".tmp",
              f"-L{HERE}", "-lnca_host", "-lnca_hip", "-L/opt/rocm/lib", "-lamdhip64",
              f"-Wl,-rpath,{HERE}", "-Wl,-rpath,/opt/rocm/lib"], "C++ host API test")
        os.replace(exe + ".tmp", exe)
    if verbose:
        print("built", HOST_LIB, mod)


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    jobs = min(len(SOURCES), max(1, int(os.environ.get("MAX_JOBS", os.cpu_count() or 4))))
    with ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs + [
        "-L/opt/rocm/lib", "-lhiprtc", "-Wl,-rpath,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    build_host(verbose)
    return LIB


if __name__ == "__main__":
    if len(sys.argv) > 1 and sys.argv[1] == "host":
        build_host(verbose=True)
    else:
        build(verbose=True)
    sys.exit(0)
