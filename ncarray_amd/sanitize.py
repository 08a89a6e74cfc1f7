"""AddressSanitizer + UndefinedBehaviorSanitizer run of the library's host code on the CPU.

Copies the package, the headers, the oracle and the tests into OUTDIR (default
/tmp/nca_sanitize), builds libnca_hip.so with its host code instrumented (device code is not:
the sanitizer flags go to the host pass only) and libnca_host.so / _ncarray with clang++ and the
same sanitizers, then runs the CPU tests there with the sanitizer runtime preloaded into the
(uninstrumented) Python interpreter.  Leak detection is off (the interpreter and torch keep
their allocations to the end); any address error or undefined behaviour ends the run.

    python -m ncarray_amd.sanitize [OUTDIR] [pytest arguments ...]

Needs no GPU.  The oracle is built without sanitizers (it is the checker, not the product).
"""
import glob
import os
import shutil
import subprocess
import sys
import sysconfig
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CLANG = "/opt/rocm/llvm/bin/clang++"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HOST_SAN = ["-fsanitize=address", "-fsanitize=undefined", "-fno-sanitize-recover=undefined",
            "-fno-sanitize=vptr", "-fno-omit-frame-pointer", "-g"]


def _runtime():
    libs = sorted(glob.glob("/opt/rocm/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so"))
    if not libs:
        raise RuntimeError("clang's AddressSanitizer runtime was not found under /opt/rocm/llvm")
    return libs[-1]


def _run(cmd, what, **kw):
    r = subprocess.run(cmd, capture_output=True, text=True, **kw)
    if r.returncode != 0:
        raise RuntimeError(f"{what} failed\n{' '.join(cmd)}\n{r.stdout[-4000:]}\n{r.stderr[-4000:]}")


def _copy_tree(out):
    if os.path.isdir(out):
        shutil.rmtree(out)
    ignore = shutil.ignore_patterns("*.so", "*.o", "__pycache__", "*.tmp", "test_host_api")
    shutil.copytree(HERE, os.path.join(out, "ncarray_amd"), ignore=ignore)
    for d in ("include", "oracle", "tests"):
        shutil.copytree(os.path.join(ROOT, d), os.path.join(out, d), ignore=ignore)


def build(out):
    """Instrumented libraries in OUTDIR/ncarray_amd; returns that folder."""
    sys.path.insert(0, HERE)
    import build as nb  # the flag and source lists of the regular build
    pkg = os.path.join(out, "ncarray_amd")
    csrc, objdir = os.path.join(pkg, "csrc"), os.path.join(out, "_build")
    os.makedirs(objdir, exist_ok=True)
    inc = [f"-I{csrc}", f"-I{os.path.join(pkg, 'include')}", f"-I{os.path.join(out, 'include')}"]
    flags = [f for f in nb.FLAGS if not f.startswith("-I") and f != "-O3"] + ["-O1"] + inc
    host_pass = []
    for f in HOST_SAN:
        host_pass += ["-Xarch_host", f]

    def compile_one(src):
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        lang = ["-x", "hip"] if src.endswith(".cpp") else []
        _run([HIPCC] + flags + host_pass + lang + ["-c", os.path.join(csrc, src), "-o", obj], f"compile {src}")
        return obj

    jobs = max(1, min(len(nb.SOURCES), int(os.environ.get("MAX_JOBS", os.cpu_count() or 4))))
    with ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(compile_one, nb.SOURCES))
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={nb.ARCH}", "-o", os.path.join(pkg, "libnca_hip.so")] + objs +
         ["-L/opt/rocm/lib", "-lhiprtc", "-Wl,-rpath,/opt/rocm/lib"], "link libnca_hip.so")
    import pybind11
    host = os.path.join(pkg, "host")
    hflags = ["-std=c++20", "-O1", "-fPIC", "-shared-libsan", f"-I{os.path.join(out, 'include')}",
              f"-I{os.path.join(pkg, 'include')}", f"-I{host}"] + HOST_SAN
    _run([CLANG] + hflags + ["-shared", os.path.join(host, "ncarray_amd.cpp"), "-o", os.path.join(pkg, "libnca_host.so"),
          f"-L{pkg}", "-lnca_hip", "-Wl,-rpath,$ORIGIN"], "libnca_host.so")
    mod = os.path.join(pkg, "_ncarray" + sysconfig.get_config_var("EXT_SUFFIX"))
    _run([CLANG] + hflags + ["-shared", f"-I{pybind11.get_include()}", f"-I{sysconfig.get_paths()['include']}",
          "-fvisibility=hidden", os.path.join(host, "pyncarray.cpp"), "-o", mod, f"-L{pkg}", "-lnca_host", "-lnca_hip",
          "-Wl,-rpath,$ORIGIN"], "_ncarray module")
    _run(["make", "-B", "-C", os.path.join(out, "oracle")], "oracle")
    return pkg


def main(argv):
    out = argv[0] if argv and not argv[0].startswith("-") else "/tmp/nca_sanitize"
    rest = argv[1:] if argv and not argv[0].startswith("-") else argv
    _copy_tree(out)
    build(out)
    env = dict(os.environ)
    env["LD_PRELOAD"] = _runtime()
    env["ASAN_OPTIONS"] = "detect_leaks=0:abort_on_error=1:halt_on_error=1:detect_odr_violation=0"
    env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1"
    # -s: a sanitizer report goes to stderr right before the abort, which pytest's capture would drop
    cmd = [sys.executable, "-m", "pytest", "tests", "-q", "-x", "-s", "-m", "not gpu", "-p", "no:cacheprovider"] + rest
    print("running", " ".join(cmd), "in", out, flush=True)
    return subprocess.run(cmd, cwd=out, env=env).returncode


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
