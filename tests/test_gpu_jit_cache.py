"""The kernel cache on disk (the reference caches compiled kernels under ~/.cache,
rtcompiler.cc:74-86; path_utils.cc:31-52, 131-150): a first process compiles and writes the code
object, a second loads it without compiling, and a truncated, damaged or foreign file is
compiled again and rewritten, with identical results every time.

Each step runs in a child process (an empty in-memory cache) with NCA_HIP_CACHE_DIR set to a
temporary directory; the child evaluates an expression whose kernel no other test compiles and
reports the cache counters (nca_hip_jit_stats) and the bytes of the result."""
import json
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CHILD = r"""
import json, sys
import numpy as np
sys.path.insert(0, sys.argv[1])
from ncarray_amd import capi
capi.init(0)
P = capi.pack
# ((((a * 3) + a) * a) - 7) / (a + 11): a tape of its own
tape = capi.Tape([P(capi.OP_LOAD_NCARR, 0), P(capi.OP_LOAD_CONST, 0), P(capi.OP_MUL), P(capi.OP_LOAD_NCARR, 0),
                  P(capi.OP_ADD), P(capi.OP_LOAD_NCARR, 0), P(capi.OP_MUL), P(capi.OP_LOAD_CONST, 1), P(capi.OP_SUB),
                  P(capi.OP_LOAD_NCARR, 0), P(capi.OP_LOAD_CONST, 2), P(capi.OP_ADD), P(capi.OP_DIV)],
                 [capi.FLOAT32], [capi.FLOAT32] * 3, capi.FLOAT32, capi.FLOAT32)
n = 4099
x = np.linspace(-3, 3, n).astype(np.float32)
nb = n * 4
dx, do = capi.ctypes.c_void_p(), capi.ctypes.c_void_p()
capi.call("nca_hip_malloc", capi.ctypes.byref(dx), nb)
capi.call("nca_hip_malloc", capi.ctypes.byref(do), nb)
capi.call("nca_hip_memcpy_h2d", dx, x.ctypes.data, nb, None)
tape.launch([capi.make_view(dx.value, capi.FLOAT32, [n])], capi.make_view(do.value, capi.FLOAT32, [n]),
            scalars=(3.0, 7.0, 11.0))
out = np.empty(n, np.float32)
capi.call("nca_hip_memcpy_d2h", out.ctypes.data, do, nb, None)
capi.call("nca_hip_free", dx)
capi.call("nca_hip_free", do)
c, d, m = capi.jit_stats()
print(json.dumps({"compiles": c, "disk": d, "memory": m, "out": out.tobytes().hex()}))
"""


def run_child(cache_dir):
    env = dict(os.environ, NCA_HIP_CACHE_DIR=str(cache_dir))
    r = subprocess.run([sys.executable, "-c", CHILD, ROOT], env=env, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    return json.loads(r.stdout.strip().splitlines()[-1])


def test_disk_cache_hit_and_damaged_files(dev, tmp_path):
    first = run_child(tmp_path)
    assert first["compiles"] == 1 and first["disk"] == 0
    files = sorted(p for p in tmp_path.iterdir() if p.suffix == ".hsaco")
    assert len(files) == 1
    path = files[0]
    good = path.read_bytes()
    assert good[:8] == b"NCAHSCO2" and good[8:72] == path.stem.encode()   # magic, then the source key
    print("cache file written", path.name, len(good), flush=True)

    second = run_child(tmp_path)                       # a fresh process loads the file
    assert second["compiles"] == 0 and second["disk"] == 1
    assert second["out"] == first["out"]

    cases = {
        "truncated": good[: len(good) // 2],
        "damaged": good[:200] + bytes(b ^ 0x5A for b in good[200:400]) + good[400:],
        "foreign": good[:8] + b"0" * 64 + good[72:],    # another source's key
        "empty": b"",
    }
    for what, data in cases.items():
        path.write_bytes(data)
        r = run_child(tmp_path)
        assert r["compiles"] == 1 and r["disk"] == 0, what
        assert r["out"] == first["out"], what
        again = path.read_bytes()                           # rewritten whole, under the same key
        assert again[:72] == good[:72] and len(again) > 136, what
        print("recompiled after a", what, "file", flush=True)
    last = run_child(tmp_path)                              # and the rewritten file serves again
    assert last["compiles"] == 0 and last["disk"] == 1 and last["out"] == first["out"]
