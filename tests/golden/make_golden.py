"""Writes the golden fixtures of this folder (run from the repository root):

    python tests/golden/make_golden.py

reference_kats.json: the known-answer cases of the reference's own tests, as data (inputs,
  operation, expected values), each with the test file and lines it comes from.
oracle_vectors.npz: small seeded inputs with the CPU oracle's outputs and FP64 values for the
  hot-path operations (C0 expression and sum, 3-D Laplacian with zero boundary, axis
  reductions, argmax with ties, per-frame binning and its backward).  The oracle itself is
  pinned by tests/test_oracle.py; these vectors freeze its outputs for the GPU parity tests.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

KATS = [
    {"source": "tests/cxx/test_gpu_reductions.cu:39-55", "kind": "full_reduce", "dtype": "float32",
     "init": {"iota": [256]}, "expect": {"sum": 32640.0, "max": 255.0, "min": 0.0}},
    {"source": "tests/cxx/test_gpu_reductions.cu:57-120", "kind": "axis_reduce", "dtype": "int32",
     "init": {"iota_plus": [[2, 3], 1]},
     "expect": {"sum0": [5, 7, 9], "max0": [4, 5, 6], "min0": [1, 2, 3], "sum1": [6, 15], "max1": [3, 6],
                "min1": [1, 4]}},
    {"source": "tests/cxx/test_gpu_reductions.cu:122-142", "kind": "sliced_fill_sum", "dtype": "float32",
     "init": {"fill": [[4, 4], 1.0], "slice_fill": [[1, 3], [1, 3], 10.0]}, "expect": {"sum": 52.0}},
    {"source": "tests/cxx/test_gpu_reductions.cu:144-156", "kind": "copy_astype", "dtype": "float32",
     "init": {"fill": [[10], 5.7]}, "to": "int32", "expect": {"values": [5] * 10}},
    {"source": "tests/cxx/test_reductions.cc:90-102", "kind": "full_reduce", "dtype": "float32",
     "init": {"fill": [[5], 1.5]}, "expect": {"sum": 7.5}},
    {"source": "tests/python/test_operations.py:56-62", "kind": "full_reduce", "dtype": "int32",
     "init": {"values": [[1, 2, 3], [4, 5, 6]]}, "expect": {"sum": 21, "max": 6, "min": 1}},
    {"source": "tests/python/test_gpu_operations.py:38-62", "kind": "full_reduce", "dtype": "uint32",
     "init": {"iota": [256]}, "expect": {"sum": 32640, "max": 255, "min": 0}},
    {"source": "tests/cxx/test_gpu_binary_operations.cu:28-55", "kind": "binary", "dtype": "float32",
     "a": 2.0, "b": 4.0, "op": "add", "expect": {"dtype": "float32", "value": 6.0}},
    {"source": "tests/python/test_gpu_operations.py:16-35", "kind": "binary", "dtype": "float32",
     "a": 2.0, "b": 4.0, "op": "mul", "expect": {"dtype": "float32", "value": 8.0}},
    {"source": "tests/cxx/test_gpu_binary_operations.cu:57-84", "kind": "binary", "dtype": "float32",
     "a": 2.0, "b": 4.0, "op": "truediv", "expect": {"dtype": "float64", "value": 0.5}},
    {"source": "tests/cxx/test_gpu_binary_operations.cu:114-141", "kind": "inplace", "dtype": "float32",
     "a": 6.0, "b": 4.0, "op": "itruediv", "expect": {"dtype": "float32", "value": 1.5}},
    {"source": "tests/cxx/test_gpu_binary_operations.cu:143-158", "kind": "scalar", "dtype": "float32",
     "a": 10.0, "b": 5.0, "op": "add", "expect": {"dtype": "float32", "value": 15.0}},
    {"source": "tests/python/test_gpu_operations.py:86-95", "kind": "scalar_chain", "dtype": "float32",
     "a": 10.0, "ops": [["iadd", 5.0], ["imul", 2.0]], "expect": {"value": 30.0}},
    {"source": "tests/cxx/test_gpu_binary_operations.cu:176-197", "kind": "values_binary", "dtype": "uint32",
     "a": [10, 11, 12, 13, 14], "b": [3, 3, 3, 3, 3], "op": "mod", "expect": {"values": [1, 2, 0, 1, 2]}},
    {"source": "tests/cxx/test_gpu_binary_operations.cu:199-217", "kind": "values_scalar", "dtype": "float32",
     "a": [0, 1, 2, 3, 4, 10, 10, 10, 10, 10], "b": 5.0, "op": "gt",
     "expect": {"dtype": "bool", "values": [False] * 5 + [True] * 5}},
    {"source": "tests/python/test_gpu_operations.py:131-156", "kind": "values_scalar", "dtype": "float32",
     "a": list(range(10)), "b": 5.0, "op": "ge", "expect": {"dtype": "bool", "values": [False] * 5 + [True] * 5}},
    {"source": "tests/cxx/test_gpu_unary_operations.cu:19-44", "kind": "unary", "dtype": "float32", "a": 5.0,
     "expect": {"neg": -5.0, "inc": 6.0, "dec": 4.0}},
    {"source": "tests/cxx/test_gpu_unary_operations.cu:46-93", "kind": "iota", "dtype": "int32", "shape": [3, 3],
     "expect": {"iota": list(range(9)), "iota_times2_plus1": [2 * i + 1 for i in range(9)]}},
]


def oracle_vectors():
    from oracle import oracle as O
    from ncarray_amd import tapes
    rng = np.random.default_rng(20240101)
    v = {}
    n = 4096
    a = rng.uniform(-1, 1, n).astype(np.float32)
    b = rng.uniform(-1, 1, n).astype(np.float32)
    v["c0_a"], v["c0_b"] = a, b
    v["c0_out"] = O.expr(tapes.fused_c0(), [a, b, a, a, b], [0.5], 10, np.float32)
    a64, b64 = a.astype(np.float64), b.astype(np.float64)
    v["c0_fp64"] = a64 * b64 + np.exp(-a64 * a64) - 0.5 * b64
    v["c0_sum_fp64"] = np.array(v["c0_out"].astype(np.float64).sum())
    x = rng.standard_normal((8, 12, 16)).astype(np.float32)
    v["lap_x"] = x
    v["lap_out"] = O.stencil_f32(x, x.shape, tapes.LAPLACIAN_OFFSETS, tapes.laplacian(), [6.0], True)
    p = np.pad(x.astype(np.float64), 1)
    v["lap_fp64"] = (p[:-2, 1:-1, 1:-1] + p[2:, 1:-1, 1:-1] + p[1:-1, :-2, 1:-1] + p[1:-1, 2:, 1:-1]
                     + p[1:-1, 1:-1, :-2] + p[1:-1, 1:-1, 2:] - 6.0 * p[1:-1, 1:-1, 1:-1])
    r = rng.integers(-20, 20, (5, 6, 7)).astype(np.float32)
    r[2, 3, 4] = 100.0
    r[4, 3, 4] = 100.0   # argmax tie along axis 0: first occurrence wins
    v["red_x"] = r
    v["red_sum0"] = O.reduce_axes_f32(r, [0], 0)
    v["red_max2"] = O.reduce_axes_f32(r, [2], 1)
    v["red_argmax0"] = O.reduce_axes_f32(r, [0], 3)
    v["red_sum_all"] = np.array(O.sum_fold(r))
    v["red_argmax_all"] = np.array(O.max_fold(r)[1])
    fr = rng.exponential(50.0, (2, 16, 16)).astype(np.float32)
    fr[:, 3, :] = 0.0
    bm = O.radial_binmap(16, 16, 8, 8, 10)
    bm[0, 0] = -1
    v["bin_frames"], v["bin_map"] = fr, bm
    c, s = O.bin_f32(fr, bm, 10, nonzero=True)
    v["bin_counts"], v["bin_sums"] = c, s
    g = rng.standard_normal((2, 10)).astype(np.float32)
    v["bin_grad"] = g
    v["bin_grad_frames"] = O.bin_backward_f32(g, bm)
    return v


def main():
    with open(os.path.join(HERE, "reference_kats.json"), "w") as f:
        json.dump(KATS, f, indent=1)
    np.savez_compressed(os.path.join(HERE, "oracle_vectors.npz"), **oracle_vectors())
    print("wrote", HERE)


if __name__ == "__main__":
    main()
