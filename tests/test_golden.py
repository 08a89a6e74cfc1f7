"""Golden fixtures (tests/golden/): the reference's known-answer cases and frozen oracle vectors.

CPU: the oracle still reproduces its frozen vectors bit for bit, and every transcribed
known answer agrees with numpy.  GPU: each known-answer case runs through the Python array
API, and the HIP kernels match the frozen vectors (bit-exact for stencils, counts, integer-
valued sums, max and argmax; stated bounds for exp and float sums).
"""
import json
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
sys.path.insert(0, GOLDEN)

KATS = json.load(open(os.path.join(GOLDEN, "reference_kats.json")))
VEC = dict(np.load(os.path.join(GOLDEN, "oracle_vectors.npz")))
NP = {"float32": np.float32, "float64": np.float64, "int32": np.int32, "uint32": np.uint32, "bool": np.bool_}
OPS = {"add": np.add, "sub": np.subtract, "mul": np.multiply, "truediv": np.true_divide, "mod": np.mod,
       "gt": np.greater, "ge": np.greater_equal}


def _host_initial(case):
    dt = NP[case["dtype"]]
    init = case.get("init", {})
    if "iota" in init:
        return np.arange(int(np.prod(init["iota"])), dtype=dt).reshape(init["iota"])
    if "iota_plus" in init:
        shape, k = init["iota_plus"]
        return (np.arange(int(np.prod(shape))) + k).astype(dt).reshape(shape)
    if "values" in init:
        return np.array(init["values"], dtype=dt)
    if "fill" in init:
        shape, v = init["fill"]
        a = np.full(shape, v, dtype=dt)
        if "slice_fill" in init:
            (r0, r1), (c0, c1), w = init["slice_fill"]
            a[r0:r1, c0:c1] = w
        return a
    raise KeyError(init)


# ---- CPU ---------------------------------------------------------------------------------------------
def test_oracle_reproduces_frozen_vectors():
    import make_golden
    fresh = make_golden.oracle_vectors()
    assert set(fresh) == set(VEC)
    for k, v in fresh.items():
        assert np.array_equal(v, VEC[k], equal_nan=True), k


@pytest.mark.parametrize("case", KATS, ids=[c["source"] for c in KATS])
def test_known_answers_agree_with_numpy(case):
    kind, exp = case["kind"], case["expect"]
    if kind in ("full_reduce", "sliced_fill_sum"):
        a = _host_initial(case)
        for k, want in exp.items():
            assert getattr(a, k)() == want
    elif kind == "axis_reduce":
        a = _host_initial(case)
        for k, want in exp.items():
            assert getattr(a, k[:-1])(axis=int(k[-1])).tolist() == want
    elif kind == "copy_astype":
        assert _host_initial(case).astype(NP[case["to"]]).tolist() == exp["values"]
    elif kind in ("binary", "scalar", "inplace"):
        dt = NP[case["dtype"]]
        r = OPS[case["op"].lstrip("i")](dt(case["a"]), dt(case["b"]))
        assert float(r) == exp["value"]
    elif kind == "scalar_chain":
        v = np.float32(case["a"])
        for op, s in case["ops"]:
            v = OPS[{"iadd": "add", "imul": "mul"}[op]](v, np.float32(s))
        assert v == exp["value"]
    elif kind in ("values_binary", "values_scalar"):
        dt = NP[case["dtype"]]
        b = np.array(case["b"], dtype=dt) if kind == "values_binary" else dt(case["b"])
        assert OPS[case["op"]](np.array(case["a"], dtype=dt), b).tolist() == exp["values"]
    elif kind == "unary":
        assert -case["a"] == exp["neg"] and case["a"] + 1 == exp["inc"] and case["a"] - 1 == exp["dec"]
    elif kind == "iota":
        n = int(np.prod(case["shape"]))
        assert list(range(n)) == exp["iota"] and [2 * i + 1 for i in range(n)] == exp["iota_times2_plus1"]
    else:
        raise KeyError(kind)


# ---- GPU ---------------------------------------------------------------------------------------------
def _device_initial(nca, case):
    dt = getattr(nca.DType, case["dtype"])
    init = case["init"]
    if "iota" in init or "iota_plus" in init:
        shape = init["iota"] if "iota" in init else init["iota_plus"][0]
        a = nca.NCDevArray(shape, dt)
        with nca.lazy_mode():
            a[...] = a.iota() if "iota" in init else a.iota() + init["iota_plus"][1]
        return a
    if "fill" in init:
        shape, v = init["fill"]
        a = nca.NCDevArray(shape, dt)
        a.fill(v)
        if "slice_fill" in init:
            (r0, r1), (c0, c1), w = init["slice_fill"]
            a[r0:r1, c0:c1] = w
        return a
    return nca.NCDevArray.from_host(np.array(init["values"], dtype=NP[case["dtype"]]))


@pytest.mark.gpu
@pytest.mark.parametrize("case", KATS, ids=[c["source"] for c in KATS])
def test_known_answers_on_device(dev, case):
    import ncarray_amd as nca
    nca.set_eager(True)
    kind, exp = case["kind"], case["expect"]
    dt = getattr(nca.DType, case["dtype"])
    if kind in ("full_reduce", "sliced_fill_sum"):
        a = _device_initial(nca, case)
        for k, want in exp.items():
            assert getattr(a, k)() == want, k
    elif kind == "axis_reduce":
        a = _device_initial(nca, case)
        for k, want in exp.items():
            assert np.array(getattr(a, k[:-1])([int(k[-1])])).tolist() == want, k
    elif kind == "copy_astype":
        a = _device_initial(nca, case)
        assert np.array(a.to_host(getattr(nca.DType, case["to"]))).tolist() == exp["values"]
    elif kind in ("binary", "inplace", "scalar"):
        a = nca.NCDevArray([4, 4, 4], dt)
        a.fill(case["a"])
        b = nca.NCDevArray([4, 4, 4], dt)
        b.fill(case["b"])
        rhs = case["b"] if kind == "scalar" else b
        if kind == "inplace":
            r = getattr(a, f"__{case['op']}__")(rhs)
        else:
            r = getattr(a, f"__{case['op']}__")(rhs)
        assert r.dtype == getattr(nca.DType, exp["dtype"])
        assert np.all(np.array(r) == exp["value"])
    elif kind == "scalar_chain":
        a = nca.NCDevArray([100], dt)
        a.fill(case["a"])
        for op, s in case["ops"]:
            a = getattr(a, f"__{op}__")(s)
        assert np.all(np.array(a) == exp["value"])
    elif kind in ("values_binary", "values_scalar"):
        a = nca.NCDevArray.from_host(np.array(case["a"], dtype=NP[case["dtype"]]))
        b = nca.NCDevArray.from_host(np.array(case["b"], dtype=NP[case["dtype"]])) if kind == "values_binary" \
            else case["b"]
        r = getattr(a, f"__{case['op']}__")(b)
        if "dtype" in exp:
            assert r.dtype == getattr(nca.DType, exp["dtype"])
        assert np.array(r).tolist() == exp["values"]
    elif kind == "unary":
        a = nca.NCDevArray([10], dt)
        a.fill(case["a"])
        assert np.all(np.array(-a) == exp["neg"])
        assert np.all(np.array(a + 1.0) == exp["inc"]) and np.all(np.array(a - 1.0) == exp["dec"])
    elif kind == "iota":
        a = nca.NCDevArray(case["shape"], dt)
        assert np.array(a.iota()).ravel().tolist() == exp["iota"]
        with nca.lazy_mode():
            e = a.iota() * 2 + 1
        assert np.array(nca.materialize(e)).ravel().tolist() == exp["iota_times2_plus1"]
    else:
        raise KeyError(kind)


@pytest.mark.gpu
def test_frozen_vectors_on_device(dev):
    import torch
    import ncarray_amd as nca
    from ncarray_amd import capi, tapes
    from ncarray_amd import torch_interop as ti
    nca.set_eager(True)
    a, b = nca.NCDevArray.from_host(VEC["c0_a"]), nca.NCDevArray.from_host(VEC["c0_b"])
    with nca.lazy_mode():
        e = a * b + nca.exp(-(a * a)) - b * 0.5
    out = np.array(nca.materialize(e))
    a64, b64 = VEC["c0_a"].astype(np.float64), VEC["c0_b"].astype(np.float64)
    scale = np.abs(a64 * b64) + np.exp(-a64 * a64) + 0.5 * np.abs(b64)
    assert np.all(np.abs(out - VEC["c0_out"]) <= 2 * np.finfo(np.float32).eps * scale)
    assert np.all(np.abs(out - VEC["c0_fp64"]) <= 3 * np.finfo(np.float32).eps * scale)
    assert abs(e.sum() - float(VEC["c0_sum_fp64"])) <= 1e-6 * np.abs(VEC["c0_out"]).sum()

    st = nca.Stencil(tapes.LAPLACIAN_OFFSETS, lambda v: v[0] + v[1] + v[2] + v[3] + v[4] + v[5] - 6.0 * v[6])
    x = nca.NCDevArray.from_host(VEC["lap_x"])
    lap = nca.NCDevArray(list(VEC["lap_x"].shape))
    st.apply(x, lap, nca.Boundary.zero)
    assert np.array_equal(np.array(lap), VEC["lap_out"])
    assert np.max(np.abs(np.array(lap) - VEC["lap_fp64"])) <= 8 * np.finfo(np.float32).eps * \
        np.abs(VEC["lap_x"]).max() * 6

    r = nca.NCDevArray.from_host(VEC["red_x"])
    assert np.array_equal(np.array(r.sum([0])), VEC["red_sum0"])
    assert np.array_equal(np.array(r.max([2])), VEC["red_max2"])
    assert np.array_equal(np.array(r.argmax([0])), VEC["red_argmax0"])
    assert r.sum() == float(VEC["red_sum_all"]) and r.argmax() == int(VEC["red_argmax_all"])

    dev_ = torch.device("cuda:0")
    fr = torch.from_numpy(VEC["bin_frames"]).to(dev_)
    bm = torch.from_numpy(VEC["bin_map"]).to(dev_)
    counts = torch.empty((2, 10), dtype=torch.int64, device=dev_)
    sums = torch.empty((2, 10), dtype=torch.float64, device=dev_)
    import ctypes
    capi.call("nca_hip_bin", ctypes.byref(ti.view(fr)), ctypes.byref(ti.view(bm)), 10, capi.COUNT_NONZERO,
              ctypes.byref(ti.view(counts)), ctypes.byref(ti.view(sums)), None)
    assert np.array_equal(counts.cpu().numpy(), VEC["bin_counts"])
    assert np.allclose(sums.cpu().numpy(), VEC["bin_sums"], rtol=1e-13)
    g = torch.from_numpy(VEC["bin_grad"]).to(dev_)
    gf = torch.empty_like(fr)
    capi.call("nca_hip_bin_backward", ctypes.byref(ti.view(g)), ctypes.byref(ti.view(bm)), ctypes.byref(ti.view(gf)),
              None)
    assert np.array_equal(gf.cpu().numpy().reshape(VEC["bin_grad_frames"].shape), VEC["bin_grad_frames"])
