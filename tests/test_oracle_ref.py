"""The oracle pinned against the reference's own host code (no GPU).

oracle/_ref/libnca_ref.so wraps the reference's header-only host routines, compiled from the
reference sources where they lie (oracle/Makefile target `ref`): host::sum/max/min/argmax/
argmin_recursive (host/reductions.hh:35-212), host::scatter_add_recursive
(host/elementwise.hh:482-505), the expression VM DynamicExprMVNode::eval
(expression/dynamicmvnode.hh:146-251) and the layouts' advance (layout.hh:451-480, 677-700).
Each test runs the same seeded inputs through the oracle's restatement and through the
reference's code.  The module is skipped where that library is not built (no reference tree).
"""
import numpy as np
import pytest

from oracle import oracle as O
from oracle import ref as R

pytestmark = pytest.mark.skipif(R.lib is None, reason="oracle/_ref not built (reference sources absent)")

P = lambda op, i=0: ((op & 0xFF) << 24) | (i & 0xFFFFFF)  # noqa: E731  (op_code.hh:174-192)
IDX, LOAD_NC, LOAD_SO, CONST = 1, 2, 3, 4
UNARY = (5, 6, 7)  # NEG INC DEC
BINARY = (14, 15, 16, 17, 18, 20, 21, 22, 23, 24, 25, 26, 27)  # ADD..MOD, EQ..GE, LAND, LOR


def same_f(a, b):
    """Equal values (zeros of either sign equal), NaN where the other is NaN."""
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return bool(np.all((a == b) | (np.isnan(a) & np.isnan(b))))


def same_bits_f32(a, b):
    """Bit-identical float32 arrays, any NaN matching any NaN."""
    a, b = np.asarray(a, np.float32), np.asarray(b, np.float32)
    nan = np.isnan(a) & np.isnan(b)
    return bool(np.all(nan | (a.view(np.uint32) == b.view(np.uint32))))


def host_image(nbytes):
    """A ByteImage over host memory: pointers written into it are real host addresses."""
    img = O.ByteImage(nbytes, 0)
    img.addr = img.buf.ctypes.data
    return img


def special_floats(rng, n, dtype):
    x = rng.standard_normal(n).astype(dtype)
    pick = rng.random(n)
    x[pick < 0.05] = np.nan
    x[(pick >= 0.05) & (pick < 0.08)] = np.inf
    x[(pick >= 0.08) & (pick < 0.11)] = -np.inf
    x[(pick >= 0.11) & (pick < 0.16)] = 0.0
    x[(pick >= 0.16) & (pick < 0.21)] = -0.0
    x[(pick >= 0.21) & (pick < 0.30)] = x[0] if n else 0  # ties
    return x


# ---- full reductions over contiguous inputs ----------------------------------------------------------
@pytest.mark.parametrize("dtype", [np.float32, np.float64, np.int32, np.int64, np.int16, np.uint8, np.bool_])
@pytest.mark.parametrize("n", [0, 1, 2, 7, 1000])
@pytest.mark.parametrize("special", [False, True])
def test_full_reductions_match_reference(dtype, n, special):
    rng = np.random.default_rng(n * 31 + int(special))
    if np.dtype(dtype).kind == "f":
        x = special_floats(rng, n, dtype) if special else rng.standard_normal(n).astype(dtype)
    elif dtype == np.bool_:
        x = rng.random(n) < (0.9 if special else 0.5)
    else:
        info = np.iinfo(dtype)
        lo, hi = (info.min, info.max) if special else (-50 if info.min < 0 else 0, 50)
        x = rng.integers(lo, hi, n, dtype=dtype, endpoint=True)
    x = np.ascontiguousarray(x)
    code = O.CODE[x.dtype]
    it = x.dtype.itemsize
    ref = {op: R.reduce(op, code, x.ctypes.data, [n], [it]) for op in (R.SUM, R.MAX, R.MIN, R.ARGMAX, R.ARGMIN)}

    s = O.sum_fold(x)
    if x.dtype.kind == "f":
        assert same_bits_f32(s, ref[R.SUM][0]) if dtype == np.float32 else same_f(s, ref[R.SUM][0])
    else:  # the sum in sum_type (oracle/ref_driver.cc hands it over as int64 bits)
        st = np.iinfo(O.SUM_TYPE[code])
        assert st.min <= s <= st.max and (s - ref[R.SUM][1]) % 2 ** 64 == 0
    mx, imx = O.max_fold(x)
    mn, imn = O.min_fold(x)
    # index rules are the reference's exactly; values equal up to the sign of a zero extremum
    # (the fold keeps the first of -0/+0, the oracle orders -0 below +0: order-independent)
    assert imx == ref[R.ARGMAX][1] and imn == ref[R.ARGMIN][1]
    rmx, rmn = ref[R.MAX], ref[R.MIN]
    if x.dtype.kind == "f":
        assert same_f(mx, rmx[0]) and same_f(mn, rmn[0])
        zeros = x[x == 0]
        if not (np.any(np.signbit(zeros)) and np.any(~np.signbit(zeros))):
            assert np.signbit(mx) == np.signbit(rmx[0]) and np.signbit(mn) == np.signbit(rmn[0])
    else:
        assert mx == rmx[1] or mx == rmx[0]
        assert mn == rmn[1] or mn == rmn[0]


def test_signed_zero_extremum_rule():
    # the one place the oracle's value departs from the host fold: both zeros at the extremum
    for seq in ([-0.0, 0.0], [0.0, -0.0], [-1.0, -0.0, 0.0, -0.0]):
        x = np.array(seq, np.float32)
        rv, _ = R.reduce(R.MAX, 10, x.ctypes.data, [x.size], [4])
        ov, oi = O.max_fold(x)
        first = x[int(np.argmax(x == 0))]
        assert np.signbit(rv) == np.signbit(first)  # the fold keeps the first zero
        assert not np.signbit(ov)  # the oracle's max is +0 whatever the order
        assert oi == R.reduce(R.ARGMAX, 10, x.ctypes.data, [x.size], [4])[1]


# ---- views: element addresses and reductions over them -----------------------------------------------
def nc_pointer_view(rng, img, shape, pax, offsets, dtype=np.float32):
    """A reference NC view with pointer axis `pax` over rows placed at random in `img`.

    The reference view: data = table, strides (bytes; ignored at pax), offsets (bytes; entries
    at pax).  Rows are C-contiguous over the axes after pax; axes before pax step through
    copies of the table."""
    it = np.dtype(dtype).itemsize
    inner = tuple(shape[pax + 1:])
    ninner = int(np.prod(inner)) if inner else 1
    extra = offsets[pax + 1:]
    row_bytes = ninner * it + sum(int(o) for o in extra) + 64
    nrows = shape[pax] + offsets[pax]
    outer = int(np.prod(shape[:pax])) if pax else 1
    tables = []
    for _ in range(outer):
        rows = [img.put(rng.standard_normal(row_bytes // it).astype(dtype), align=it) for _ in range(nrows)]
        rows = [rows[k] for k in rng.permutation(nrows)]
        tables.append(img.put_ptrs(rows))
    # tables are laid out one after another with a constant gap: the outer axes' stride
    tstride = (tables[1] - tables[0]) if outer > 1 else 0
    assert all(t == tables[0] + k * tstride for k, t in enumerate(tables))
    strides = []
    for d in range(len(shape)):
        if d < pax:
            strides.append(tstride * int(np.prod(shape[d + 1:pax])))
        elif d == pax:
            strides.append(0)
        else:
            strides.append(it * int(np.prod(shape[d + 1:])))
    return img.addr + tables[0], strides


@pytest.mark.parametrize("shape,pax,offsets", [
    ((5, 9), 0, [0, 0]),
    ((5, 9), 0, [2, 12]),
    ((3, 4, 6), 1, [0, 1, 8]),
    ((2, 3, 4, 5), 2, [0, 0, 3, 4]),
])
def test_nc_pointer_axis_addresses_and_reductions(shape, pax, offsets):
    rng = np.random.default_rng(len(shape) * 7 + pax)
    img = host_image(1 << 20)
    data, strides = nc_pointer_view(rng, img, list(shape), pax, offsets)
    ref_off = R.offsets(data, img.addr, shape, strides, offsets, R.NC, pax)
    # the device view form: offsets of the axes before pax folded into the base, offsets[pax] the
    # table entry offset, those after pax the row offset; strides[pax] = 1 entry
    base = data + sum(offsets[:pax])
    dstrides = list(strides)
    dstrides[pax] = 1
    mine = O.nc_offsets(img.buf, img.addr, base, shape, dstrides, pax, offsets[pax], sum(offsets[pax + 1:]))
    assert np.array_equal(mine, ref_off)
    vals = img.buf[mine[..., None] + np.arange(4)].copy().view(np.float32)[..., 0]
    for op, fn in ((R.SUM, None), (R.MAX, O.max_fold), (R.ARGMIN, O.min_fold)):
        rv, ri = R.reduce(op, 10, data, shape, strides, offsets, R.NC, pax)
        if op == R.SUM:
            assert same_bits_f32(O.sum_fold(vals), rv)
        elif op == R.MAX:
            assert same_f(fn(vals)[0], rv)
        else:
            assert fn(vals)[1] == ri


@pytest.mark.parametrize("shape,strides_sign", [((4, 6), 1), ((3, 5, 7), -1)])
def test_nc_strided_offsets_fold_into_base(shape, strides_sign):
    # no pointer axis: every axis' byte offset is added at each advance, i.e. once into the base
    rng = np.random.default_rng(3)
    x = rng.standard_normal(4096).astype(np.float32)
    it = 4
    nd = len(shape)
    strides = [strides_sign * it * (2 + d) * int(np.prod(shape[d + 1:])) for d in range(nd)]
    offsets = [4 * (d + 1) for d in range(nd)]
    start = 2048 * it
    data = x.ctypes.data + start
    ref_off = R.offsets(data, x.ctypes.data, shape, strides, offsets, R.NC, -1)
    mine = O.nc_offsets(x.view(np.uint8), x.ctypes.data, data + sum(offsets), shape, strides)
    assert np.array_equal(mine, ref_off)
    vals = x.view(np.uint8)[mine[..., None] + np.arange(4)].copy().view(np.float32)[..., 0]
    assert same_bits_f32(O.sum_fold(vals), R.reduce(R.SUM, 10, data, shape, strides, offsets)[0])


@pytest.mark.parametrize("case", range(4))
def test_so_addresses_and_reductions(case):
    rng = np.random.default_rng(100 + case)
    img = host_image(1 << 20)
    if case == 0:  # one indirect axis (a table of rows), suboffset into each row
        rows = [img.put(rng.standard_normal(40).astype(np.float32)) for _ in range(6)]
        base = img.put_ptrs(rows)
        shape, strides, subs = (6, 9), (8, 4), (12, -1)
    elif case == 1:  # indirect at the last axis too: a table of element pointers per row
        elems = [img.put(rng.standard_normal(3).astype(np.float32)) for _ in range(20)]
        rowt = [img.put_ptrs(elems[k * 5:(k + 1) * 5]) for k in range(4)]
        base = img.put_ptrs(rowt)
        shape, strides, subs = (4, 5), (8, 8), (0, 4)
    elif case == 2:  # direct, reversed, then indirect
        rows = [img.put(rng.standard_normal(16).astype(np.float32)) for _ in range(12)]
        tbl = img.put_ptrs(rows)
        base = tbl + 11 * 8
        shape, strides, subs = (3, 4, 5), (-32, -8, 4), (-1, 0, -1)
    else:  # stepped pointer table with a suboffset, stepped inner axis
        rows = [img.put(rng.standard_normal(64).astype(np.float32)) for _ in range(10)]
        base = img.put_ptrs(rows)
        shape, strides, subs = (5, 7), (16, 8), (8, -1)
    addr = img.addr + base
    ref_off = R.offsets(addr, img.addr, shape, strides, subs, R.SO)
    mine = O.so_offsets(img.buf, img.addr, addr, shape, strides, subs)
    assert np.array_equal(mine, ref_off)
    vals = O.so_gather(img.buf, img.addr, addr, np.float32, shape, strides, subs)
    assert same_bits_f32(O.sum_fold(vals), R.reduce(R.SUM, 10, addr, shape, strides, subs, R.SO)[0])
    assert O.max_fold(vals)[1] == R.reduce(R.ARGMAX, 10, addr, shape, strides, subs, R.SO)[1]


# ---- scatter_add ---------------------------------------------------------------------------------------
@pytest.mark.parametrize("n,nd", [(0, 5), (1, 1), (1000, 37), (5000, 3)])
def test_scatter_add_matches_reference(n, nd):
    rng = np.random.default_rng(n + nd)
    dest = rng.standard_normal(nd).astype(np.float32)
    idx = rng.integers(-3, nd + 3, n).astype(np.int32)  # out-of-range indices are skipped
    src = rng.standard_normal(n).astype(np.float32)
    assert same_bits_f32(O.scatter_add_f32(dest, idx, src), R.scatter_add_f32(dest, idx, src))


# ---- the expression VM ------------------------------------------------------------------------------------
def random_tape(rng, n_arrays, n_scalars, depth, load):
    if depth == 0 or rng.random() < 0.25:
        r = rng.random()
        if r < 0.6:
            return [P(load, int(rng.integers(n_arrays)))]
        if r < 0.9:
            return [P(CONST, int(rng.integers(n_scalars)))]
        return [P(IDX)]
    if rng.random() < 0.2:
        return random_tape(rng, n_arrays, n_scalars, depth - 1, load) + [P(int(rng.choice(UNARY)))]
    return (random_tape(rng, n_arrays, n_scalars, depth - 1, load)
            + random_tape(rng, n_arrays, n_scalars, depth - 1, load) + [P(int(rng.choice(BINARY)))])


@pytest.mark.parametrize("seed", range(40))
def test_vm_matches_reference(seed):
    rng = np.random.default_rng(seed)
    n = 257
    arrays = [special_floats(rng, n, np.float32) for _ in range(3)]
    arrays[1][::7] = 0.0
    scalars = [float(np.float32(v)) for v in (rng.standard_normal(), 0.0, -2.5)]
    load = LOAD_SO if seed % 2 else LOAD_NC
    tape = random_tape(rng, 3, 3, 4, load)
    mine = O.expr(tape, arrays, scalars, 10, np.float32)
    ref = R.vm_f32(tape, arrays, scalars, so_loads=(load == LOAD_SO))
    assert same_bits_f32(mine, ref), [hex(t) for t in tape]


def test_vm_division_rule_matches_reference():
    # DIV by zero (interface.hh:150-156): NaN for a finite numerator, the numerator otherwise
    a = np.array([1.0, -1.0, 0.0, np.inf, -np.inf, np.nan, 3.0, -0.0], np.float32)
    b = np.array([0.0, -0.0, 0.0, 0.0, -0.0, 0.0, 2.0, 0.0], np.float32)
    tape = [P(LOAD_NC, 0), P(LOAD_NC, 1), P(17)]
    assert same_bits_f32(O.expr(tape, [a, b], [], 10, np.float32), R.vm_f32(tape, [a, b], []))
