"""Permuted (transposing), strided and indirect copies and expressions on the GPU through the C ABI.

A strided expression whose output is unit-stride along its last axis while an operand is
unit-stride along another axis runs in the tiled transpose form (64 x 64 tiles staged through
LDS).  Copies are checked bit-exactly against torch's own permuted copy; expressions are checked
bit-exactly against the same tape evaluated in the contiguous form on contiguous copies of the
operands (the per-element code is the same, only the addressing differs).  Strided and NC layouts
whose last axis holds 64 or more elements run in the row form (one wave per row)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _capi():
    from ncarray_amd import capi, tapes, torch_interop as ti
    return capi, tapes, ti


_TORCH = {torch.float32: "FLOAT32", torch.float64: "FLOAT64", torch.bfloat16: "BFLOAT16",
          torch.float16: "FLOAT16", torch.int32: "INT32", torch.uint8: "UINT8", torch.int64: "INT64"}


def _copy(x, out_dtype=None):
    capi, _, ti = _capi()
    sd = getattr(capi, _TORCH[x.dtype])
    od = getattr(capi, _TORCH[out_dtype or x.dtype])
    cd = capi.FLOAT32 if x.dtype in (torch.bfloat16, torch.float16) else sd
    out = torch.empty(x.shape, dtype=out_dtype or x.dtype, device=x.device)
    capi.Tape([capi.pack(capi.OP_LOAD_NCARR, 0)], [sd], [], cd, od).launch([ti.view(x)], ti.view(out))
    return out


def _rand(shape, dtype, dev, seed):
    g = torch.Generator(device="cpu").manual_seed(seed)
    if dtype.is_floating_point:
        return torch.randn(shape, generator=g, dtype=torch.float32).to(dtype).to(dev)
    hi = 256 if dtype == torch.uint8 else 1 << 30
    return torch.randint(0, hi, shape, generator=g, dtype=torch.int64).to(dtype).to(dev)


@pytest.mark.parametrize("shape,perm", [
    ((37, 70, 129), (0, 2, 1)),
    ((37, 70, 129), (2, 1, 0)),
    ((64, 64, 64), (1, 0, 2)),      # last axis kept: plain strided form
    ((16, 16), (1, 0)),
    ((300, 17), (1, 0)),
    ((8, 100), (1, 0)),             # an axis below 16: plain strided form
    ((3, 20, 7, 65), (0, 3, 2, 1)),
    ((5, 33, 18, 70), (3, 1, 2, 0)),
])
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64, torch.bfloat16, torch.int32, torch.uint8])
def test_permuted_copy_exact(dev, shape, perm, dtype):
    x = _rand(shape, dtype, dev, seed=len(shape) * 100 + perm[0])
    xp = x.permute(*perm)
    got = _copy(xp)
    assert torch.equal(got, xp.contiguous()), (shape, perm, dtype)


def test_permuted_cast_matches_contiguous_cast(dev):
    x = _rand((130, 67, 96), torch.float32, dev, seed=5) * 1e3
    xt = x.transpose(0, 2)
    got = _copy(xt, torch.bfloat16)
    ref = _copy(xt.contiguous(), torch.bfloat16)
    assert torch.equal(got, ref)
    got16 = _copy(xt, torch.float16)
    assert torch.equal(got16, _copy(xt.contiguous(), torch.float16))


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_transposed_expression_matches_contiguous_form(dev, dtype):
    """fused C0 tape (exp, divisions, scalars) on transposed operands; for FLOAT64 the two staged
    tiles exceed the 64 KiB LDS budget, so one operand is staged and the other read directly."""
    capi, tapes, ti = _capi()
    cd = capi.FLOAT32 if dtype == torch.float32 else capi.FLOAT64
    a = _rand((90, 45, 77), dtype, dev, seed=1).transpose(0, 2)
    b = _rand((90, 45, 77), dtype, dev, seed=2).transpose(0, 2)
    tape = capi.Tape(tapes.fused_c0(), [cd] * 5, [capi.FLOAT32], cd, cd)
    outs = []
    for aa, bb in ((a, b), (a.contiguous(), b.contiguous())):
        o = torch.empty(a.shape, dtype=dtype, device=dev)
        ops = {"a": ti.view(aa), "b": ti.view(bb)}
        tape.launch([ops[k] for k in tapes.C0_OPERANDS], ti.view(o), scalars=tapes.C0_SCALARS)
        outs.append(o)
    assert torch.equal(outs[0], outs[1])


def test_mixed_layout_operands(dev):
    """one transposed operand (staged), one contiguous operand (read directly), one broadcast."""
    capi, _, ti = _capi()
    P = capi.pack
    x = _rand((200, 130), torch.float32, dev, seed=7)
    y = _rand((130, 200), torch.float32, dev, seed=8)
    s = torch.tensor(0.25, dtype=torch.float32, device=dev)
    ins = [P(capi.OP_LOAD_NCARR, 0), P(capi.OP_LOAD_NCARR, 1), P(capi.OP_SUB), P(capi.OP_LOAD_NCARR, 2), P(capi.OP_MUL)]
    out = torch.empty((130, 200), dtype=torch.float32, device=dev)
    capi.Tape(ins, [capi.FLOAT32] * 3, [], capi.FLOAT32, capi.FLOAT32).launch(
        [ti.view(x.t()), ti.view(y), ti.view(s)], ti.view(out))
    ref = (x.t().cpu().numpy() - y.cpu().numpy()) * np.float32(0.25)
    assert np.array_equal(out.cpu().numpy(), ref)


def test_large_transpose_exact(dev):
    x = torch.arange(512 * 512 * 512, dtype=torch.int32, device=dev).view(512, 512, 512)
    for perm in ((2, 1, 0), (0, 2, 1)):
        got = _copy(x.permute(*perm))
        assert torch.equal(got, x.permute(*perm).contiguous()), perm
        del got


# ---- row form: one wave per row of the last axis (strided and indirect layouts) ----------------

@pytest.mark.parametrize("shape,sl", [
    ((40, 33, 300), (slice(None), slice(None), slice(None, None, 2))),
    ((7, 5, 3, 130), (slice(None), slice(1, 4), slice(None), slice(None, 129))),
    ((65, 191), (slice(None, None, 3), slice(None))),
    ((9, 64), (slice(None, None, 2), slice(None))),
])
@pytest.mark.parametrize("dtype", [torch.float32, torch.bfloat16, torch.int64, torch.uint8])
def test_strided_rows_exact(dev, shape, sl, dtype):
    x = _rand(shape, dtype, dev, seed=len(shape))
    xs = x[sl]
    assert torch.equal(_copy(xs), xs.contiguous())


def test_nc_rows_read_write_exact(dev):
    capi, _, ti = _capi()
    P = capi.pack
    ident = capi.Tape([P(capi.OP_LOAD_NCARR, 0)], [capi.FLOAT32], [], capi.FLOAT32, capi.FLOAT32)
    x = _rand((37, 20, 130), torch.float32, dev, seed=11)
    tbl = torch.tensor([x[i].data_ptr() for i in range(37)], dtype=torch.int64, device=dev)
    vx = capi.make_nc_view(tbl.data_ptr(), capi.FLOAT32, (37, 20, 130), [1, 130 * 4, 4], 0)
    out = torch.empty_like(x)
    ident.launch([vx], ti.view(out))
    assert torch.equal(out, x)
    # row_offset: every row read from its third element on
    vo8 = capi.make_nc_view(tbl.data_ptr(), capi.FLOAT32, (37, 20, 128), [1, 130 * 4, 4], 0, row_offset=8)
    out8 = torch.empty((37, 20, 128), dtype=torch.float32, device=dev)
    ident.launch([vo8], ti.view(out8))
    assert torch.equal(out8, x[:, :, 2:])
    # an NC output: planes of y written through their table, from a transposed operand and a
    # plain one
    y = torch.zeros_like(x)
    tbly = torch.tensor([y[i].data_ptr() for i in range(37)], dtype=torch.int64, device=dev)
    vy = capi.make_nc_view(tbly.data_ptr(), capi.FLOAT32, (37, 20, 130), [1, 130 * 4, 4], 0)
    z = _rand((130, 20, 37), torch.float32, dev, seed=12)
    add = capi.Tape([P(capi.OP_LOAD_NCARR, 0), P(capi.OP_LOAD_NCARR, 1), P(capi.OP_ADD)],
                    [capi.FLOAT32] * 2, [], capi.FLOAT32, capi.FLOAT32)
    add.launch([ti.view(x), ti.view(z.permute(2, 1, 0))], vy)
    assert torch.equal(y, x + z.permute(2, 1, 0))


def test_rows_linear_index(dev):
    """IDX is the C-order linear index of the output element in the row form too."""
    capi, _, ti = _capi()
    P = capi.pack
    ins = [P(capi.OP_IDX), P(capi.OP_LOAD_CONST, 0), P(capi.OP_MUL), P(capi.OP_LOAD_CONST, 1), P(capi.OP_ADD)]
    big = torch.zeros((3, 10, 200), dtype=torch.int64, device=dev)
    o = big[:, ::2, :]
    capi.Tape(ins, [], [capi.INT64, capi.INT64], capi.INT64, capi.INT64).launch([], ti.view(o), scalars=(2, 1))
    assert torch.equal(o.cpu(), (torch.arange(3 * 5 * 200) * 2 + 1).view(3, 5, 200))
    assert (big[:, 1::2, :] == 0).all()


@pytest.mark.parametrize("dtype", [torch.bfloat16, torch.float16])
def test_two_byte_transpose_pairs(dev, dtype):
    """2-byte data with 4-byte aligned strides moves as 4-byte pairs in the tiled form; odd extents
    along the staged axis (97) and the stored axis (67) take the single-element tails, and a base
    one element off 4-byte alignment takes the unpaired form."""
    base = _rand((66, 130, 98), dtype, dev, seed=21)
    for x in (base, base[..., :97], base[:67]):
        for perm in ((2, 1, 0), (0, 2, 1)):
            xp = x.permute(*perm)
            assert torch.equal(_copy(xp), xp.contiguous()), (tuple(x.shape), perm)
    flat = _rand((1 + 66 * 130 * 98,), dtype, dev, seed=22)
    off = flat[1:].view(66, 130, 98)
    assert off.data_ptr() % 4 == 2
    assert torch.equal(_copy(off.transpose(0, 2)), off.transpose(0, 2).contiguous())
    # 4-byte source cast to a 2-byte output: the staged operand is not 2-byte, unpaired form
    x32 = _rand((66, 130, 98), torch.float32, dev, seed=23).transpose(0, 2)
    assert torch.equal(_copy(x32, dtype), _copy(x32.contiguous(), dtype))
