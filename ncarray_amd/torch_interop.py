"""Zero-copy views of torch tensors for the C ABI (torch provides device memory and
streams; every computation runs in libnca_hip)."""
import torch

from . import capi

TORCH_TO_CODE = {
    torch.bool: capi.BOOL, torch.uint8: capi.UINT8, torch.int8: capi.INT8, torch.int16: capi.INT16,
    torch.int32: capi.INT32, torch.int64: capi.INT64, torch.float32: capi.FLOAT32, torch.float64: capi.FLOAT64,
    torch.bfloat16: capi.BFLOAT16, torch.float16: capi.FLOAT16,
}
for _name, _code in (("uint16", capi.UINT16), ("uint32", capi.UINT32), ("uint64", capi.UINT64)):
    if hasattr(torch, _name):
        TORCH_TO_CODE[getattr(torch, _name)] = _code
CODE_TO_TORCH = {v: k for k, v in TORCH_TO_CODE.items()}


def view(t, dtype_code=None):
    """capi.View over tensor t (any strides; element strides converted to bytes)."""
    code = TORCH_TO_CODE[t.dtype] if dtype_code is None else dtype_code
    item = t.element_size()
    v = capi.make_view(t.data_ptr(), code, list(t.shape), [s * item for s in t.stride()])
    v._owner = t  # the view keeps its tensor (and so its device memory) alive
    return v


def shape_view(shape, dtype_code):
    """A view with no data (shape and dtype only: outputs that are not materialised)."""
    return capi.make_view(None, dtype_code, list(shape))


def stream_ptr(stream=None):
    """The raw hipStream_t of a torch stream (None: the current stream)."""
    s = torch.cuda.current_stream() if stream is None else stream
    # torch's default stream is the legacy null stream: order with it without synchronising
    return s.cuda_stream or capi.STREAM_DEFAULT_ASYNC
