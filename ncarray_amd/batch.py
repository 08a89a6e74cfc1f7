"""Ragged batches: one persistent launch over many uneven 1-D arrays (plus hipGraph
capture and the autotuner), over nca_hip_batch_* of the C ABI."""
import ctypes

import torch

from . import capi
from . import torch_interop as ti


class RaggedBatch:
    """Evaluates `tape` on items i = 0..n-1 with inputs[k][i] (1-D contiguous tensors of
    equal length per item), optional outputs[i], and per-item sum/max/min (`fuse`)."""

    def __init__(self, tape, inputs, outputs=None, fuse=capi.FUSE_SUM | capi.FUSE_MAX, mode=capi.ACCUM_PLAIN,
                 tile=0, vec=0, waves_per_cu=0):
        capi.init(0)
        self.tape = tape
        n_in = len(inputs)
        n = len(inputs[0]) if n_in else len(outputs)
        self.n_items = n
        self.inputs, self.outputs = inputs, outputs
        dev = (inputs[0][0] if n_in else outputs[0]).device
        sizes = [int((inputs[0][i] if n_in else outputs[i]).numel()) for i in range(n)]
        for k in range(n_in):
            for i in range(n):
                t = inputs[k][i]
                assert t.is_contiguous() and t.numel() == sizes[i] and ti.TORCH_TO_CODE[t.dtype] == tape.arr_dtypes[k]
        self.sizes = sizes
        sum_code = capi.reduce_result_dtype(capi.REDUCE_SUM, tape.dest_dtype)
        self.sums = torch.empty(n, dtype=ti.CODE_TO_TORCH[sum_code], device=dev)
        self.maxs = torch.empty(n, dtype=ti.CODE_TO_TORCH[tape.dest_dtype], device=dev)
        self.mins = torch.empty(n, dtype=ti.CODE_TO_TORCH[tape.dest_dtype], device=dev)
        ptrs = (ctypes.c_void_p * max(1, n * n_in))(*[inputs[k][i].data_ptr() for i in range(n) for k in range(n_in)])
        outs = None
        if outputs is not None:
            outs = (ctypes.c_void_p * max(1, n))(*[o.data_ptr() for o in outputs])
        sz = (ctypes.c_int64 * max(1, n))(*sizes)
        h = ctypes.c_uint64()
        capi.call("nca_hip_batch_plan", tape.handle, n, n_in, ptrs, outs, sz, fuse, mode,
                  self.sums.data_ptr() if fuse & capi.FUSE_SUM else None,
                  self.maxs.data_ptr() if fuse & capi.FUSE_MAX else None,
                  self.mins.data_ptr() if fuse & capi.FUSE_MIN else None, tile, vec, waves_per_cu, ctypes.byref(h))
        self.handle = h.value
        self._graph = None

    def config(self):
        t, v, w, nt = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        capi.call("nca_hip_batch_config", self.handle, ctypes.byref(t), ctypes.byref(v), ctypes.byref(w), ctypes.byref(nt))
        return {"tile": t.value, "vec": v.value, "waves_per_cu": w.value, "n_tiles": nt.value}

    def launch(self, scalars=(), stream=None):
        sa = capi.scalar_array(scalars, self.tape.scalar_dtypes)
        capi.call("nca_hip_batch_launch", self.handle, sa, len(scalars), stream)

    def capture(self, scalars=()):
        """Capture the batch (both launches) into an instantiated hipGraph."""
        sa = capi.scalar_array(scalars, self.tape.scalar_dtypes)
        g = ctypes.c_uint64()
        capi.call("nca_hip_batch_graph", self.handle, sa, len(scalars), ctypes.byref(g))
        self._graph = g.value
        return self._graph

    def replay(self, stream=None):
        capi.call("nca_hip_graph_launch", self._graph, stream)

    def autotune(self, scalars=(), reps=3):
        sa = capi.scalar_array(scalars, self.tape.scalar_dtypes)
        ms = ctypes.c_float()
        capi.call("nca_hip_batch_autotune", self.handle, sa, len(scalars), reps, ctypes.byref(ms))
        return ms.value

    def close(self):
        if self._graph is not None:
            capi.call("nca_hip_graph_destroy", self._graph)
            self._graph = None
        if self.handle:
            capi.call("nca_hip_batch_destroy", self.handle)
            self.handle = 0

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def ragged_sizes(n_items=4096, lo=1e3, hi=4e6, seed=0):
    """Item sizes log-uniform in [lo, hi] (seeded)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    return np.exp(rng.uniform(np.log(lo), np.log(hi), n_items)).astype(np.int64)
