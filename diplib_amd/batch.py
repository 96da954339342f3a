"""Batched uneven form: many 2-D images of different sizes and per-image sigmas
through one persistent kernel (C ABI: dipamd_batch_*).

    plan = GaussBatch(inputs, outputs, sigmas)   # builds the work queue, autotunes the tile
    plan.run()          # one persistent launch on the current stream
    plan.run_graph()    # the same, replayed from a hipGraph captured on first use
    plan.run_per_image()  # the per-image loop of the plain drop-in (the baseline)

Each output equals dip::Gauss(in, {sx, sy}, {0}, "FIR", bc, truncation) bit for bit
in the default exact mode.  Inputs are 2-D device tensors (uint8, float32 or float64)
with unit x stride; outputs are preallocated float32 (float64 for float64 input).
"""
import ctypes

import torch

from . import _lib
from ._lib import DipamdBatchImage, check, lib
from .gauss import _bc_codes, _stream_handle

_IN = {torch.uint8: _lib.DT_UINT8, torch.float32: _lib.DT_SFLOAT, torch.float64: _lib.DT_DFLOAT}


def _descriptors(inputs, outputs, sigmas):
    """(ctypes array of dipamd_batch_image, input dtype, output dtype)."""
    if len(inputs) != len(outputs) or len(inputs) != len(sigmas) or not inputs:
        raise ValueError("inputs, outputs and sigmas must be non-empty lists of equal length")
    dt = inputs[0].dtype
    if dt not in _IN:
        raise TypeError(f"batched form takes uint8, float32 or float64 input, not {dt}")
    odt = torch.float64 if dt == torch.float64 else torch.float32
    arr = (DipamdBatchImage * len(inputs))()
    for i, (a, b, s) in enumerate(zip(inputs, outputs, sigmas)):
        if a.dim() != 2 or a.dtype != dt or a.stride(1) != 1 or not a.is_cuda:
            raise ValueError("batch inputs must be 2-D device tensors of one dtype with unit x stride")
        if b.shape != a.shape or b.dtype != odt or b.stride(1) != 1:
            raise ValueError("batch outputs must match the inputs' shapes, dtype " + str(odt))
        sx, sy = (s, s) if not hasattr(s, "__len__") else (s[0], s[1])
        arr[i].inp = a.data_ptr()
        arr[i].out = b.data_ptr()
        arr[i].width = a.shape[1]
        arr[i].height = a.shape[0]
        arr[i].in_stride_y = a.stride(0)
        arr[i].out_stride_y = b.stride(0)
        arr[i].sigma_x = float(sx)
        arr[i].sigma_y = float(sy)
        arr[i].order_x = 0
        arr[i].order_y = 0
    return arr, dt, odt


def gauss_batch_noalloc(inputs, outputs, sigmas, boundary="mirror", truncation=3.0, fast=False, tile_config=-1):
    """The batched form in one call that allocates no device memory
    (dipamd_batch_gauss_noalloc): the work queue and descriptors go to the library's
    fixed device array.  Same results as GaussBatch(...).run(); calls that overlap on
    one GPU race."""
    arr, dt, odt = _descriptors(inputs, outputs, sigmas)
    bcs = _bc_codes(boundary)
    b_arr = (ctypes.c_int32 * max(1, len(bcs)))(*bcs)
    dev = inputs[0].device
    with torch.cuda.device(dev):
        check(lib().dipamd_batch_gauss_noalloc(arr, len(inputs), _IN[dt],
                                               _lib.DT_DFLOAT if odt == torch.float64 else _lib.DT_SFLOAT, b_arr,
                                               len(bcs), float(truncation), _lib.FAST if fast else _lib.EXACT,
                                               int(tile_config), _stream_handle(dev)))


class GaussBatch:
    def __init__(self, inputs, outputs, sigmas, boundary="mirror", truncation=3.0, fast=False, tile_config=-1):
        arr, dt, odt = _descriptors(inputs, outputs, sigmas)
        self._keep = (list(inputs), list(outputs))
        self.device = inputs[0].device
        bcs = _bc_codes(boundary)
        b_arr = (ctypes.c_int32 * max(1, len(bcs)))(*bcs)
        self._plan = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            check(lib().dipamd_batch_create(ctypes.byref(self._plan), arr, len(inputs), _IN[dt],
                                            _lib.DT_DFLOAT if odt == torch.float64 else _lib.DT_SFLOAT, b_arr,
                                            len(bcs), float(truncation), _lib.FAST if fast else _lib.EXACT,
                                            int(tile_config), _stream_handle(self.device)))
        nt = ctypes.c_int64()
        cfg = ctypes.c_int32()
        check(lib().dipamd_batch_info(self._plan, ctypes.byref(nt), ctypes.byref(cfg)))
        self.ntiles = nt.value
        self.config = cfg.value

    def run(self):
        with torch.cuda.device(self.device):
            check(lib().dipamd_batch_run(self._plan, _stream_handle(self.device)))

    def run_graph(self):
        with torch.cuda.device(self.device):
            check(lib().dipamd_batch_run_graph(self._plan, _stream_handle(self.device)))

    def run_per_image(self):
        with torch.cuda.device(self.device):
            check(lib().dipamd_batch_run_per_image(self._plan, _stream_handle(self.device)))

    def close(self):
        if self._plan:
            lib().dipamd_batch_destroy(self._plan)
            self._plan = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
