"""Learned warm start of SKM registrations (no reference counterpart: the
reference starts every registration from zero coefficients,
src/SKMReg.cxx:375-430).

WarmStartNet wraps the library's network (skm_ws_*: a 3D conv
encoder-decoder with BF16 weights and FP32 accumulation, trained and run by
libskm.so's own kernels).  Device arrays are torch tensors on the GPU (torch
is used for device memory and streams only; every computation is a library
kernel).
"""
import ctypes

import numpy as np

from . import _capi
from ._capi import check


class WarmStartNet:
    """in_size (x, y, z) of the network input (multiples of 8); control points
    in input voxel coordinates (cp_size, cp_origin, cp_spacing, each x, y, z)."""

    def __init__(self, in_size, cp_size, cp_origin, cp_spacing, max_batch=1):
        import torch

        self._torch = torch
        self._lib = _capi.load()
        d = _capi.SkmWsDesc()
        for a in range(3):
            d.in_size[a] = int(in_size[a])
            d.cp_size[a] = int(cp_size[a])
            d.cp_origin[a] = float(cp_origin[a])
            d.cp_spacing[a] = float(cp_spacing[a])
        d.max_batch = int(max_batch)
        h = ctypes.c_void_p()
        check(self._lib.skm_ws_create(ctypes.byref(d), ctypes.byref(h)))
        self._h = h
        self.in_size = tuple(int(v) for v in in_size)
        self.cp_size = tuple(int(v) for v in cp_size)
        self.max_batch = int(max_batch)
        self.n_params = int(self._lib.skm_ws_num_params(h))
        check(self._lib.skm_ws_set_stream(h, ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))

    def close(self):
        if self._h:
            self._lib.skm_ws_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001 -- interpreter shutdown
            pass

    # -- parameters (host float32, the library's layout)
    def set_params(self, params):
        p = np.ascontiguousarray(params, np.float32)
        if p.size != self.n_params:
            raise ValueError(f"expected {self.n_params} parameters, got {p.size}")
        check(self._lib.skm_ws_set_params(self._h, _capi.fptr(p)))

    def get_params(self):
        p = np.empty(self.n_params, np.float32)
        check(self._lib.skm_ws_get_params(self._h, _capi.fptr(p)))
        return p

    def save(self, path):
        self.get_params().tofile(path)

    def load(self, path):
        self.set_params(np.fromfile(path, np.float32))

    # -- device calls (torch cuda float32 tensors)
    def _shape_in(self, batch):
        return (batch, self.in_size[2], self.in_size[1], self.in_size[0], 2)

    def _shape_out(self, batch):
        return (batch, self.cp_size[2], self.cp_size[1], self.cp_size[0], 3)

    def _check(self, t, shape):
        if tuple(t.shape) != tuple(shape) or t.dtype != self._torch.float32 or not t.is_cuda or not t.is_contiguous():
            raise ValueError(f"expected a contiguous float32 cuda tensor of shape {shape}, got {tuple(t.shape)}")

    def forward(self, x):
        batch = x.shape[0]
        self._check(x, self._shape_in(batch))
        out = self._torch.empty(self._shape_out(batch), device=x.device, dtype=self._torch.float32)
        check(self._lib.skm_ws_forward(self._h, ctypes.c_void_p(x.data_ptr()), batch, ctypes.c_void_p(out.data_ptr())))
        return out

    def backward(self, dout):
        """Parameter gradients for the output gradient of the last forward."""
        grad = self._torch.empty(self.n_params, device=dout.device, dtype=self._torch.float32)
        check(self._lib.skm_ws_backward(self._h, ctypes.c_void_p(dout.data_ptr()), ctypes.c_void_p(grad.data_ptr())))
        return grad

    LAYER_CHANNELS = (16, 16, 32, 32, 48, 48, 64, 64, 48, 32)
    LAYER_LEVELS = (0, 0, 1, 1, 2, 2, 3, 3, 2, 1)

    def activation(self, layer, batch):
        """Output of conv `layer` in the last forward (torch tensor)."""
        lev = self.LAYER_LEVELS[layer]
        shape = (batch, self.in_size[2] >> lev, self.in_size[1] >> lev, self.in_size[0] >> lev,
                 self.LAYER_CHANNELS[layer])
        out = self._torch.empty(shape, device="cuda", dtype=self._torch.float32)
        check(self._lib.skm_ws_get_activation(self._h, int(layer), ctypes.c_void_p(out.data_ptr())))
        return out

    def train_step(self, x, target, learning_rate):
        batch = x.shape[0]
        self._check(x, self._shape_in(batch))
        self._check(target, self._shape_out(batch))
        loss = ctypes.c_double()
        check(self._lib.skm_ws_train_step(self._h, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(target.data_ptr()),
                                          batch, float(learning_rate), ctypes.byref(loss)))
        return loss.value
