"""Learned per-column tendency correction: a 32 -> 128 -> 128 -> 32 GELU MLP on libsht's
MFMA kernels (BF16 weights, FP32 master copy and accumulation, Adam).

No counterpart in SpeedyWeather: the network maps the 32 values of one grid column
(8 levels x {u, v, T, ln ps}) to 32 tendency corrections.  Device tensors only; the
kernels run on the current torch stream.
"""
from __future__ import annotations

import ctypes as C

import torch

from ._lib import check, lib

N_IN, N_OUT = 32, 32


def nparams() -> int:
    return int(lib().sht_mlp_nparams())


def _stream(device):
    return C.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def _check_rows(x: torch.Tensor, name: str, device) -> None:
    if not x.is_cuda:
        raise ValueError(f"{name} must be a device tensor (there is no CPU path)")
    if x.dtype != torch.float32 or x.dim() != 2 or x.shape[1] != 32 or not x.is_contiguous():
        raise ValueError(f"{name} must be a contiguous float32 (B, 32) tensor")
    if x.device != device:
        raise ValueError(f"{name} is on {x.device}, the network on {device}")


class ColumnMLP:
    """Owns the device parameters, Adam moments and training scratch for batches up to
    ``max_batch`` columns."""

    def __init__(self, params, max_batch: int, device="cuda:0"):
        self.device = torch.device(device)
        if self.device.type != "cuda":
            raise ValueError("ColumnMLP needs a GPU device (there is no CPU path)")
        idx = self.device.index if self.device.index is not None else torch.cuda.current_device()
        self.device = torch.device("cuda", idx)
        self.max_batch = int(max_batch)
        h = C.c_void_p()
        check(lib().sht_mlp_create(idx, self.max_batch, C.byref(h)), "sht_mlp_create")
        self._h = h
        self._loss = torch.zeros(1, dtype=torch.float32, device=self.device)
        self.set_params(params)

    def set_params(self, params) -> None:
        p = torch.as_tensor(params, dtype=torch.float32).detach().cpu().contiguous()
        if p.numel() != nparams():
            raise ValueError(f"expected {nparams()} parameters, got {p.numel()}")
        torch.cuda.synchronize(self.device)
        check(lib().sht_mlp_set_params(self._h, C.c_void_p(p.data_ptr())), "sht_mlp_set_params")

    def params(self) -> torch.Tensor:
        out = torch.empty(nparams(), dtype=torch.float32)
        torch.cuda.synchronize(self.device)
        check(lib().sht_mlp_get_params(self._h, C.c_void_p(out.data_ptr())), "sht_mlp_get_params")
        return out

    def forward(self, x: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        _check_rows(x, "x", self.device)
        if out is None:
            out = torch.empty_like(x)
        _check_rows(out, "out", self.device)
        if out.shape[0] != x.shape[0] or x.shape[0] < 1:
            raise ValueError("x and out need the same number (>= 1) of rows")
        check(lib().sht_mlp_forward(self._h, C.c_void_p(x.data_ptr()), C.c_void_p(out.data_ptr()),
                                    int(x.shape[0]), _stream(self.device)), "sht_mlp_forward")
        return out

    __call__ = forward

    def _xt(self, x, t):
        _check_rows(x, "x", self.device)
        _check_rows(t, "t", self.device)
        if t.shape[0] != x.shape[0] or not 1 <= x.shape[0] <= self.max_batch:
            raise ValueError(f"x and t need the same number of rows, 1..{self.max_batch}")

    def gradient(self, x: torch.Tensor, t: torch.Tensor):
        """(loss tensor (1,), gradient (nparams,)) of mean((mlp(x) - t)^2); no update."""
        self._xt(x, t)
        g = torch.empty(nparams(), dtype=torch.float32, device=self.device)
        loss = torch.empty(1, dtype=torch.float32, device=self.device)
        check(lib().sht_mlp_gradient(self._h, C.c_void_p(x.data_ptr()), C.c_void_p(t.data_ptr()),
                                     int(x.shape[0]), C.c_void_p(g.data_ptr()), C.c_void_p(loss.data_ptr()),
                                     _stream(self.device)), "sht_mlp_gradient")
        return loss, g

    def train_step(self, x: torch.Tensor, t: torch.Tensor, lr: float) -> torch.Tensor:
        """One Adam step; returns the loss before the step as a device tensor (1,)."""
        self._xt(x, t)
        loss = torch.empty(1, dtype=torch.float32, device=self.device)
        check(lib().sht_mlp_train_step(self._h, C.c_void_p(x.data_ptr()), C.c_void_p(t.data_ptr()),
                                       int(x.shape[0]), C.c_float(lr), C.c_void_p(loss.data_ptr()),
                                       _stream(self.device)), "sht_mlp_train_step")
        return loss

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().sht_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
