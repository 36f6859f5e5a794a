"""Python handle of the warm-start U-Net (pa_unet_* in include/plato_mi355x.h).

All compute runs in the library's HIP kernels (BF16 MFMA forward and backward,
FP32 Adam); tensors are torch CUDA float32, channels last: input [batch, nx, ny, nz,
IN_CH], output [batch, nx, ny, nz, OUT_CH].
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _capi

IN_CH = 8
OUT_CH = 16

# (cin, cout, input level, output level) of the 15 convolutions in launch order, as
# pa_unet.hip's build_layers defines them for base width C (level l = grid halved l times,
# rounded up; stride-2 layers go down a level, transposed ones up).
def conv_layers(C: int) -> list[tuple[int, int, int, int]]:
    return [(IN_CH, C, 0, 0), (C, C, 0, 0), (C, 2 * C, 0, 1), (2 * C, 2 * C, 1, 1), (2 * C, 4 * C, 1, 2),
            (4 * C, 4 * C, 2, 2), (4 * C, 8 * C, 2, 3), (8 * C, 8 * C, 3, 3), (8 * C, 4 * C, 3, 2),
            (8 * C, 4 * C, 2, 2), (4 * C, 2 * C, 2, 1), (4 * C, 2 * C, 1, 1), (2 * C, C, 1, 0),
            (2 * C, C, 0, 0), (C, OUT_CH, 0, 0)]


def _p(t: torch.Tensor):
    if not t.is_cuda or t.dtype != torch.float32 or not t.is_contiguous():
        raise ValueError("expected a contiguous float32 CUDA tensor")
    return C.c_void_p(t.data_ptr())


class UNet:
    def __init__(self, grid: tuple[int, int, int], base_channels: int = 16, seed: int = 0):
        self.grid = tuple(int(n) for n in grid)
        self.base = int(base_channels)
        h = C.c_void_p()
        _capi.ucall("pa_unet_create", C.byref(h), *self.grid, self.base, C.c_ulonglong(seed))
        self.h = h
        n = C.c_size_t()
        _capi.ucall("pa_unet_num_params", self.h, C.byref(n))
        self.num_params = int(n.value)
        self.voxels = self.grid[0] * self.grid[1] * self.grid[2]

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            self.h = None
            try:
                _capi.lib().pa_unet_destroy(h)
            except Exception:  # interpreter shutdown: the library may already be gone
                pass

    def _stream(self):
        _capi.ucall("pa_unet_set_stream", self.h, C.c_void_p(torch.cuda.current_stream().cuda_stream))

    def _check(self, x: torch.Tensor) -> int:
        if x.shape[1:] != (*self.grid, IN_CH):
            raise ValueError(f"input must be [batch, {self.grid}, {IN_CH}], got {tuple(x.shape)}")
        return int(x.shape[0])

    def forward_flops(self, batch: int) -> int:
        """FLOPs of one forward pass as the implicit-GEMM convolutions compute them:
        2 x (batch x output voxels) x cout x 27 cin per layer (all 27 taps, padding included)."""
        g = [self.grid]
        for _ in range(3):
            g.append(tuple((n + 1) // 2 for n in g[-1]))
        vox = [a * b * c for a, b, c in g]
        return sum(2 * batch * vox[lo] * co * 27 * ci for ci, co, _, lo in conv_layers(self.base))

    def params(self) -> torch.Tensor:
        out = torch.empty(self.num_params, dtype=torch.float32, device="cuda")
        self._stream()
        _capi.ucall("pa_unet_params", self.h, _p(out), None)
        return out

    def set_params(self, p: torch.Tensor) -> None:
        p = p.detach().to(device="cuda", dtype=torch.float32).contiguous()
        assert p.numel() == self.num_params
        self._stream()
        _capi.ucall("pa_unet_params", self.h, None, _p(p))

    def forward(self, x: torch.Tensor) -> torch.Tensor:
        b = self._check(x)
        y = torch.empty((b, *self.grid, OUT_CH), dtype=torch.float32, device=x.device)
        self._stream()
        _capi.ucall("pa_unet_forward", self.h, _p(x), b, _p(y))
        return y

    def gradient(self, x: torch.Tensor, target: torch.Tensor) -> tuple[float, torch.Tensor]:
        b = self._check(x)
        g = torch.empty(self.num_params, dtype=torch.float32, device=x.device)
        loss = C.c_double()
        self._stream()
        _capi.ucall("pa_unet_gradient", self.h, _p(x), _p(target), b, _p(g), C.byref(loss))
        return loss.value, g

    def train_step(self, x: torch.Tensor, target: torch.Tensor, lr: float) -> float:
        b = self._check(x)
        loss = C.c_double()
        self._stream()
        _capi.ucall("pa_unet_train_step", self.h, _p(x), _p(target), b, C.c_float(lr), C.byref(loss))
        return loss.value
