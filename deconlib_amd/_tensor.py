"""Device-buffer helpers: PyTorch owns memory, the HIP library computes."""

from __future__ import annotations

import numpy as np
import torch


def device() -> torch.device:
    if not torch.cuda.is_available():
        raise RuntimeError("deconlib_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch.device("cuda", torch.cuda.current_device())


def as_f32(x) -> torch.Tensor:
    """Contiguous float32 device tensor view/copy of ``x`` (numpy, torch or scalar)."""
    if isinstance(x, torch.Tensor):
        t = x
    else:
        t = torch.from_numpy(np.ascontiguousarray(np.asarray(x, dtype=np.float32)))
    if t.dtype != torch.float32:
        t = t.to(torch.float32)
    if t.device.type != "cuda":
        t = t.to(device())
    return t.contiguous()


def empty_f32(shape) -> torch.Tensor:
    return torch.empty(tuple(int(s) for s in shape), dtype=torch.float32, device=device())


def to_numpy(x) -> np.ndarray:
    if isinstance(x, torch.Tensor):
        return x.detach().cpu().numpy()
    return np.asarray(x)
