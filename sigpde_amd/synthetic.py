"""Seeded synthetic inputs: Brownian paths x[:, 0] = 0, x[:, 1:] = cumsum of i.i.d. N(0, var)
increments, generated in float64 on the CPU and then cast (so every dtype sees the same draw)."""
from __future__ import annotations

import torch


def brownian_paths(batch: int, length: int, dim: int, var: float, seed: int,
                   dtype: torch.dtype = torch.float32, device="cpu") -> torch.Tensor:
    g = torch.Generator().manual_seed(int(seed))
    inc = torch.randn((batch, length - 1, dim), generator=g, dtype=torch.float64) * (var ** 0.5)
    x = torch.zeros((batch, length, dim), dtype=torch.float64)
    x[:, 1:, :] = torch.cumsum(inc, dim=1)
    return x.to(dtype=dtype, device=device)
