"""Device resolution with the reference's priority order
(Atom3D atom3d/core/device_utils.py:12-57): explicit ``device`` argument, then
the first tensor argument's device, then the default ("cuda", or "cpu" when no
GPU is visible)."""
from __future__ import annotations

from typing import Optional, Union

import torch


def resolve_device(*tensors: Optional[torch.Tensor], device: Optional[Union[str, torch.device]] = None,
                   default: str = "cuda") -> torch.device:
    if device is not None:
        return device if isinstance(device, torch.device) else torch.device(device)
    for t in tensors:
        if isinstance(t, torch.Tensor):
            return t.device
    if default == "cuda" and not torch.cuda.is_available():
        return torch.device("cpu")
    return torch.device(default)


def ensure_same_device(*tensors: torch.Tensor, target_device: Optional[torch.device] = None) -> tuple:
    if not tensors:
        raise ValueError("At least one tensor required")
    target = tensors[0].device if target_device is None else target_device
    return tuple(t if t.device == target else t.to(target) for t in tensors)


def get_default_cuda_device() -> torch.device:
    return torch.device("cuda:0") if torch.cuda.is_available() else torch.device("cpu")
