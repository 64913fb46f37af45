"""Host <-> device helpers for the GPU parity tests."""

import numpy as np
import torch


def to_dev(a: np.ndarray) -> torch.Tensor:
    """Copies a numpy array to cuda:0, reinterpreting unsigned types as signed."""
    a = np.ascontiguousarray(a)
    view = {np.dtype(np.uint64): np.int64, np.dtype(np.uint32): np.int32}.get(a.dtype)
    if view is not None:
        a = a.view(view)
    return torch.from_numpy(a.copy()).to("cuda:0")


def to_host(t: torch.Tensor, dtype) -> np.ndarray:
    torch.cuda.synchronize()
    return t.cpu().numpy().view(dtype)
