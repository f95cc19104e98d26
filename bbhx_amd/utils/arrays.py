"""Small helpers to move numpy / scalar inputs onto the backend's device."""

from __future__ import annotations

import numpy as np


def as_np1d(x, dtype=np.float64):
    """numpy 1D copy of a scalar, list, numpy array or torch tensor."""
    try:
        import torch

        if isinstance(x, torch.Tensor):
            x = x.detach().cpu().numpy()
    except ImportError:  # pragma: no cover
        pass
    return np.array(np.atleast_1d(x), dtype=dtype, copy=True)


def to_dev(x, device, dtype=None):
    """Contiguous torch tensor on ``device`` (float64 unless ``dtype``)."""
    import torch

    dtype = torch.float64 if dtype is None else dtype
    if isinstance(x, torch.Tensor):
        return x.to(device=device, dtype=dtype).contiguous()
    return torch.as_tensor(np.ascontiguousarray(x), dtype=dtype, device=device).contiguous()


def to_host(x):
    """numpy view/copy of a torch tensor (or pass-through)."""
    try:
        import torch

        if isinstance(x, torch.Tensor):
            return x.detach().cpu().numpy()
    except ImportError:  # pragma: no cover
        pass
    return x


def np_logspace_rows(start, stop, num):
    """numpy.logspace(start, stop, num, axis=-1) for array endpoints, in torch
    (base 10): start + arange(num) * step with the last point set to stop."""
    import torch

    step = (stop - start) / (num - 1)
    ar = torch.arange(num, dtype=torch.float64, device=start.device)
    y = start[..., None] + ar * step[..., None]
    y[..., -1] = stop
    return torch.pow(10.0, y)
