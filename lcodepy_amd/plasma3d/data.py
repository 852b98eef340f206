"""Attribute containers of device tensors (plasma3d/data.py:6-34).

As in the reference, construction and ``copy()`` take copies, so results
have value semantics; ``save()`` writes a compressed npz of host copies."""
from __future__ import annotations

import numpy as np


class Arrays:
    def __init__(self, xp=None, **kwargs):
        import torch
        self.xp = torch if xp is None else xp
        for name, array in kwargs.items():
            setattr(self, name, _to_tensor(array, copy=True))

    @classmethod
    def wrap(cls, **kwargs) -> 'Arrays':
        """Group existing tensors without copying (used for fresh outputs)."""
        obj = cls.__new__(cls)
        import torch
        obj.xp = torch
        for name, array in kwargs.items():
            setattr(obj, name, array)
        return obj

    def names(self):
        return [k for k in self.__dict__ if k != 'xp']

    def copy(self) -> 'Arrays':
        return Arrays.wrap(**{k: _copy(getattr(self, k)) for k in self.names()})

    def save(self, *args, **kwargs):
        data = {k: _host(getattr(self, k)) for k in self.names()}
        np.savez_compressed(*args, **kwargs, **data)


def _copy(a):
    import torch
    if isinstance(a, torch.Tensor):
        return a.clone()
    if isinstance(a, np.ndarray):
        return a.copy()
    if isinstance(a, dict):
        return dict(a)
    return a


def _to_tensor(a, copy=True):
    import torch
    from ..runtime import device
    if isinstance(a, torch.Tensor):
        return a.clone() if copy else a
    if isinstance(a, np.ndarray):
        return torch.from_numpy(np.ascontiguousarray(a)).to(device())
    return a


def _host(a):
    import torch
    if isinstance(a, torch.Tensor):
        return a.detach().cpu().numpy()
    return a
