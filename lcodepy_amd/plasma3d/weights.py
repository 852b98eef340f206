"""Plasma deposit backed by the HIP cell-sorted gather deposit
(plasma3d/weights.py:193-287)."""
from __future__ import annotations

import ctypes

from .. import runtime
from .._native import ptr, stream_ptr
from ..config import Config
from . import _bind
from .data import Arrays


def get_deposit_plasma(config: Config):
    """``deposit_plasma(all_particles, const_arrays) -> (ro, jx, jy, jz)``,
    each (2, N, N), sort column from ``const_arrays.sorts``."""
    ctx = runtime.context(config)
    dt = runtime.Dtypes(config.precision())
    N = config.getint('window-width-steps')
    L = ctx.lib
    if config.getbool('dual-plasma-approach'):
        raise NotImplementedError('dual-plasma-approach is not supported in 3D')

    def deposit_plasma(all_particles: dict, const_arrays: Arrays):
        import torch
        dev = runtime.device()
        ro, jx, jy, jz = (torch.zeros(2, N, N, dtype=dt.current, device=dev) for _ in range(4))
        for sort, column in const_arrays.sorts.items():
            a = all_particles[sort]
            _bind.check_sort(a, dt)
            s = _bind.sort_struct(a, a, column)
            ctx.check(L.lc3d_deposit(ctx.handle, ctypes.byref(s), ptr(ro[column]),
                                     ptr(jx[column]), ptr(jy[column]), ptr(jz[column]),
                                     None, None, None, None, None, stream_ptr()), 'lc3d_deposit')
        return ro, jx, jy, jz

    return deposit_plasma
