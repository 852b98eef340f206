"""Field solver backed by the HIP DST/DCT solve (plasma3d/fields.py:269-324)."""
from __future__ import annotations

import ctypes

from .. import runtime
from .._native import ptr, stream_ptr
from ..config import Config
from . import _bind
from .data import Arrays

FIELD_KEYS = ('Ex', 'Ey', 'Ez', 'Bx', 'By', 'Bz', 'Phi')


def new_fields(N: int, dtype) -> Arrays:
    import torch
    dev = runtime.device()
    return Arrays.wrap(**{k: torch.empty(N, N, dtype=dtype, device=dev) for k in FIELD_KEYS})


def get_field_computer(config: Config):
    """``compute_fields(fields, fields_prev, const, rho_beam_full,
    rho_beam_prev, currents_prev, currents_full) -> (fields_full,
    fields_half)``; ``fields`` is the Helmholtz predictor."""
    ctx = runtime.context(config)
    dt = runtime.Dtypes(config.precision())
    N = config.getint('window-width-steps')
    L = ctx.lib

    def compute_fields(fields, fields_prev, const, rho_beam_full, rho_beam_prev,
                       currents_prev, currents_full):
        full = new_fields(N, dt.full)
        half = new_fields(N, dt.half)
        pred32 = int(fields.Ex.dtype != dt.full)
        rbf = rho_beam_full.to(dt.beam_rho).contiguous()
        rbp = rho_beam_prev.to(dt.beam_rho).contiguous()
        args = (ctypes.byref(_bind.fields7(fields)), pred32,
                ctypes.byref(_bind.fields7(fields_prev)), ptr(rbf), ptr(rbp),
                ctypes.byref(_bind.currents(currents_prev)),
                ctypes.byref(_bind.currents(currents_full)),
                ctypes.byref(_bind.fields7(full)), ctypes.byref(_bind.fields7(half)))
        ctx.check(L.lc3d_compute_fields(ctx.handle, *args, stream_ptr()), 'lc3d_compute_fields')
        return full, half

    return compute_fields
