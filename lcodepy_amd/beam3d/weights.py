"""Beam deposit backed by the HIP kernel (beam3d/weights.py:118-165)."""
from __future__ import annotations

from .. import runtime
from .._native import ptr, stream_ptr
from ..config import Config


def get_deposit_beam(config: Config):
    """``deposit_beam(plasma_layer_idx, x, y, xi, q_norm, out_ro0, out_ro1)``
    adds the layer's charge into the running (xi_loc) and next (1 - xi_loc)
    layers, in place."""
    ctx = runtime.context(config)

    def deposit_beam(plasma_layer_idx, x, y, xi, q_norm, out_ro0, out_ro1):
        n = int(q_norm.numel())
        if n == 0:
            return
        ctx.check(ctx.lib.lc3d_deposit_beam(ctx.handle, n, int(plasma_layer_idx), ptr(x), ptr(y),
                                            ptr(xi), ptr(q_norm), ptr(out_ro0), ptr(out_ro1),
                                            stream_ptr()), 'lc3d_deposit_beam')

    return deposit_beam
