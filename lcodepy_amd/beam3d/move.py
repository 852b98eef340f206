"""Beam pusher backed by the HIP kernel (beam3d/move.py:279-311).

Runs the reference's CPU scheme (explicit midpoint, beam3d/move.py:31-127),
the one its golden fixtures were made with."""
from __future__ import annotations

import ctypes

from .. import runtime
from .._native import ptr, stream_ptr
from ..config import Config
from ..plasma3d import _bind


def get_move_beam_particles(config: Config):
    ctx = runtime.context(config)

    def move_beam_particles(beam_layer_idx, beam_layer, fields_k, fields_k_1, lost_idxes,
                            moved_idxes, fell_idxes):
        n = beam_layer.size
        if n == 0:
            return
        fk, fk1 = _bind.fields6(fields_k), _bind.fields6(fields_k_1)
        for f in (fields_k, fields_k_1):
            if f.Ex.dtype.itemsize != 8:
                raise TypeError('the beam moves in the float64 fields of layers k and k+1')
        b = beam_layer
        ctx.check(ctx.lib.lc3d_move_beam(
            ctx.handle, n, int(beam_layer_idx), ptr(b.q_m), ptr(b.dt), ctypes.byref(fk),
            ctypes.byref(fk1), ptr(b.remaining_steps), ptr(b.id), ptr(b.x), ptr(b.y), ptr(b.xi),
            ptr(b.px), ptr(b.py), ptr(b.pz), ptr(lost_idxes), ptr(moved_idxes), ptr(fell_idxes),
            stream_ptr()), 'lc3d_move_beam')

    return move_beam_particles
