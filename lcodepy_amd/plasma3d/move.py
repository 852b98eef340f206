"""Plasma particle movers backed by the HIP push (plasma3d/move.py:316-383).

``get_plasma_particles_mover(config)`` returns the same two closures as the
reference: ``move_particles_wo_fields(particles, const_arrays)`` (a copy of
the particles, free-streamed) and ``move_particles_smart(fields,
particles_prev, particles_full, const_arrays)`` (in place on
``particles_full``, which it also returns)."""
from __future__ import annotations

import ctypes

from .. import runtime
from .._native import ptr, stream_ptr
from ..config import Config
from . import _bind
from .data import Arrays


def get_plasma_particles_mover(config: Config):
    ctx = runtime.context(config)
    dt = runtime.Dtypes(config.precision())
    L = ctx.lib

    def move_particles_wo_fields(particles: dict, const_arrays: Arrays) -> dict:
        particles_full = {sort: particles[sort].copy() for sort in const_arrays.sorts}
        for sort in const_arrays.sorts:
            a = particles_full[sort]
            _bind.check_sort(a, dt)
            ctx.check(L.lc3d_move_wo_fields(ctx.handle, a.m.numel(), ptr(a.m), ptr(a.x_offt),
                                            ptr(a.y_offt), ptr(a.px), ptr(a.py), ptr(a.pz),
                                            stream_ptr()), 'lc3d_move_wo_fields')
        return particles_full

    def move_particles_smart(fields: Arrays, particles_prev: dict, particles_full: dict,
                             const_arrays: Arrays) -> dict:
        f32 = int(fields.Ex.dtype == dt.half and dt.half != dt.full)
        f6 = _bind.fields6(fields)
        for sort in const_arrays.sorts:
            _bind.check_sort(particles_prev[sort], dt)
            _bind.check_sort(particles_full[sort], dt)
            s = _bind.sort_struct(particles_prev[sort], particles_full[sort], 0)
            ctx.check(L.lc3d_move_smart(ctx.handle, ctypes.byref(s), ctypes.byref(f6), f32,
                                        stream_ptr()), 'lc3d_move_smart')
        return particles_full

    return move_particles_wo_fields, move_particles_smart
