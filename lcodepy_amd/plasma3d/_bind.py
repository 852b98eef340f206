"""Packing of Arrays groups into the C ABI's structs (include/lcodepy_amd.h)."""
from __future__ import annotations

from .. import _native
from .._native import ptr

PARTICLE_KEYS = ('x_init', 'y_init', 'x_offt', 'y_offt', 'px', 'py', 'pz', 'q', 'm')


def check_sort(a, dt) -> None:
    for k in PARTICLE_KEYS:
        t = getattr(a, k)
        want = (dt.position if k in ('x_init', 'y_init', 'x_offt', 'y_offt')
                else dt.pz if k == 'pz' else dt.bulk)
        if t.dtype != want or not t.is_cuda or not t.is_contiguous():
            raise TypeError(f'particle array {k}: need a contiguous {want} device tensor, '
                            f'got {t.dtype} on {t.device}')


def sort_struct(prev, full, column: int) -> _native.Sort:
    """Layer-k arrays from ``prev``, layer-k+1 arrays (and m) from ``full``."""
    s = _native.Sort()
    s.n = prev.m.numel()
    s.x_init, s.y_init = ptr(prev.x_init), ptr(prev.y_init)
    s.m = ptr(full.m)
    s.x_offt_prev, s.y_offt_prev = ptr(prev.x_offt), ptr(prev.y_offt)
    s.px_prev, s.py_prev, s.pz_prev, s.q_prev = ptr(prev.px), ptr(prev.py), ptr(prev.pz), ptr(prev.q)
    s.x_offt, s.y_offt = ptr(full.x_offt), ptr(full.y_offt)
    s.px, s.py, s.pz, s.q = ptr(full.px), ptr(full.py), ptr(full.pz), ptr(full.q)
    s.column = column
    return s


def fields6(f) -> _native.Fields6:
    return _native.Fields6(*[ptr(getattr(f, k)) for k in ('Ex', 'Ey', 'Ez', 'Bx', 'By', 'Bz')])


def fields7(f) -> _native.Fields:
    return _native.Fields(*[ptr(getattr(f, k)) for k in ('Ex', 'Ey', 'Ez', 'Bx', 'By', 'Bz', 'Phi')])


def currents(c) -> _native.Currents:
    for k in ('ro', 'jx', 'jy', 'jz'):
        t = getattr(c, k)
        if t.dim() != 3 or t.shape[0] != 2:
            raise ValueError(f'current {k} must have shape (2, N, N), got {tuple(t.shape)}')
    return _native.Currents(ptr(c.ro), ptr(c.jx), ptr(c.jy), ptr(c.jz))
