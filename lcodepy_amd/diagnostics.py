"""Device-side diagnostics reductions (diagnostics/FXi/Strategy3D.py:6-25).

``calculate_energy_fluxes(grid_step_size, plasma_particles, plasma_fields)``
keeps the reference's name, arguments and result (the total energy flux of a
plasma layer) but reduces on the device through ``lc3d_energy_flux`` and
returns a 0-d float64 tensor on the device, so a per-xi-step diagnostic can
accumulate it without a host copy per step."""
from __future__ import annotations

import ctypes

from . import _native, runtime
from ._native import ptr, stream_ptr
from .plasma3d import _bind


def _context_for(N: int, h: float, precision: str, device: int) -> _native.Context:
    """The context of a running solver with this grid and precision."""
    with runtime._LOCK:
        for key, ctx in runtime._CACHE.items():
            if key[0] == device and key[1] == precision and ctx.N == N and abs(ctx.h - h) <= 1e-12 * h:
                return ctx
    raise RuntimeError(f'calculate_energy_fluxes: no lcodepy_amd context for N={N}, h={h}, '
                       f'{precision}: build the plasma solver for this config first')


def calculate_energy_fluxes(grid_step_size, plasma_particles, plasma_fields):
    """Sum over the grid of (Ez^2 + Bz^2 + (Ex - By)^2 + (Ey + Bx)^2) h^2 / 2
    plus the electrons' sum of sqrt(m^2 + px^2 + py^2 + pz^2) - m
    (Strategy3D.py:6-25), as a 0-d float64 device tensor."""
    import torch
    el = plasma_particles['electrons']
    N = int(plasma_fields.Ez.shape[0])
    precision = 'fp32-bulk' if el.px.dtype == torch.float32 else 'fp64'
    ctx = _context_for(N, float(grid_step_size), precision, plasma_fields.Ez.device.index)
    for k in ('Ex', 'Ey', 'Ez', 'Bx', 'By', 'Bz'):
        t = getattr(plasma_fields, k)
        if t.dtype != torch.float64 or not t.is_contiguous():
            raise ValueError(f'calculate_energy_fluxes: {k} must be contiguous float64')
    arrs = [el.m, el.px, el.py, el.pz]
    if not all(a.is_contiguous() for a in arrs):
        raise ValueError('calculate_energy_fluxes: particle arrays must be contiguous')
    psi = torch.empty((), dtype=torch.float64, device=plasma_fields.Ez.device)
    f6 = _bind.fields6(plasma_fields)
    ctx.check(ctx.lib.lc3d_energy_flux(ctx.handle, ctypes.byref(f6), el.m.numel(), ptr(el.m),
                                       ptr(el.px), ptr(el.py), ptr(el.pz), ptr(psi),
                                       stream_ptr()), 'lc3d_energy_flux')
    return psi
