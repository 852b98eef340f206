"""One full xi step of the plasma (plasma3d/solver.py:9-78) as one native call.

``step_dxi`` keeps the reference's signature and value semantics (fresh
particle, field and current objects every step) but runs the free stream, the
three pushes, three deposits and two field solves as one stream-ordered
sequence of HIP kernels (lc3d_plasma_step), with no host round trip."""
from __future__ import annotations

import ctypes

import torch

from .. import _native, runtime
from .._native import ptr, stream_ptr
from ..config import Config
from . import _bind
from .data import Arrays
from .fields import get_field_computer, new_fields
from .move import get_plasma_particles_mover
from .rhoj import get_rhoj_computer


class Plane2d3vPlasmaSolver:
    def __init__(self, config: Config):
        self.config = config
        self.ctx = runtime.context(config)
        self.dt = runtime.Dtypes(config.precision())
        self.N = config.getint('window-width-steps')
        self.ion_model = config.get('ion-model')
        # The reference's per-stage factories, for callers that use them.
        self.compute_rhoj = get_rhoj_computer(config)
        self.compute_fields = get_field_computer(config)
        self.move_particles_wo_fields, self.move_particles = get_plasma_particles_mover(config)
        # The electrons' noise filter after the last push (solver.py:16-17,
        # 69-73) runs inside the native step.
        self.declustering_enabled = config.getbool('declustering-enabled')
        if self.declustering_enabled:
            from .noise_filter import get_noise_filter
            self.noise_filter = get_noise_filter(config)

    def step_dxi(self, particles_prev: dict, fields_prev: Arrays, currents_prev: Arrays,
                 const_arrays: Arrays, rho_beam_full, rho_beam_prev):
        N, dt, dev = self.N, self.dt, runtime.device()
        sorts = list(const_arrays.sorts.items())
        particles_full = {}
        structs = (_native.Sort * len(sorts))()
        for i, (sort, column) in enumerate(sorts):
            prev = particles_prev[sort]
            _bind.check_sort(prev, dt)
            full = Arrays.wrap(
                x_init=prev.x_init, y_init=prev.y_init, m=prev.m,
                x_offt=torch.empty_like(prev.x_offt), y_offt=torch.empty_like(prev.y_offt),
                px=torch.empty_like(prev.px), py=torch.empty_like(prev.py),
                pz=torch.empty_like(prev.pz), q=torch.empty_like(prev.q),
                **{k: getattr(prev, k) for k in prev.names()
                   if k not in _bind.PARTICLE_KEYS})
            particles_full[sort] = full
            structs[i] = _bind.sort_struct(prev, full, column)
            if self.declustering_enabled and i == 0:
                z = lambda: torch.zeros_like(prev.x_offt, dtype=torch.float64)
                dxp = getattr(prev, 'dx_chaotic', None)
                dyp = getattr(prev, 'dy_chaotic', None)
                dxp = z() if dxp is None else dxp.to(torch.float64).contiguous()
                dyp = z() if dyp is None else dyp.to(torch.float64).contiguous()
                full.dx_chaotic, full.dy_chaotic = z(), z()
                s = structs[i]
                s.dx_chaotic_prev, s.dy_chaotic_prev = dxp.data_ptr(), dyp.data_ptr()
                s.dx_chaotic, s.dy_chaotic = full.dx_chaotic.data_ptr(), full.dy_chaotic.data_ptr()
        fields_full = new_fields(N, dt.full)
        fields_half = new_fields(N, dt.half)
        cur = Arrays.wrap(**{k: torch.empty(2, N, N, dtype=dt.current, device=dev)
                             for k in ('ro', 'jx', 'jy', 'jz')})
        bg = None
        if self.ion_model == 'background':
            bg = const_arrays.ro_initial
            if bg.dtype != dt.current:
                bg = bg.to(dt.current)
            bg = bg.contiguous()
        elif len(sorts) == 1:
            for k in ('ro', 'jx', 'jy', 'jz'):
                getattr(cur, k)[1].zero_()
        rbf = rho_beam_full.to(dt.beam_rho).contiguous()
        rbp = rho_beam_prev.to(dt.beam_rho).contiguous()
        self.ctx.check(self.ctx.lib.lc3d_plasma_step(
            self.ctx.handle, len(sorts), structs, ctypes.byref(_bind.fields7(fields_prev)),
            ctypes.byref(_bind.currents(currents_prev)), ptr(rbf), ptr(rbp), ptr(bg),
            ctypes.byref(_bind.fields7(fields_full)), ctypes.byref(_bind.fields7(fields_half)),
            ctypes.byref(_bind.currents(cur)), stream_ptr()), 'lc3d_plasma_step')
        return particles_full, fields_full, cur
