"""Diagnostics of the xi loop.

* ``calculate_energy_fluxes(grid_step_size, plasma_particles, plasma_fields)``
  (diagnostics/FXi/Strategy3D.py:6-25) keeps the reference's name, arguments
  and result (the total energy flux of a plasma layer) but reduces on the
  device through ``lc3d_energy_flux`` and returns a 0-d float64 tensor on the
  device, so a per-xi-step diagnostic can accumulate it without a host copy
  per step.
* ``SaveRunState`` (diagnostics/diagnostics_3d.py:623-738): plasma snapshots
  during the xi loop and the beam / plasma state at the end of a time step,
  in files ``Simulation`` resumes from (``load_beamfile``, and
  ``external_plasmastate`` + ``path_to_plasmastate`` through load_plasma)."""
from __future__ import annotations

import ctypes
import math
from pathlib import Path

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


def absremainder(x, y):
    """|IEEE remainder of x by y|: the distance of x from the nearest multiple
    of y, which floating-point % does not give for multiples of a decimal
    step (diagnostics_3d.py:34-47)."""
    return abs(math.remainder(x, y))


class SaveRunState:
    """Saves the run state (diagnostics_3d.py:623-738).

    ``after_step_dxi`` writes ``snapshots<name>/plasmastate_<t>_<xi>.npz``
    when the time is within half a time step of a multiple of
    ``output_period`` and xi within half an xi step of a multiple of
    ``saving_xi_period``; ``dump`` (end of a time step) writes
    ``run_states<name>/beamfile_<t>.npz`` and ``plasmastate_<t>.npz``.  The
    plasma files hold every sort as load_plasma reads it and the electrons'
    flat keys as the reference writes them (write_plasmastate)."""

    def __init__(self, output_period=1000, saving_xi_period=math.inf, save_beam=False,
                 save_plasma=False, unique_directory_name=''):
        self._output_period = output_period
        self._saving_xi_period = saving_xi_period
        self._save_beam = bool(save_beam)
        self._save_plasma = bool(save_plasma)
        self._name = unique_directory_name
        self._ro_initial = None

    def __repr__(self) -> str:
        return (f'SaveRunState(output_period={self._output_period}, '
                f'saving_xi_period={self._saving_xi_period}, save_beam={self._save_beam}, '
                f'save_plasma={self._save_plasma}, unique_directory_name={self._name})')

    def pull_config(self, config):
        self._time_step = config.getfloat('time-step')
        self._xi_step = config.getfloat('xi-step')
        self._output_period = max(self._output_period, self._time_step)
        self._saving_xi_period = max(self._saving_xi_period, self._xi_step)
        if config.getbool('rigid-beam') and self._save_beam:
            raise ValueError('SaveRunState: a rigid beam has no particles to save; '
                             'use save_beam=False')

    def set_ro_initial(self, ro_initial) -> None:
        """Background-ion density to store with the plasma (optional: the
        reference's files do not carry it and load_plasma then rebuilds it)."""
        self._ro_initial = ro_initial

    def conditions_check(self, current_time, xi_plasma_layer) -> bool:
        return (absremainder(current_time, self._output_period) <= self._time_step / 2 and
                absremainder(xi_plasma_layer, self._saving_xi_period) <= self._xi_step / 2)

    def after_step_dxi(self, current_time, xi_plasma_layer, plasma_particles, plasma_fields,
                       plasma_currents, ro_beam):
        if self._save_plasma and self.conditions_check(current_time, xi_plasma_layer):
            from .plasma3d.initialization import write_plasmastate
            d = Path(f'./snapshots{self._name}')
            d.mkdir(parents=True, exist_ok=True)
            write_plasmastate(d / f'plasmastate_{current_time:08.2f}_{xi_plasma_layer:+09.2f}.npz',
                              plasma_particles, plasma_fields, plasma_currents, xi_plasma_layer,
                              self._ro_initial)

    def dump(self, current_time, xi_plasma_layer, plasma_particles, plasma_fields,
             plasma_currents, beam_drain, clean_data=True):
        if absremainder(current_time, self._output_period) > self._time_step / 2:
            return
        d = Path(f'./run_states{self._name}')
        d.mkdir(parents=True, exist_ok=True)
        if self._save_beam:
            beam_drain.beam_buffer.save(d / f'beamfile_{current_time:08.2f}.npz')
        if self._save_plasma:
            from .plasma3d.initialization import write_plasmastate
            write_plasmastate(d / f'plasmastate_{current_time:08.2f}.npz', plasma_particles,
                              plasma_fields, plasma_currents, xi_plasma_layer, self._ro_initial)


class OutputType:
    """What a diagnostic writes (diagnostics/utils.py:3-9)."""
    __slots__ = ()
    NUMBERS = 0x1
    PICTURES = 0x2
    ALL = NUMBERS | PICTURES


class FXiType:
    """Quantities of the f(xi) diagnostic, as bit flags (FXi/FXi.py:15-46)."""
    __slots__ = ()
    Ex = 0x1
    Ey = 0x2
    Ez = 0x4
    E = Ex | Ey | Ez
    Er = Ex
    Ef = Ey
    Bx = 0x8
    By = 0x10
    Bz = 0x20
    B = Bx | By | Bz
    Bf = Bx
    EB = E | B
    Phi = 0x40
    ne = 0x80
    ni = 0x100
    n = ne | ni
    rho_beam = 0x200
    Sf = 0x400
    dx_chaotic = 0x800
    dy_chaotic = 0x1000
    dx_chaotic_perp = 0x2000
    dy_chaotic_perp = 0x4000


# Grid-sampled quantities in the reference's processing order (FXi.py:89-112).
_GRID_QUANTITIES = ('Ex', 'Ey', 'Ez', 'Bx', 'By', 'Bz', 'ne', 'ni', 'rho_beam', 'Phi')
_MAX_QUANTITIES = ('dx_chaotic', 'dy_chaotic', 'dx_chaotic_perp', 'dy_chaotic_perp')


class FXiDiag:
    """f(xi): quantities at probe points along the window, one value per xi
    step (diagnostics/FXi/FXi.py:49-135 with Strategy3D.py:28-95).

    The values stay on the device for the whole time step: every xi step one
    launch samples the grids at the probe points into the next row of a
    device buffer (``lc3d_probe_sample``), the energy flux Sf is reduced into
    that row (``lc3d_energy_flux``) and the chaotic-displacement maxima too
    (``lc3d_abs_max``).  ``dump`` reads the buffer back once and writes
    ``diagnostics[/<directory_name>]/f_xi_<t>.npz`` with one array per
    quantity (length = xi steps, times the probe points when
    ``probe_lines`` lists several), plus ``xi``; with
    ``OutputType.PICTURES`` also a plot per quantity."""

    def __init__(self, output_period=100, saving_xi_period=100, f_xi=FXiType.Ez,
                 output_type=OutputType.NUMBERS, probe_lines=None, directory_name=None):
        self._f_xi = int(f_xi)
        self._output_period = output_period
        self._saving_xi_period = saving_xi_period   # not used, as in the reference
        self._output_type = output_type
        self._probe_lines = probe_lines
        self._directory = Path('diagnostics') if directory_name is None else \
            Path('diagnostics') / directory_name
        self._is_first_step = True
        self._reset()

    def _reset(self):
        self._rows, self._xi, self._buf = 0, [], None

    def pull_config(self, config):
        import numpy as np
        import torch
        from . import runtime
        self._ctx = runtime.context(config)
        self._h = config.getfloat('transverse-step')
        self._time_step = config.getfloat('time-step')
        self._xi_step = config.getfloat('xi-step')
        self._output_period = max(self._output_period, self._time_step)
        self._saving_xi_period = max(self._saving_xi_period, self._xi_step)
        N = config.getint('window-width-steps')
        self._N = N
        probe = self._probe_lines
        if probe is None:
            ax, ay = np.zeros(1), np.zeros(1)
            self._scalar = True
        else:
            probe = np.asarray(probe, dtype=np.float64)
            if probe.ndim < 1 or probe.ndim > 2:
                raise ValueError('probe_lines must be 1D or 2D')
            ax, ay = (probe, np.zeros_like(probe)) if probe.ndim == 1 else (probe[0], probe[1])
            self._scalar = False
        ix = (N // 2 + np.round(ax / self._h)).astype(np.int32)
        iy = (N // 2 + np.round(ay / self._h)).astype(np.int32)
        if np.any(ix < 0) or np.any(ix >= N) or np.any(iy < 0) or np.any(iy >= N):
            raise ValueError(f'probe points outside the {N} x {N} grid')
        dev = runtime.device()
        self._ix = torch.as_tensor(ix).to(dev)
        self._iy = torch.as_tensor(iy).to(dev)
        self._npts = len(ix)
        f = self._f_xi
        self._grid = [q for q in _GRID_QUANTITIES if f & getattr(FXiType, q)]
        self._max = [q for q in _MAX_QUANTITIES if f & getattr(FXiType, q)]
        self._sf = bool(f & FXiType.Sf)
        self._width = len(self._grid) * self._npts + int(self._sf) + len(self._max)

    def _active(self, current_time) -> bool:
        return self._is_first_step or \
            absremainder(current_time, self._output_period) <= self._time_step / 2

    def _row(self):
        import torch
        from . import runtime
        if self._buf is None or self._rows == self._buf.shape[0]:
            cap = 1024 if self._buf is None else 2 * self._buf.shape[0]
            buf = torch.empty(cap, max(1, self._width), dtype=torch.float64,
                              device=runtime.device())
            if self._buf is not None:
                buf[:self._rows].copy_(self._buf[:self._rows])
            self._buf = buf
        row = self._buf[self._rows]
        self._rows += 1
        return row

    def after_step_dxi(self, current_time, xi_plasma_layer, plasma_particles, plasma_fields,
                       plasma_currents, rho_beam):
        if not self._active(current_time):
            return
        import torch
        row = self._row()
        ctx = self._ctx
        if self._grid:
            src = _native.ProbeSources()
            src.n = len(self._grid)
            for k, q in enumerate(self._grid):
                if q in ('ne', 'ni'):
                    t = plasma_currents.ro[0 if q == 'ne' else 1]
                elif q == 'rho_beam':
                    t = rho_beam
                else:
                    t = getattr(plasma_fields, q)
                if t.dtype not in (torch.float32, torch.float64) or not t.is_contiguous() or \
                        tuple(t.shape) != (self._N, self._N):
                    raise ValueError(f'FXiDiag: {q} must be a contiguous ({self._N}, {self._N}) '
                                     f'float32/float64 device array')
                src.src[k] = ptr(t)
                src.is_f32[k] = int(t.dtype == torch.float32)
            ctx.check(ctx.lib.lc3d_probe_sample(ctx.handle, ctypes.byref(src), self._npts,
                                                ptr(self._ix), ptr(self._iy), ptr(row),
                                                stream_ptr()), 'lc3d_probe_sample')
        col = len(self._grid) * self._npts
        if self._sf:
            psi = calculate_energy_fluxes(self._h, plasma_particles, plasma_fields)
            row[col].copy_(psi)
            col += 1
        for q in self._max:
            name = q.replace('_perp', '')
            a = getattr(plasma_particles['electrons'], name)
            if a.dtype != torch.float64 or not a.is_contiguous() or a.dim() != 2:
                raise ValueError(f'FXiDiag: electrons.{name} must be a contiguous 2-D float64 '
                                 'device array')
            nr, nc = a.shape
            if q.endswith('_perp'):
                # The central window the reference takes (Strategy3D.py:91-93):
                # a1 = int(n / 6) .. 5 a1 in both directions.
                a1 = int(nr / 6)
                r0, r1, c0, c1 = a1, 5 * a1, a1, min(nc, 5 * a1)
            else:
                r0, r1, c0, c1 = 0, nr, 0, nc
            ctx.check(ctx.lib.lc3d_abs_max(ctx.handle, ptr(a), nr, nc, r0, r1, c0, c1,
                                           ptr(row[col:col + 1]), stream_ptr()), 'lc3d_abs_max')
            col += 1
        self._xi.append(xi_plasma_layer)

    def data(self) -> dict:
        """The values gathered since the last dump, on the host (one copy)."""
        import numpy as np
        host = self._buf[:self._rows].cpu().numpy() if self._rows else \
            np.zeros((0, max(1, self._width)))
        out, col = {}, 0
        for q in self._grid:
            v = host[:, col:col + self._npts]
            out[q] = v[:, 0].copy() if self._scalar else v.copy()
            col += self._npts
        if self._sf:
            out['Sf'] = host[:, col].copy()
            col += 1
        for q in self._max:
            out[q] = host[:, col].copy()
            col += 1
        out['xi'] = np.array(self._xi)
        return out

    def dump(self, current_time, xi_plasma_layer, plasma_particles, plasma_fields,
             plasma_currents, beam_drain):
        if not self._active(current_time):
            return
        import numpy as np
        data = self.data()
        self._directory.mkdir(parents=True, exist_ok=True)
        if self._output_type & OutputType.NUMBERS:
            np.savez(self._directory / f'f_xi_{current_time:08.2f}.npz', **data)
        if self._output_type & OutputType.PICTURES:
            import matplotlib
            matplotlib.use('Agg')
            import matplotlib.pyplot as plt
            for name, values in data.items():
                if name == 'xi':
                    continue
                plt.plot(data['xi'], values)
                plt.title(name)
                plt.savefig(self._directory / f'{name}_{current_time:08.2f}.jpg')
                plt.close()
        self._reset()
        self._is_first_step = False
