"""Device contexts keyed by run geometry, and dtype rules of the two precisions.

A context (``_native.Context``) owns the kernels' workspaces for one grid
size, step sizes, precision and particle capacity.  Every factory built from
the same config shares one context, as the reference's factories share one
config (plasma3d/solver.py:10-17).
"""
from __future__ import annotations

import math
import threading

import numpy as np

from . import _native
from .config import Config

_CACHE: dict = {}
_LOCK = threading.Lock()


_HAVE_GPU = False


def device():
    """torch's current device; raises without a GPU (no CPU fallback).  The
    availability check runs until it first succeeds (it is slow)."""
    global _HAVE_GPU
    import torch
    if not _HAVE_GPU:
        if not torch.cuda.is_available():
            raise RuntimeError('lcodepy_amd needs a ROCm GPU: torch.cuda.is_available() is '
                               'False (the HIP path has no CPU fallback)')
        _HAVE_GPU = True
    return torch.device('cuda', torch.cuda.current_device())


class Dtypes:
    """Tensor dtypes of one precision form (see include/lcodepy_amd.h)."""

    def __init__(self, precision: str):
        import torch
        self.precision = precision
        bulk = precision == 'fp32-bulk'
        self.position = torch.float64          # x_init, y_init, x_offt, y_offt
        self.bulk = torch.float32 if bulk else torch.float64   # px py q m
        self.pz = torch.float64                # gamma - pz must not cancel
        self.full = torch.float64              # fields of layers k, k+1
        self.half = torch.float32 if bulk else torch.float64   # half-step fields
        self.current = torch.float32 if bulk else torch.float64
        self.beam_rho = torch.float64


def lattice_size(config: Config) -> int:
    """Macro-particles per sort of the reference's lattice
    (plasma3d/initialization.py:72-135, 326-330)."""
    from .plasma3d.initialization import make_plasma_grid
    fin = config.getfloat('plasma-fineness')
    fin = int(fin) if fin > 1 else fin
    steps = config.getint('window-width-steps') - 2 * config.getint('plasma-padding-steps')
    n = len(make_plasma_grid(np, steps, config.getfloat('transverse-step'), fin))
    return n * n


def eigenvalues(N: int, h: float) -> np.ndarray:
    """4/h^2 sin^2(k pi / (2(N-1))), k = 0..N-1, as numpy evaluates it in
    plasma3d/initialization.py:24, 43-44, 63."""
    k = np.arange(N)
    return 4 / h ** 2 * np.sin(k * np.pi / (2 * (N - 1))) ** 2


def geometry(config: Config) -> dict:
    h = config.getfloat('transverse-step')
    N = config.getint('window-width-steps')
    max_radius = h * N / 2
    return dict(N=N, h=h, xi_step=config.getfloat('xi-step'),
                boundary=h * (N / 2 - config.getint('bound-padding-steps')),
                subtraction_coeff=config.getfloat('field-solver-subtraction-coefficient'),
                lost_radius=max(0.9 * max_radius, max_radius - 1))


def context(config: Config, max_particles: int | None = None,
            private: bool = False) -> _native.Context:
    """Shared context for an adjusted 3D config; with ``private`` a new one
    that nothing else uses (work of several contexts may then run
    concurrently on different streams: a context's workspaces serve one
    stream at a time)."""
    import torch
    config.check_hip_backend()
    g = geometry(config)
    prec = config.precision()
    cap = max(lattice_size(config), int(max_particles or 0))
    radix = config.getraw('fft-radix-plan')
    radix = tuple(int(r) for r in radix) if radix else None
    dev = device().index
    solver = config.field_solver()
    shape = (int(config.get('rfft-cap', 0) or 0), int(config.get('rfft-tpb', 0) or 0))
    # mask of the FFT kernels by waves (1 rows forward, 2 columns, 4 rows
    # inverse); true = all three
    wave = str(config.get('rfft-wave', 0) or 0).strip().lower()
    wave = 7 if wave in ('true', 'yes', 'on') else (0 if wave in ('false', 'no', 'off') else int(wave))
    if solver == 'auto':
        from . import autotune
        choice = autotune.solve_choice(dev, g['N'], g['h'], g['xi_step'],
                                       g['subtraction_coeff'], prec)
        solver, radix = choice['field_solver'], choice['fft_radix']
        shape = (choice['rfft_cap'], choice['rfft_tpb'])
        wave = int(choice.get('rfft_wave') or 0)
    decl = None
    if config.getbool('declustering-enabled'):
        from .plasma3d.noise_filter import declustering_params
        side = math.isqrt(lattice_size(config))
        decl = declustering_params(config, side)
    dkey = None if decl is None else (decl['side'], config.getint('declustering-averaging'),
                                      decl['force'], decl['damping'], decl['limit'])
    key = (dev, prec, radix, solver, shape, wave, dkey) + tuple(sorted(g.items()))
    with _LOCK:
        ctx = None if private else _CACHE.get(key)
        if ctx is None or ctx.max_particles < cap:
            with torch.cuda.device(dev):
                ctx = _native.Context(eigenvalues=eigenvalues(g['N'], g['h']), precision=prec,
                                      max_particles=cap, fft_radix=radix,
                                      field_solver=solver, rfft_cap=shape[0],
                                      rfft_tpb=shape[1], rfft_wave=wave, declustering=decl, **g)
            if not private:
                _CACHE[key] = ctx
    return ctx
