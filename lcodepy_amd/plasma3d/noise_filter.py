"""Declustering noise filter (plasma3d/noise_filter.py:82-238) on the GPU.

``get_noise_filter(config)`` returns ``noise_filter(particles,
particles_prev) -> Arrays`` with the reference's signature and value
semantics: px and py of ``particles`` are kicked in place and the returned
Arrays carries the new ``dx_chaotic`` / ``dy_chaotic``.  The kernels
(lc3d_noise_filter) compute the lattice inhomogeneity, its Gaussian
smoothing as a circular convolution with the kernel that scipy.ndimage's
``fourier_gaussian`` applies in frequency space (:func:`gaussian_taps`), and
the restoring and damping kicks.  ``Plane2d3vPlasmaSolver.step_dxi`` runs the
same kernels inside its native step when ``declustering-enabled`` is set.

The damping coefficient is read from the key ``'damping-declustering'``, as
the reference does (noise_filter.py:89); the default config spells it
``'declustering-damping'``, so the damping is 0 unless that spelling is set.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from .. import _native, runtime
from .._native import stream_ptr
from ..config import Config
from .data import Arrays


def gaussian_taps(side: int, sigma: float, rel: float = 1e-17):
    """(taps, tap0): the circular convolution that ``fourier_gaussian(.,
    sigma)`` is on a periodic lattice of ``side`` points, g[m] = (1/n)
    sum_k exp(-2 pi^2 sigma^2 (k/n)^2) exp(2 pi i k m / n) over the signed
    FFT frequencies k, as float64 taps for the offsets tap0 .. tap0 +
    len(taps) - 1.  When the response at the Nyquist frequency,
    exp(-pi^2 sigma^2 / 2), is below rel, g is (by Poisson summation) the
    sampled Gaussian of width sigma wrapped onto the period to within rel,
    and the offsets beyond R = sigma sqrt(2 ln(1/rel)), where it has fallen
    below rel of its peak, are dropped (taps -R..R).  Otherwise (sigma below
    about 2.8 lattice cells) the cut spectrum rings and g keeps the whole
    period (tap0 = 0), as it does when -R..R covers it."""
    n = int(side)
    sigma = float(sigma)
    k = np.fft.fftfreq(n) * n
    g = np.fft.ifft(np.exp(-2 * np.pi ** 2 * sigma ** 2 * (k / n) ** 2)).real
    R = int(np.ceil(sigma * np.sqrt(2 * np.log(1 / rel))))
    if 2 * R + 1 >= n or np.exp(-np.pi ** 2 * sigma ** 2 / 2) >= rel:
        return np.ascontiguousarray(g), 0
    return np.ascontiguousarray(np.concatenate([g[n - R:], g[:R + 1]])), -R


def declustering_params(config: Config, side: int) -> dict:
    """The context parameters of the filter for a lattice of side ``side``
    (noise_filter.py:82-90)."""
    taps, tap0 = gaussian_taps(side, config.getint('declustering-averaging'))
    return dict(side=side, taps=taps, tap0=tap0,
                force=config.getfloat('declustering-force'),
                damping=config.getfloat('damping-declustering'),
                limit=config.getfloat('declustering-limit'))


def get_noise_filter(config: Config):
    ctx = runtime.context(config)
    dt = runtime.Dtypes(config.precision())

    def noise_filter(particles: Arrays, particles_prev: Arrays) -> Arrays:
        from . import _bind
        n = particles.m.numel()
        z = lambda: torch.zeros_like(particles.x_offt, dtype=torch.float64)
        dxp = getattr(particles_prev, 'dx_chaotic', None)
        dyp = getattr(particles_prev, 'dy_chaotic', None)
        dxp = z() if dxp is None else dxp.to(torch.float64).contiguous()
        dyp = z() if dyp is None else dyp.to(torch.float64).contiguous()
        dxc, dyc = z(), z()
        _bind.check_sort(particles, dt)
        s = _bind.sort_struct(particles, particles, 0)
        s.dx_chaotic_prev, s.dy_chaotic_prev = dxp.data_ptr(), dyp.data_ptr()
        s.dx_chaotic, s.dy_chaotic = dxc.data_ptr(), dyc.data_ptr()
        if n != ctx.declustering_side ** 2:
            raise ValueError(f'noise_filter: {n} particles is not the configured lattice '
                             f'({ctx.declustering_side}^2)')
        ctx.check(ctx.lib.lc3d_noise_filter(ctx.handle, ctypes.byref(s), stream_ptr()),
                  'lc3d_noise_filter')
        keep = {k: getattr(particles, k) for k in particles.names()
                if k not in ('dx_chaotic', 'dy_chaotic')}
        return Arrays.wrap(**keep, dx_chaotic=dxc, dy_chaotic=dyc)

    return noise_filter
