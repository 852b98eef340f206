"""String-typed simulation config with lcodePy's keys and 3D adjustments.

Mirrors ``lcode/config/config.py`` (Config class, :13-182) and
``lcode/config/default_config_values.py`` (defaults, :3-103).  The keys,
default values, ``get*`` coercion rules and the 3D adjustments are kept
identical because they decide the grid size N, the particles per cell and the
Helmholtz subtraction coefficient, and a drop-in must resolve a user's config
to the same numbers:

* ``window-width`` is doubled in 3D (config.py:191);
* N is the "good" FFT size nearest to ``int(width / step)``, ties to the lower
  value, searched in ``[estim - 100, estim + 100)`` (config.py:198-220,
  good_size :239-258);
* fineness = round(sqrt(ppc)), ppc = fineness**2 (config.py:222-234);
* the subtraction coefficient becomes 1 + xi-step only if its string form is
  exactly ``'1'`` (config.py:194-196), so ``1.0`` stays 1.0.

``processing-unit-type`` accepts ``'mi355x'`` (aliases ``'gpu'``, ``'hip'``)
for this package's HIP path; ``'cpu'`` is the reference's own Numba path and
is rejected here.
"""
from __future__ import annotations

import copy as _copy
from typing import Any

import numpy as np

DEFAULT_CONFIG_VALUES: dict[str, Any] = {
    'geometry': 'circ',
    'processing-unit-type': 'mi355x',
    'window-width': 5.0,
    'transverse-step': 0.05,
    'window-length': 15.0,
    'xi-step': 0.05,
    'time-limit': 200.5,
    'time-step': 25,
    'plasma-particles-per-cell': 10,
    'ion-model': 'mobile',
    'ion-mass': 1836,
    'plasma-shape': 'legacy',
    'plasma-rshape': 'uniform',
    'plasma-width': 1,
    'plasma-width-2': 0,
    'plasma-density-2': 0.5,
    'plasma-zshape': '''
        ''',
    'declustering-enabled': False,
    'declustering-averaging': 3,
    'declustering-force': 0.3,
    'declustering-damping': 0.1,
    'declustering-limit': 1e-3,
    'beam-substepping-energy': 2,
    'magnetic-field': 0,
    'substepping-depth': 3,
    'substepping-sensitivity': 0.2,
    'bound-padding-steps': 10,
    'plasma-padding-steps': 15,
    'dual-plasma-approach': False,
    'plasma-coarseness': 5,
    'rigid-beam': False,
    'field-solver-subtraction-coefficient': 1,
    'trapped-path-limit': 0,
    'correctotransverse-steps': 2,
    # Added by this package: arithmetic of the HIP path.
    #   'fp64'      -- float64 everywhere (the reference's precision);
    #   'fp32-bulk' -- float32 push/deposit/transforms with float64 positions,
    #                  gamma - pz, 1/(1 - v_z) and predictor-corrector averages.
    'precision': 'fp64',
    # Field-solve implementation of the HIP path: 'matrix' (DST-I/DCT-I as
    # products with N x N transform matrices) or 'fft' (LDS FFTs; only in
    # libraries built with it, not the GPU library).
    'field-solver': 'matrix',
}

HIP_PU_TYPES = ('mi355x', 'gpu', 'hip')


def _factorize(number: int) -> list[int]:
    """Prime factors of ``number`` in ascending order (config.py:229-237)."""
    factors = []
    d = 2
    while number > 1:
        while number % d == 0:
            factors.append(d)
            number //= d
        d += 1
    return factors


def good_size(number: int) -> bool:
    """True when N is odd and N - 1 factors into {2,3,5,7,11,13} with at most
    one 11 or 13 (config.py:240-248)."""
    factors = _factorize(number - 1)
    return (all(f in (2, 3, 4, 5, 7, 11, 13) for f in factors)
            and factors.count(11) + factors.count(13) < 2 and bool(number % 2))


def choose_grid_steps(window_width: float, step: float) -> int:
    """Grid size the reference picks for a (doubled) width (config.py:198-216)."""
    estim = int(window_width / step)
    lower = estim - 100 if estim - 100 > 0 else 0
    good = np.array([a for a in range(lower, estim + 100) if good_size(a)])
    return int(good[np.abs(good - estim).argmin()])


class Config:
    """Dictionary of raw values read back through ``str()`` coercion."""

    def __init__(self, new_config_values: dict | None = None, default_config: bool = True):
        self.config_values: dict[str, Any] = {}
        if default_config:
            self.update(DEFAULT_CONFIG_VALUES)
        if new_config_values is not None:
            if isinstance(new_config_values, Config):
                new_config_values = new_config_values.config_values
            self.update(new_config_values)

    # -- coercing getters (config.py:35-64) ---------------------------------
    def get(self, option_name: str, fallback: str = '') -> str:
        if option_name in self.config_values:
            return str(self.config_values.get(option_name))
        return fallback

    def getraw(self, option_name: str):
        return self.config_values.get(option_name)

    def getbool(self, option_name: str, fallback: bool = False) -> bool:
        value = self.get(option_name).lower()
        if value == 'true':
            return True
        if value == 'false':
            return False
        return fallback

    def getint(self, option_name: str, fallback: int = 0) -> int:
        try:
            return int(float(self.get(option_name)))
        except ValueError:
            return fallback

    def getfloat(self, option_name: str, fallback: float = 0.0) -> float:
        try:
            return float(self.get(option_name))
        except ValueError:
            return fallback

    # -- setters ------------------------------------------------------------
    def set(self, option_name: str, option_value: Any) -> None:
        self.config_values[option_name] = option_value
        if option_name == 'geometry' and self.config_values['geometry'] == 'circ':
            self.config_values['geometry'] = '2d'

    def update(self, new_config_values: dict | None = None) -> None:
        if new_config_values is not None:
            for key in new_config_values:
                self.set(key, new_config_values[key])

    def __copy__(self) -> 'Config':
        ret = Config(default_config=False)
        ret.config_values = _copy.copy(self.config_values)
        return ret

    # -- 3D adjustments (config.py:185-234) ---------------------------------
    def _adjust_config_values_3d(self) -> None:
        if self.get('geometry') == '3d':
            self.set('window-width', 2 * self.getfloat('window-width'))
            self.adjust_window_width_and_steps_3d()
            self.adjust_plasma_fineness()
            if self.get('field-solver-subtraction-coefficient') == '1':
                self.set('field-solver-subtraction-coefficient',
                         1 + self.getfloat('xi-step'))

    def adjust_window_width_and_steps_3d(self) -> None:
        step = self.getfloat('transverse-step')
        optimal = choose_grid_steps(self.getfloat('window-width'), step)
        self.set('window-width', optimal * step)
        self.set('window-width-steps', optimal)

    def adjust_plasma_fineness(self) -> None:
        sqrt_per_cell = np.sqrt(self.getfloat('plasma-particles-per-cell'))
        self.set('plasma-fineness', round(sqrt_per_cell))
        self.set('plasma-particles-per-cell', round(sqrt_per_cell) ** 2)
        if 0. < sqrt_per_cell < 1.:
            self.set('plasma-fineness', 1 / round(1 / sqrt_per_cell))
            self.set('plasma-particles-per-cell', 1 / round(1 / sqrt_per_cell) ** 2)

    # -- backend --------------------------------------------------------------
    def pu_type(self) -> str:
        return self.get('processing-unit-type').lower()

    def check_hip_backend(self) -> None:
        pu = self.pu_type()
        if pu not in HIP_PU_TYPES:
            raise ValueError(
                f"processing-unit-type {pu!r} is not served by lcodepy_amd; "
                f"use one of {HIP_PU_TYPES} (this package implements the HIP path only)")

    def field_solver(self) -> str:
        """'auto' (the default: the faster of the two GPU solves for this grid
        and precision, measured once per process by lcodepy_amd.autotune),
        'rfft' (paired LDS FFTs), 'matrix' (transform-matrix products on the
        matrix cores) or 'fft' (refused by the GPU library)."""
        s = self.get('field-solver', 'auto').lower()
        if s not in ('auto', 'matrix', 'fft', 'rfft'):
            raise ValueError(f"field-solver must be 'auto', 'matrix', 'fft' or 'rfft', got {s!r}")
        return s

    def precision(self) -> str:
        p = self.get('precision', 'fp64').lower()
        if p not in ('fp64', 'fp32-bulk'):
            raise ValueError(f"precision must be 'fp64' or 'fp32-bulk', got {p!r}")
        return p
