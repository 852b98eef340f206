"""Initial transverse/longitudinal plasma density shaping (plasma3d/profile.py).

Host numpy, run once per time step on the lattice: built-in radial profiles
'uniform' / 'stepwise' / 'channel' / 'parabolic-channel' / 'gaussian'
(profile.py:3-47), time segments (:50-109) and a user callable
``plasma-shape(t, x, y)`` (:124-125)."""
from __future__ import annotations

import numpy as np


def _radial(time_middle, profile, width, width2, density2, r):
    coef = np.ones_like(r)
    if profile in ('1', 'uniform'):
        pass
    elif profile in ('2', 'stepwise'):
        delta = width2 - width
        if delta > 0:
            mask = np.logical_and(width < r, r < width2)
            r0 = np.min(r[mask])
            coef[mask] = 1 - (r[mask] - r0) / delta
        coef[r > max(width, width2)] = 0
    elif profile in ('3', 'channel'):
        coef[r <= width] = density2
        delta = width2 - width
        if delta > 0:
            mask = np.logical_and(width < r, r < width2)
            r0 = np.min(r[mask])
            coef[mask] = (density2 * (1 + delta - r[mask]) + (r[mask] - r0)) / delta
    elif profile in ('4', 'parabolic-channel'):
        coef = density2 * (1 + (r / width) ** 2)
        coef[coef > 1] = 1
    elif profile in ('5', 'gaussian'):
        coef = np.exp(-r ** 2 / 2 / width ** 2)
        coef[r > 5 * width] = 0
    else:
        print(f'plasma-rshape: no radial profile called {profile!r} (t = {time_middle}); '
              'the plasma is taken as uniform')
    return coef


_BUILTIN = {'1', 'uniform', '2', 'stepwise', '3', 'channel', '4', 'parabolic-channel',
            '5', 'gaussian'}


def _rshape_coef(config, time_middle, x_init, y_init):
    rshape = config.get('plasma-rshape')
    r = np.sqrt(x_init ** 2 + y_init ** 2)
    coef = np.ones_like(x_init)
    if rshape in _BUILTIN:
        return _radial(time_middle, rshape, config.getfloat('plasma-width'),
                       config.getfloat('plasma-width-2'), config.getfloat('plasma-density-2'), r)
    for seg in [line.split() for line in rshape.splitlines() if line.split()]:
        if time_middle < float(seg[0]):
            return _radial(time_middle, seg[1], float(seg[2]), float(seg[3]), float(seg[4]), r)
        time_middle -= float(seg[0])
    return coef


def _zshape_coef(config, time_middle):
    zshape = config.get('plasma-zshape')
    coef = 1
    if not zshape.isspace():
        for seg in [line.split() for line in zshape.splitlines() if line.split()]:
            if time_middle < float(seg[0]):
                if seg[2] == 'L':
                    coef = (float(seg[1]) * (float(seg[0]) - time_middle)
                            + float(seg[3]) * time_middle) / float(seg[0])
                else:
                    print(f'plasma-zshape: segment kind {seg[2]!r} is not supported; the density '
                          'coefficient is taken as 1')
                break
            time_middle -= float(seg[0])
    return coef


def profile_initial_plasma(config, current_time, x_init, y_init, q, m):
    """(q * coef, m * coef) for the lattice (profile.py:111-145)."""
    time_middle = current_time - config.getfloat('time-step') / 2
    shape = config.getraw('plasma-shape')
    if callable(shape):
        coef = shape(current_time, x_init, y_init)
    else:
        coef = _zshape_coef(config, time_middle) * _rshape_coef(config, time_middle, x_init, y_init)
    if np.any(np.asarray(coef) < 0):
        print('plasma-shape gives a negative density somewhere; the plasma is taken as uniform')
        coef = np.ones_like(x_init)
    return q * coef, m * coef
