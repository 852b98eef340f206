"""Plasma state initialisation for the HIP path (plasma3d/initialization.py).

The lattice, the density profile and the solver matrices are host-side numpy
(they run once per time step, initialization.py:278-414); the background-ion
density is deposited on the GPU with the same kernel as every step
(initialization.py:351-356), so the initial net charge is exactly zero.
"""
from __future__ import annotations

import numpy as np

from .. import runtime
from ..config import Config
from .data import Arrays
from .profile import profile_initial_plasma

ELECTRON_CHARGE = -1
ELECTRON_MASS = 1


def dirichlet_matrix(xp, grid_steps, grid_step_size):
    """1/(l_i + l_j) / (2(N-1))^2 on the interior (initialization.py:15-27)."""
    k = np.arange(1, grid_steps - 1)
    lam = 4 / grid_step_size ** 2 * np.sin(k * np.pi / (2 * (grid_steps - 1))) ** 2
    return (1 / (lam[:, None] + lam[None, :])) / (2 * (grid_steps - 1)) ** 2


def mixed_matrix(xp, grid_steps, grid_step_size, subtraction_coeff):
    """1/(l_i + l_j + c) / (2(N-1))^2, i interior, j all (initialization.py:32-47)."""
    ki, kj = np.arange(1, grid_steps - 1), np.arange(grid_steps)
    li = 4 / grid_step_size ** 2 * np.sin(ki * np.pi / (2 * (grid_steps - 1))) ** 2
    lj = 4 / grid_step_size ** 2 * np.sin(kj * np.pi / (2 * (grid_steps - 1))) ** 2
    return (1 / (li[:, None] + lj[None, :] + subtraction_coeff)) / (2 * (grid_steps - 1)) ** 2


def neumann_matrix(xp, grid_steps, grid_step_size):
    """1/(l_i + l_j) / (2(N-1))^2 with [0, 0] = 0 (initialization.py:53-67)."""
    k = np.arange(0, grid_steps)
    lam = 4 / grid_step_size ** 2 * np.sin(k * np.pi / (2 * (grid_steps - 1))) ** 2
    with np.errstate(divide='ignore'):
        mul = 1 / (lam[:, None] + lam[None, :])
    mul[0, 0] = 0
    return mul / (2 * (grid_steps - 1)) ** 2


def make_plasma_grid(xp, steps, step_size, fineness):
    """Initial coordinates of the lattice along one axis, mirror-symmetric
    about 0 (initialization.py:72-104).  With spacing h / fineness and
    K = (steps // 2) * fineness points a side: an odd or fractional fineness
    puts a point on the axis (the whole multiples of the spacing below K in
    size), an even one straddles it (the odd multiples of half the spacing).  Each value
    is formed as the reference forms it, a tick count times the spacing, so
    the lattice is the same bit for bit."""
    spacing = step_size / fineness
    per_side = steps // 2 * fineness
    if fineness % 2:   # true for a fractional fineness too (K need not be whole)
        count = np.arange(per_side)
        ticks = np.concatenate([-count[:0:-1], count])
    else:
        outer = .5 + np.arange(per_side)
        ticks = np.concatenate([-outer[::-1], outer])
    lattice = ticks * spacing
    if not np.array_equal(lattice, -lattice[::-1]):
        raise AssertionError('plasma lattice is not symmetric about the axis')
    return lattice


def make_plasma_single(config: Config, current_time, steps, cell_size, fineness=2):
    """Host lattice of fineness^2 particles per cell, x-major
    (initialization.py:107-135).  Returns numpy arrays."""
    grid = make_plasma_grid(np, steps, cell_size, fineness)
    n = len(grid)
    y_init = np.ascontiguousarray(np.broadcast_to(grid, (n, n)))
    x_init = y_init.T.copy()
    q = np.ones_like(x_init) * ELECTRON_CHARGE / fineness ** 2 * cell_size ** 2
    m = np.ones_like(x_init) * ELECTRON_MASS / fineness ** 2 * cell_size ** 2
    q, m = profile_initial_plasma(config, current_time, x_init, y_init, q, m)
    z = np.zeros_like(x_init)
    return x_init, y_init, z.copy(), z.copy(), z.copy(), z.copy(), z.copy(), q, m


def particles_to_device(config: Config, host: dict) -> Arrays:
    """Sort arrays on the device with the precision form's dtypes."""
    import torch
    dt = runtime.Dtypes(config.precision())
    dev = runtime.device()
    out = {}
    for k, v in host.items():
        dtype = (dt.position if k in ('x_init', 'y_init', 'x_offt', 'y_offt')
                 else dt.pz if k == 'pz' else dt.bulk)
        if k.startswith('d') and 'chaotic' in k:
            dtype = dt.position
        out[k] = torch.as_tensor(np.ascontiguousarray(v)).to(device=dev, dtype=dtype)
    return Arrays.wrap(**out)


def init_plasma(config: Config, current_time=0):
    """(fields, particles, currents, const_arrays, xi_plasma_layer) on the
    device, single-plasma approach (initialization.py:278-414)."""
    import torch
    from .weights import get_deposit_plasma
    grid_steps = config.getint('window-width-steps')
    grid_step_size = config.getfloat('transverse-step')
    bound_padding_steps = config.getint('bound-padding-steps')
    plasma_padding_steps = config.getint('plasma-padding-steps')
    fineness = config.getfloat('plasma-fineness')
    subtraction_coeff = config.getfloat('field-solver-subtraction-coefficient')
    xi_step_size = config.getfloat('xi-step')
    ion_model = config.get('ion-model')
    if fineness > 1:
        fineness = int(fineness)
    assert grid_steps % 2 == 1
    if bound_padding_steps <= 2:
        raise Exception(f'bound-padding-steps = {bound_padding_steps}: at least 3 cells (10 by '
                        'default) must separate the plasma from the window edge, so that no '
                        'particle stencil reaches the edge cells')
    if config.getbool('dual-plasma-approach'):
        raise NotImplementedError('dual-plasma-approach is not supported in 3D '
                                  '(it is broken in the reference as well)')
    x_init, y_init, x_offt, y_offt, px, py, pz, q, m = make_plasma_single(
        config, current_time, steps=grid_steps - plasma_padding_steps * 2,
        cell_size=grid_step_size, fineness=fineness)
    z = np.zeros_like(x_init)
    host = dict(x_init=x_init, y_init=y_init, x_offt=x_offt, y_offt=y_offt, px=px, py=py, pz=pz,
                q=q, m=m, dx_chaotic=z, dy_chaotic=z, dx_chaotic_perp=z, dy_chaotic_perp=z)
    particles = {'electrons': particles_to_device(config, host)}

    dt = runtime.Dtypes(config.precision())
    dev = runtime.device()
    const_arrays = Arrays.wrap()
    const_arrays.sorts = {'electrons': 0}
    ro_e, _, _, _ = get_deposit_plasma(config)(particles, const_arrays)
    ro_initial = -ro_e[0].clone()

    def t(a, dtype=torch.float64):
        return torch.as_tensor(a).to(device=dev, dtype=dtype)

    grid = (np.arange(grid_steps) - grid_steps // 2) * grid_step_size
    const_arrays = Arrays.wrap(
        grid=t(grid), ro_initial=ro_initial,
        dirichlet_matrix=t(dirichlet_matrix(np, grid_steps, grid_step_size)),
        field_mixed_matrix=t(mixed_matrix(np, grid_steps, grid_step_size, subtraction_coeff)),
        neumann_matrix=t(neumann_matrix(np, grid_steps, grid_step_size)))

    def zeros(dtype, *shape):
        return torch.zeros(*shape, grid_steps, grid_steps, dtype=dtype, device=dev)

    fields = Arrays.wrap(**{k: zeros(dt.full) for k in ('Ex', 'Ey', 'Ez', 'Bx', 'By', 'Bz', 'Phi')})
    currents = Arrays.wrap(**{k: zeros(dt.current, 2) for k in ('ro', 'jx', 'jy', 'jz')})
    if ion_model == 'background':
        const_arrays.sorts = {'electrons': 0}
    elif ion_model == 'mobile':
        ion_mass = config.getfloat('ion-mass')
        ions = dict(host)
        ions['q'] = -q
        ions['m'] = ion_mass * m
        particles['ions'] = particles_to_device(config, ions)
        const_arrays.sorts = {'electrons': 0, 'ions': 1}
    else:
        raise ValueError(f'unknown ion-model {ion_model!r}')
    xi_plasma_layer = xi_step_size
    return fields, particles, currents, const_arrays, xi_plasma_layer


STATE_FIELDS = ('Ex', 'Ey', 'Ez', 'Bx', 'By', 'Bz', 'Phi')
STATE_CURRENTS = ('ro', 'jx', 'jy', 'jz')
STATE_PARTICLE = ('x_init', 'y_init', 'q', 'm', 'x_offt', 'y_offt', 'px', 'py', 'pz',
                  'dx_chaotic', 'dy_chaotic', 'dx_chaotic_perp', 'dy_chaotic_perp')


def _host64(a):
    import torch
    if isinstance(a, torch.Tensor):
        a = a.detach().cpu().numpy()
    return np.asarray(a, dtype=np.float64)


def _particle_record(arrs) -> np.ndarray:
    """One sort as a structured (n, n) array with a float64 field per
    quantity: the per-sort entry ``state[sort][field]`` load_plasma reads
    (initialization.py:430-445).  Quantities a sort lacks are zeros."""
    first = _host64(getattr(arrs, 'x_init'))
    rec = np.zeros(first.shape, dtype=[(k, np.float64) for k in STATE_PARTICLE])
    for k in STATE_PARTICLE:
        v = getattr(arrs, k, None)
        if v is not None:
            rec[k] = _host64(v)
    return rec


def write_plasmastate(path, particles: dict, fields, currents, xi_plasma_layer,
                      ro_initial=None) -> None:
    """Writes a plasma state that load_plasma reads back.

    The file holds both layouts of the reference: every sort as a structured
    record under its name (what load_plasma reads, initialization.py:417-460)
    and the electrons' quantities as flat keys beside the fields and
    currents (what SaveRunState writes, diagnostics_3d.py:666-689), plus
    ``xi_plasma_layer`` and, when given, ``rho_initial``."""
    data = {'xi_plasma_layer': np.array([float(xi_plasma_layer)])}
    for sort, arrs in particles.items():
        data[sort] = _particle_record(arrs)
    el = particles.get('electrons')
    if el is not None:
        for k in STATE_PARTICLE:
            data[k] = data['electrons'][k].copy()
    for k in STATE_FIELDS:
        data[k] = _host64(getattr(fields, k))
    for k in STATE_CURRENTS:
        data[k] = _host64(getattr(currents, k))
    if ro_initial is not None:
        data['rho_initial'] = _host64(ro_initial)
    np.savez_compressed(path, **data)


def read_plasmastate(path, sorts) -> tuple:
    """(fields, particles, currents, rho_initial or None, xi_plasma_layer or
    None) of a plasma state file as host float64 arrays.  A sort is read from
    its structured record; a file in SaveRunState's flat layout (no record)
    gives its flat quantities to the electrons."""
    with np.load(path) as state:
        files = set(state.files)
        fields = {k: np.array(state[k], dtype=np.float64) for k in STATE_FIELDS}
        currents = {k: np.array(state[k], dtype=np.float64) for k in STATE_CURRENTS}
        particles = {}
        for sort in sorts:
            if sort in files:
                rec = state[sort]
                particles[sort] = {k: np.array(rec[k], dtype=np.float64) for k in STATE_PARTICLE}
            elif sort == 'electrons' and 'x_init' in files:
                particles[sort] = {k: np.array(state[k], dtype=np.float64)
                                   for k in STATE_PARTICLE}
            else:
                raise KeyError(f'plasma state {path}: no particle record for sort {sort!r}')
        rho_initial = np.array(state['rho_initial']) if 'rho_initial' in files else None
        xi = float(np.asarray(state['xi_plasma_layer']).ravel()[0]) \
            if 'xi_plasma_layer' in files else None
    return fields, particles, currents, rho_initial, xi


def load_plasma(config: Config, path_to_plasmastate: str):
    """initialization.py:417-460: the fields, every sort's particles and the
    currents of a saved state, with rho_initial and xi_plasma_layer when the
    file holds them (see write_plasmastate / read_plasmastate)."""
    import torch
    fields, particles, currents, const_arrays, xi_plasma_layer = init_plasma(config)
    dt = runtime.Dtypes(config.precision())
    dev = runtime.device()
    f, parts, cur, rho_initial, xi = read_plasmastate(path_to_plasmastate, const_arrays.sorts)
    fields = Arrays.wrap(**{k: torch.as_tensor(f[k]).to(device=dev, dtype=dt.full)
                            for k in STATE_FIELDS})
    for sort, host in parts.items():
        particles[sort] = particles_to_device(config, host)
    currents = Arrays.wrap(**{k: torch.as_tensor(cur[k]).to(device=dev, dtype=dt.current)
                              for k in STATE_CURRENTS})
    if rho_initial is not None:
        const_arrays.ro_initial = torch.as_tensor(rho_initial).to(device=dev, dtype=dt.current)
    if xi is not None:
        xi_plasma_layer = xi
    return fields, particles, currents, const_arrays, xi_plasma_layer
