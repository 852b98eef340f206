"""Independent pins of each oracle function (what the fixtures leave open).

Every check is a mathematical fact or a special case that a plausible slip
(dropped term, wrong sign, swapped axis, wrong normalisation) would break:
discrete-operator residuals of the three solves, the cardinal quartic
B-spline, charge conservation, single-particle motion, and bit-identity of
the banded-parallel deposit with the serial loop order."""
import math

import numpy as np
import pytest

from lcodepy_amd.config import Config, choose_grid_steps, good_size
from oracle import lcode_oracle as O


# -- config (config.py:185-258) ----------------------------------------------
@pytest.mark.parametrize('width,step,n', [(16, 0.05, 321), (12.8, 0.1, 127), (12.9, 0.1, 129),
                                          (25.6, 0.05, 513), (25.6, 0.025, 1025),
                                          (30.72, 0.04, 769)])
def test_grid_size_choice(width, step, n):
    assert choose_grid_steps(width, step) == n
    assert good_size(n)


def test_good_size_rules():
    assert good_size(513) and good_size(769) and good_size(385)
    assert not good_size(512)            # even
    assert not good_size(2 * 11 * 13 + 1)  # two of {11, 13}
    assert not good_size(17 * 2 + 1)     # factor 17


@pytest.mark.parametrize('ppc,fin,ppc_out', [(1, 1, 1), (2, 1, 1), (3, 2, 4), (4, 2, 4),
                                             (9, 3, 9), (10, 3, 9)])
def test_fineness_rounding(ppc, fin, ppc_out):
    c = Config({'geometry': '3d', 'plasma-particles-per-cell': ppc})
    c._adjust_config_values_3d()
    assert c.getint('plasma-fineness') == fin
    assert c.getint('plasma-particles-per-cell') == ppc_out


@pytest.mark.parametrize('value,expected', [(1, 1.05), ('1', 1.05), (1.0, 1.0), (0.5, 0.5)])
def test_subtraction_coefficient_string_rule(value, expected):
    c = Config({'geometry': '3d', 'xi-step': 0.05,
                'field-solver-subtraction-coefficient': value})
    c._adjust_config_values_3d()
    assert c.getfloat('field-solver-subtraction-coefficient') == pytest.approx(expected, abs=0)


# -- quartic spline (weights.py:12-31) ----------------------------------------
def _cardinal_bspline4(t):
    """Centred degree-4 cardinal B-spline from the truncated-power formula."""
    s = t + 2.5
    return sum((-1) ** j * math.comb(5, j) * max(s - j, 0.0) ** 4 for j in range(6)) / 24.0


@pytest.mark.parametrize('x', [-0.5, -0.37, -0.1, 0.0, 0.013, 0.25, 0.4999])
def test_weight4_is_the_cardinal_bspline(x):
    w = [O.weight4(x, k) for k in range(-2, 3)]
    for k, wk in zip(range(-2, 3), w):
        assert wk == pytest.approx(_cardinal_bspline4(k - x), abs=1e-13)  # formula cancels ~1e-14
    assert sum(w) == pytest.approx(1.0, abs=1e-15)
    assert sum(k * wk for k, wk in zip(range(-2, 3), w)) == pytest.approx(x, abs=1e-15)
    for k in range(-2, 3):
        assert O.weight4(-x, -k) == pytest.approx(O.weight4(x, k), abs=1e-16)


def test_weight4_at_cell_centre():
    assert [O.weight4(0.0, k) for k in range(-2, 3)] == pytest.approx(
        [1 / 384, 19 / 96, 115 / 192, 19 / 96, 1 / 384], abs=0)


# -- transforms and solves (fields.py) -----------------------------------------
def _lap(u, h):
    return (u[2:, 1:-1] + u[:-2, 1:-1] + u[1:-1, 2:] + u[1:-1, :-2] - 4 * u[1:-1, 1:-1]) / h ** 2


@pytest.mark.parametrize('N', [9, 17, 33])
def test_dst_dct_self_inverse(N):
    rng = np.random.default_rng(N)
    a = rng.standard_normal((N - 2, N - 2))
    assert np.allclose(O.dst2d(O.dst2d(a)), (2 * (N - 1)) ** 2 * a, rtol=1e-12, atol=1e-10)
    b = rng.standard_normal((N, N))
    assert np.allclose(O.dct2d(O.dct2d(b)), (2 * (N - 1)) ** 2 * b, rtol=1e-12, atol=1e-10)


@pytest.mark.parametrize('N,h', [(17, 0.1), (33, 0.05)])
def test_dirichlet_solve_inverts_the_5point_laplacian(N, h):
    rng = np.random.default_rng(1)
    f = rng.standard_normal((N - 2, N - 2))
    u = np.pad(O.dst2d(O.dst2d(f) * O.dirichlet_matrix(N, h)), 1)
    assert np.allclose(-_lap(u, h), f, rtol=1e-10, atol=1e-10)


@pytest.mark.parametrize('N,h,c', [(17, 0.1, 1.0), (33, 0.05, 1.05)])
def test_mixed_solve_is_helmholtz_dirichlet_x_neumann_y(N, h, c):
    rng = np.random.default_rng(2)
    f = rng.standard_normal((N - 2, N))
    u = O.mix2d(O.mixed_matrix(N, h, c) * O.mix2d(f)[1:-1, :])
    assert np.all(u[0] == 0) and np.all(u[-1] == 0)
    ue = np.concatenate([u[:, 1:2], u, u[:, -2:-1]], axis=1)   # even reflection in y
    res = -(ue[2:, 1:-1] + ue[:-2, 1:-1] + ue[1:-1, 2:] + ue[1:-1, :-2]
            - 4 * ue[1:-1, 1:-1]) / h ** 2 + c * u[1:-1, :]
    assert np.allclose(res, f, rtol=1e-10, atol=1e-10)


@pytest.mark.parametrize('N,h', [(17, 0.1), (33, 0.05)])
def test_neumann_solve_inverts_laplacian_on_the_non_constant_modes(N, h):
    rng = np.random.default_rng(3)
    f = rng.standard_normal((N, N))
    u = O.dct2d(O.dct2d(f) * O.neumann_matrix(N, h))
    w = np.ones(N)
    w[0] = w[-1] = 0.5
    fbar = (w[:, None] * w[None, :] * f).sum() / (w.sum() ** 2)
    ue = np.pad(u, 1, mode='reflect')
    assert np.allclose(-_lap(ue, h), f - fbar, rtol=1e-9, atol=1e-9)


def test_eigenmode_closed_form():
    N, h = 33, 0.05
    i = np.arange(1, N - 1)
    a, b = 3, 5
    mode = np.sin(a * np.pi * i / (N - 1))[:, None] * np.sin(b * np.pi * i / (N - 1))[None, :]
    lam = O.eigenvalues(N, h)
    u = O.dst2d(O.dst2d(mode) * O.dirichlet_matrix(N, h))
    assert np.allclose(u, mode / (lam[a] + lam[b]), rtol=1e-12, atol=1e-14)


def test_compute_fields_zero_sources_give_zero_fields():
    p = O.Params(N=17, h=0.1, xi_step=0.05, subtraction_coeff=1.0)
    const = O.Const(ro_initial=np.zeros((17, 17)), dirichlet=O.dirichlet_matrix(17, 0.1),
                    mixed=O.mixed_matrix(17, 0.1, 1.0), neumann=O.neumann_matrix(17, 0.1))
    z = O.zero_fields(17)
    full, half = O.compute_fields(p, z, z, const, np.zeros((17, 17)), np.zeros((17, 17)),
                                  O.zero_currents(17), O.zero_currents(17))
    for k in full:
        assert not np.any(full[k]) and not np.any(half[k])


# -- deposit (weights.py:35-67) -------------------------------------------------
def _one_particle(x, y, px=0.0, py=0.0, pz=0.0, q=-0.0025, m=0.0025):
    f = lambda v: np.array([[v]], dtype=np.float64)
    return dict(x_init=f(x), y_init=f(y), x_offt=f(0.0), y_offt=f(0.0), px=f(px), py=f(py),
                pz=f(pz), q=f(q), m=f(m))


def test_deposit_single_particle_at_cell_centre():
    N, h = 17, 0.1
    p = O.Params(N=N, h=h, xi_step=0.05, subtraction_coeff=1.0)
    a = _one_particle(0.2, -0.3)     # centre of cell (ix, iy) = (10, 5)
    g = [np.zeros((N, N)) for _ in range(4)]
    O.deposit_sort(p, a, *g)
    w = np.array([1 / 384, 19 / 96, 115 / 192, 19 / 96, 1 / 384])
    expect = np.zeros((N, N))
    expect[8:13, 3:8] = np.outer(w, w) / h ** 2 * -0.0025
    assert np.allclose(g[0], expect, rtol=1e-13, atol=1e-18)
    assert not np.any(g[1]) and not np.any(g[2]) and not np.any(g[3])


def test_deposit_conserves_charge_and_current():
    N, h = 33, 0.1
    p = O.Params(N=N, h=h, xi_step=0.05, subtraction_coeff=1.0)
    rng = np.random.default_rng(4)
    n = 500
    a = dict(x_init=rng.uniform(-1, 1, n), y_init=rng.uniform(-1, 1, n), x_offt=np.zeros(n),
             y_offt=np.zeros(n), px=rng.normal(0, .3, n), py=rng.normal(0, .3, n),
             pz=rng.normal(0, .3, n), q=np.full(n, -0.01), m=np.full(n, 0.01))
    g = [np.zeros((N, N)) for _ in range(4)]
    O.deposit_sort(p, a, *g)
    gamma = np.sqrt(a['m'] ** 2 + a['px'] ** 2 + a['py'] ** 2 + a['pz'] ** 2)
    dro = a['q'] / (1 - a['pz'] / gamma)
    assert g[0].sum() * h * h == pytest.approx(dro.sum(), rel=1e-12)
    assert g[1].sum() * h * h == pytest.approx((a['px'] * dro / gamma).sum(), rel=1e-11)
    assert g[3].sum() * h * h == pytest.approx((a['pz'] * dro / gamma).sum(), rel=1e-11)


def test_banded_parallel_deposit_equals_serial_bitwise():
    N, h = 65, 0.05
    p = O.Params(N=N, h=h, xi_step=0.05, subtraction_coeff=1.0)
    rng = np.random.default_rng(5)
    n = 20000
    a = dict(x_init=rng.uniform(-1.3, 1.3, n), y_init=rng.uniform(-1.3, 1.3, n),
             x_offt=np.zeros(n), y_offt=np.zeros(n), px=rng.normal(0, .1, n),
             py=rng.normal(0, .1, n), pz=rng.normal(0, .1, n), q=np.full(n, -1e-3),
             m=np.full(n, 1e-3))
    a['q'][::7] = 0
    g1 = [np.zeros((N, N)) for _ in range(4)]
    g8 = [np.zeros((N, N)) for _ in range(4)]
    O.deposit_sort(p, a, *g1, nthreads=1)
    O.deposit_sort(p, a, *g8, nthreads=8)
    for u, v in zip(g1, g8):
        assert np.array_equal(u, v)


# -- plasma push (move.py:28-165) -----------------------------------------------
def _push(p, fields, a, full=None):
    prev = {'electrons': a}
    full = O.move_wo_fields(p, prev) if full is None else full
    return O.move_smart(p, fields, prev, full)['electrons']


def test_free_stream_estimate():
    p = O.Params(N=33, h=0.1, xi_step=0.05, subtraction_coeff=1.0)
    a = _one_particle(0.0, 0.0, px=0.3, py=-0.2, pz=0.1)
    g = math.sqrt(0.0025 ** 2 + 0.1 ** 2 + 0.3 ** 2 + 0.2 ** 2)
    full = O.move_wo_fields(p, {'electrons': a})['electrons']
    assert full['x_offt'][0, 0] == pytest.approx(0.3 / (g - 0.1) * 0.05, rel=1e-15)
    assert full['y_offt'][0, 0] == pytest.approx(-0.2 / (g - 0.1) * 0.05, rel=1e-15)
    assert a['x_offt'][0, 0] == 0.0     # the previous state is not touched


def test_push_without_fields_is_free_streaming():
    p = O.Params(N=33, h=0.1, xi_step=0.05, subtraction_coeff=1.0)
    a = _one_particle(0.1, 0.2, px=0.3, py=-0.2, pz=0.1)
    out = _push(p, O.zero_fields(33), a)
    g = math.sqrt(0.0025 ** 2 + 0.1 ** 2 + 0.3 ** 2 + 0.2 ** 2)
    assert out['px'][0, 0] == 0.3 and out['pz'][0, 0] == 0.1
    assert out['x_offt'][0, 0] == pytest.approx(0.3 / (g - 0.1) * 0.05, rel=1e-14)


def test_push_in_uniform_ez():
    p = O.Params(N=33, h=0.1, xi_step=0.05, subtraction_coeff=1.0)
    f = O.zero_fields(33)
    f['Ez'][:] = 0.7
    q, m = -0.0025, 0.0025
    out = _push(p, f, _one_particle(0.0, 0.0, q=q, m=m))
    p1 = q * 0.05 * 0.7 / 2
    g1 = math.sqrt(m * m + p1 * p1)
    assert out['pz'][0, 0] == pytest.approx(q * 0.05 / (1 - p1 / g1) * 0.7, rel=1e-14)
    assert out['px'][0, 0] == 0 and out['x_offt'][0, 0] == 0


def test_push_reflects_and_absorbs_at_the_wall():
    p = O.Params(N=33, h=0.1, xi_step=0.05, subtraction_coeff=1.0)
    b = p.boundary                               # 0.1 * (16.5 - 10) = 0.65
    a = _one_particle(b - 0.01, 0.0, px=5.0, pz=0.0)
    out = _push(p, O.zero_fields(33), a)
    x = b - 0.01 + out['x_offt'][0, 0]
    assert x < b and out['px'][0, 0] == 0 and out['q'][0, 0] != 0
    a = _one_particle(5.5 * b, 0.0)              # still outside after both reflections
    out = _push(p, O.zero_fields(33), a, full={'electrons': {k: v.copy() for k, v in a.items()}})
    assert out['q'][0, 0] == 0 and out['x_offt'][0, 0] == 0


def test_push_skips_absorbed_particles():
    p = O.Params(N=33, h=0.1, xi_step=0.05, subtraction_coeff=1.0)
    a = _one_particle(0.1, 0.1, px=0.2, q=0.0)
    f = O.zero_fields(33)
    f['Ex'][:] = 1.0
    full = {'electrons': {k: v.copy() for k, v in a.items()}}
    out = O.move_smart(p, f, {'electrons': a}, full)['electrons']
    assert out['px'][0, 0] == 0.2 and out['x_offt'][0, 0] == 0


# -- beam (beam3d) ------------------------------------------------------------------
def _beam(n, rng, xi0, xi_step, r=0.5):
    b = O.empty_beam()
    b = {k: v for k, v in b.items()}
    b['xi'] = rng.uniform(xi0 - xi_step, xi0, n)
    b['x'] = rng.uniform(-r, r, n)
    b['y'] = rng.uniform(-r, r, n)
    b['px'] = rng.normal(0, 0.01, n)
    b['py'] = rng.normal(0, 0.01, n)
    b['pz'] = np.full(n, 1000.0)
    b['q_m'] = np.full(n, -1.0)
    b['q_norm'] = np.full(n, 0.01)
    b['id'] = np.arange(1, n + 1, dtype=np.int64)
    b['dt'] = np.zeros(n)
    b['remaining_steps'] = np.zeros(n, dtype=np.int64)
    return b


def test_beam_deposit_two_layer_split_conserves_charge():
    p = O.Params(N=33, h=0.1, xi_step=0.05, subtraction_coeff=1.0)
    rng = np.random.default_rng(6)
    calc = O.BeamCalc(p)
    b = _beam(300, rng, -0.05 * 4, 0.05)
    rho_now = calc.layout_beam_layer(b, 4)
    total = (rho_now.sum() + calc.rho_layout.sum() / p.h ** 2) * p.h ** 2
    assert total == pytest.approx(b['q_norm'].sum(), rel=1e-12)
    xi_loc = (-0.05 * 4 - b['xi']) / 0.05
    assert rho_now.sum() * p.h ** 2 == pytest.approx((b['q_norm'] * xi_loc).sum(), rel=1e-12)


def test_beam_push_zero_field_drift_and_flags():
    p = O.Params(N=33, h=0.1, xi_step=0.05, subtraction_coeff=1.0, time_step=0.5)
    rng = np.random.default_rng(7)
    b = _beam(4, rng, -0.05 * 3, 0.05)
    b['xi'][:] = -0.17                            # inside layer [-0.2, -0.15)
    b['x'][1] = 0.999 * p.lost_radius             # will cross the wall
    b['px'][1] = 50.0
    b['pz'][2] = 1.0                              # slow: falls to the next layer
    z = O.zero_fields(33)
    x0, px0 = b['x'].copy(), b['px'].copy()
    lost, moved, fell = O.BeamCalc(p).move_beam_layer(b, 0, 4, z, z)
    assert list(lost['id']) == [-2]
    assert 3 in list(fell['id'])
    assert 1 in list(moved['id'])
    i = list(moved['id']).index(1)
    g = math.sqrt(1 + px0[0] ** 2 + b['py'][0] ** 2 + 1000.0 ** 2)
    assert moved['x'][i] == pytest.approx(x0[0] + 0.5 * px0[0] / g, rel=1e-12)
    assert moved['px'][i] == px0[0]


def test_beam_substep_multiplier():
    qm = np.array([-1.0, -1.0, -1.0])
    pz = np.array([1000.0, 1.0, 0.1])
    dt = np.zeros(3)
    O.lib().oracle_beam_substep(O.ctypes.c_int64(3), O._p(qm), O._p(pz), O.ctypes.c_double(2.0),
                                O._p(dt))
    for d, z in zip(dt, pz):
        max_dt = math.sqrt(math.sqrt(1 + z * z) / 2.0)
        assert d <= max_dt and (d == 1.0 or 2 * d > max_dt)


# ---- energy flux (diagnostics/FXi/Strategy3D.py:6-25) ----------------------------

def _zero_fields(N):
    return {k: np.zeros((N, N)) for k in ('Ex', 'Ey', 'Ez', 'Bx', 'By', 'Bz')}


def _rest(n, m=0.25):
    return {'m': np.full(n, m), 'px': np.zeros(n), 'py': np.zeros(n), 'pz': np.zeros(n)}


def test_energy_flux_zero_state():
    assert O.energy_flux(0.1, _rest(10), _zero_fields(7)) == 0.0


def test_energy_flux_uniform_fields():
    # Uniform Ez = a, Bz = b over N^2 cells: N^2 (a^2 + b^2) h^2 / 2.
    N, h, a, b = 9, 0.05, 0.3, -0.7
    f = _zero_fields(N)
    f['Ez'][:] = a
    f['Bz'][:] = b
    assert np.isclose(O.energy_flux(h, _rest(4), f), N * N * (a * a + b * b) * h * h / 2,
                      rtol=1e-14)


def test_energy_flux_transverse_combination_cancels():
    # Ex = By and Ey = -Bx carry no flux; Ex = -By doubles it: (2c)^2 h^2 / 2 per cell.
    rng = np.random.default_rng(5)
    N, h = 11, 0.1
    f = _zero_fields(N)
    c = rng.normal(size=(N, N))
    d = rng.normal(size=(N, N))
    f['Ex'], f['By'], f['Ey'], f['Bx'] = c, c.copy(), d, -d
    assert O.energy_flux(h, _rest(3), f) == 0.0
    f['By'] = -c
    assert np.isclose(O.energy_flux(h, _rest(3), f), np.sum((2 * c) ** 2) * h * h / 2,
                      rtol=1e-13)


def test_energy_flux_single_particle():
    # m = 3, |p| = 4 (px, py, pz = 0, 0, 4 and 2, 2, 2*sqrt(2)): gamma m - m = 5 - 3.
    e = {'m': np.array([3.0, 3.0]), 'px': np.array([0.0, 2.0]), 'py': np.array([0.0, 2.0]),
         'pz': np.array([4.0, 2 * np.sqrt(2.0)])}
    assert np.isclose(O.energy_flux(0.1, e, _zero_fields(5)), 4.0, rtol=1e-15)


# -- lattice branches (initialization.py:72-104) -----------------------------------
@pytest.mark.parametrize('steps,fineness', [(99, 1), (739, 3), (27, 5), (99, 0.5), (40, 0.25)])
def test_plasma_grid_odd_and_fractional_fineness(steps, fineness):
    """Odd and fractional fineness put a particle on the axis: the lattice
    has spacing h / fineness, contains 0 exactly, is symmetric, stays inside
    the plasma width steps * h and has 2 ceil(floor(steps / 2) fineness) - 1
    points (one fewer than the even-fineness lattice, whose points straddle
    the axis).  A fraction 0 < f < 1 has f % 2 = f != 0, so it takes the odd
    branch too: the reference's separate fractional branch
    (initialization.py:95-97) is never reached."""
    h = 0.1
    g = O.plasma_grid(steps, h, fineness)
    k = math.ceil(steps // 2 * fineness)
    assert len(g) == 2 * k - 1
    assert np.any(g == 0.0)
    assert np.array_equal(g, -g[::-1])
    assert np.allclose(np.diff(g), h / fineness, rtol=0, atol=1e-12)
    assert np.max(np.abs(g)) < steps * h / 2


@pytest.mark.parametrize('steps,fineness', [(99, 2), (483, 2), (30, 4)])
def test_plasma_grid_even_fineness_straddles_the_axis(steps, fineness):
    h = 0.05
    g = O.plasma_grid(steps, h, fineness)
    assert len(g) == 2 * (steps // 2 * fineness)
    assert not np.any(g == 0.0)
    assert g[len(g) // 2] == pytest.approx(h / fineness / 2, abs=0)
    assert np.array_equal(g, -g[::-1])


def test_lattice_sizes_of_the_workloads():
    """97^2 electrons at N = 129, 1 ppc; 964^2 at N = 513, 4 ppc; 2213^2 per
    sort at N = 769, 9 ppc; 1988^2 at N = 1025, 4 ppc (plasma width N - 30 cells)."""
    for N, fin, n in [(129, 1, 97), (513, 2, 964), (769, 3, 2213), (1025, 2, 1988)]:
        assert len(O.plasma_grid(N - 30, 0.05, fin)) == n


def test_density_profile_scales_charge_and_mass():
    p = O.Params(N=65, h=0.1, xi_step=0.05, subtraction_coeff=1.0)
    shape = lambda x, y: np.minimum(1 + 0.1 * (x ** 2 + y ** 2), 2.0)
    a, u = O.uniform_plasma(p, density=shape), O.uniform_plasma(p)
    coef = np.minimum(1 + 0.1 * (u['x_init'] ** 2 + u['y_init'] ** 2), 2.0)
    assert np.array_equal(a['q'], u['q'] * coef) and np.array_equal(a['m'], u['m'] * coef)
    assert a['q'][0, 0] == pytest.approx(-p.h ** 2 / 4 * coef[0, 0], rel=1e-15)


# -- mobile ions (initialization.py:395-407, rhoj.py:11-20) ---------------------------
def test_mobile_ions_cancel_the_electron_charge_exactly():
    """The ion sort is the electron lattice with q -> -q, so at t = 0 the
    two columns of every current are exact negatives (the deposit sums the
    same terms with the sign flipped): net rho = 0 bit for bit, j = 0."""
    p = O.Params(N=65, h=0.1, xi_step=0.05, subtraction_coeff=1.0, ion_model='mobile',
                 ion_mass=1836.0)
    fields, particles, currents, const, _ = O.init_plasma(p)
    assert const.sorts == {'electrons': 0, 'ions': 1}
    ions, el = particles['ions'], particles['electrons']
    assert np.array_equal(ions['q'], -el['q']) and np.array_equal(ions['m'], 1836.0 * el['m'])
    cur = O.compute_rhoj(p, particles, const)
    assert np.any(cur['ro'][0] != 0)
    assert np.array_equal(cur['ro'][0] + cur['ro'][1], np.zeros((p.N, p.N)))
    for k in ('jx', 'jy', 'jz'):
        assert not np.any(cur[k])


@pytest.mark.parametrize('ion_mass', [100.0, 1836.0])
def test_mobile_ion_in_uniform_ez(ion_mass):
    """An ion of the mobile sort at rest in a uniform Ez: the two-pass
    predictor gives dp_z = q Ez dxi / (1 - v_z) with v_z taken at the
    half-step momentum q Ez dxi / 2 and gamma m = sqrt(m^2 + p^2), where
    q = +h^2 / fineness^2 and m = ion_mass h^2 / fineness^2."""
    p = O.Params(N=33, h=0.1, xi_step=0.05, subtraction_coeff=1.0, ion_model='mobile',
                 ion_mass=ion_mass)
    _, particles, _, const, _ = O.init_plasma(p)
    ions = {k: np.ascontiguousarray(v[:1, :1]) for k, v in particles['ions'].items()}
    q, m = p.h ** 2 / 4, ion_mass * p.h ** 2 / 4
    assert ions['q'][0, 0] == pytest.approx(q, rel=1e-15)
    assert ions['m'][0, 0] == pytest.approx(m, rel=1e-15)
    f = O.zero_fields(33)
    f['Ez'][:] = 40.0
    full = O.move_wo_fields(p, {'ions': ions})
    out = O.move_smart(p, f, {'ions': ions}, full)['ions']
    ph = q * 40.0 * 0.05 / 2
    vz = ph / math.sqrt(m * m + ph * ph)
    assert out['pz'][0, 0] == pytest.approx(q * 40.0 * 0.05 / (1 - vz), rel=1e-14)
    assert out['px'][0, 0] == 0 and out['x_offt'][0, 0] == 0


# -- rigid beam (beam3d/beam_calculator.py:148-181) ------------------------------------
def test_rigid_density_is_the_callable_on_the_xy_meshgrid():
    """Element [i, j] of the density is the callable at x = grid[j],
    y = grid[i] (meshgrid's 'xy' indexing), at xi = -layer * xi_step: an
    asymmetric density tells the two axes apart, and each point is
    evaluated on its own."""
    p = O.Params(N=17, h=0.25, xi_step=0.05, subtraction_coeff=1.0)
    f = lambda xp, xi, x, y: (x + 2.0) * xp.exp(-(y - 0.5) ** 2) * (1.0 + xi)
    rho = O.rigid_beam_density(p, f, 7)
    grid = [(k - 8) * 0.25 for k in range(17)]
    for i in range(17):
        for j in range(17):
            want = (grid[j] + 2.0) * math.exp(-(grid[i] - 0.5) ** 2) * (1.0 - 7 * 0.05)
            assert rho[i, j] == pytest.approx(want, rel=1e-14, abs=1e-300)
    assert not np.allclose(rho, rho.T)


def test_rigid_xi_loop_feeds_each_step_its_layer_and_the_previous_one(monkeypatch):
    """The rigid loop hands step_dxi the density of layer i as the full
    density and that of layer i - 1 as the previous one (zero before the
    first layer), for layers 0..xi_steps; no beam particle is moved."""
    p = O.Params(N=17, h=0.25, xi_step=0.1, subtraction_coeff=1.0, window_length=0.5,
                 plasma_padding_steps=3, bound_padding_steps=3)
    seen = []
    real = O.step_dxi

    def spy(pp, parts, fields, cur, const, full, prev, trace=None):
        seen.append((full.copy(), prev.copy()))
        return real(pp, parts, fields, cur, const, full, prev)

    monkeypatch.setattr(O, 'step_dxi', spy)
    f = lambda xp, xi, x, y: -0.1 * xp.exp(-(x ** 2 + 2 * y ** 2)) * (xi - 1.0)
    _, _, _, ez = O.simulate_rigid(p, f)
    assert len(seen) == 6 and len(ez) == 6
    grid = (np.arange(17) - 8) * 0.25
    for i, (full, prev) in enumerate(seen):
        xi = -i * 0.1
        want = np.array([[-0.1 * math.exp(-(grid[j] ** 2 + 2 * grid[r] ** 2)) * (xi - 1.0)
                          for j in range(17)] for r in range(17)])
        assert np.allclose(full, want, rtol=1e-14, atol=0)
        if i == 0:
            assert not np.any(prev)
        else:
            assert np.array_equal(prev, seen[i - 1][0])


# -- declustering noise filter (plasma3d/noise_filter.py) ------------------------
def _inhomogeneity_loops(d, axis):
    """calculate_inhomogeneity_numba's loops, element by element."""
    d = d.T if axis else d
    n = d.shape[0]
    out = np.zeros_like(d)
    for i in range(n):
        for j in range(n - 2):
            if d[i, j + 1] != 0 and d[i, j] != 0 and d[i, j + 2] != 0:
                out[i, j + 1] = d[i, j + 1] - (d[i, j] + d[i, j + 2]) / 2
        if d[i, 0] != 0 and d[i, 1] != 0 and d[i, 2] != 0:
            out[i, 0] = d[i, 0] - 2 * d[i, 1] + d[i, 2]
        if d[i, -1] != 0 and d[i, -2] != 0 and d[i, -3] != 0:
            out[i, -1] = d[i, -1] - 2 * d[i, -2] + d[i, -3]
    return out.T if axis else out


@pytest.mark.parametrize('axis', [0, 1])
def test_inhomogeneity_closed_forms_and_loops(axis):
    n = 9
    j = np.arange(n, dtype=float)
    ramp = np.broadcast_to(3.0 + 0.5 * j, (n, n)).copy()       # linear along the 2nd index
    square = np.broadcast_to((j + 1) ** 2, (n, n)).copy()       # (j + 1)^2
    if axis:
        ramp, square = ramp.T.copy(), square.T.copy()
    assert np.array_equal(O.inhomogeneity(ramp, axis), np.zeros((n, n)))
    want = -np.ones((n, n))
    want[:, 0] = want[:, -1] = 2.0            # d0 - 2 d1 + d2 of (j + 1)^2
    assert np.array_equal(O.inhomogeneity(square, axis), want.T if axis else want)
    rng = np.random.default_rng(3)
    d = rng.standard_normal((n, n))
    d[rng.random((n, n)) < 0.2] = 0.0          # absorbed (zero) offsets switch terms off
    assert np.array_equal(O.inhomogeneity(d, axis), _inhomogeneity_loops(d, axis))


def test_gaussian_smoothing_of_constants_and_fourier_modes():
    n, sigma = 16, 1.7
    c = np.full((n, n), 2.5)
    assert np.allclose(O.gaussian_smooth(c, sigma), c, rtol=1e-15, atol=0)
    i = np.arange(n)
    for k, l in ((1, 0), (3, 2), (8, 5)):
        mode = np.cos(2 * np.pi * k * i / n)[:, None] * np.cos(2 * np.pi * l * i / n)[None, :]
        damp = np.exp(-2 * np.pi ** 2 * sigma ** 2 * ((k / n) ** 2 + (l / n) ** 2))
        assert np.allclose(O.gaussian_smooth(mode, sigma), damp * mode, rtol=0, atol=1e-15)


def _filter_params(damping=0.0):
    p = O.Params(N=65, h=0.1, xi_step=0.05, subtraction_coeff=1.0)
    p.declustering = O.Declustering(averaging=3, force=0.3, damping=damping, limit=1e-3)
    return p


def _lattice(n, rng, x_offt, y_offt):
    return dict(x_offt=x_offt, y_offt=y_offt, px=1e-3 * rng.standard_normal((n, n)),
                py=1e-3 * rng.standard_normal((n, n)))


@pytest.mark.parametrize('shift', [0.0, 2.5e-4])
def test_noise_filter_leaves_an_undisturbed_or_uniformly_shifted_lattice_alone(shift):
    n, rng = 24, np.random.default_rng(4)
    el = _lattice(n, rng, np.full((n, n), shift), np.full((n, n), -shift))
    px0, py0 = el['px'].copy(), el['py'].copy()
    out = O.noise_filter(_filter_params(damping=0.7), el, {})
    assert np.array_equal(out['dx_chaotic'], np.zeros((n, n)))
    assert np.array_equal(out['dy_chaotic'], np.zeros((n, n)))
    assert np.array_equal(el['px'], px0) and np.array_equal(el['py'], py0)


def test_noise_filter_single_displaced_particle():
    """One particle displaced by delta in a uniformly shifted lattice: its
    inhomogeneity is delta, its four neighbours' -delta/4; the smoothing is
    the wrapped sampled Gaussian (sigma 3 cells, computed directly); the
    kicks follow by the stencil written out per particle.  Total px is
    unchanged (the stencil conserves momentum; no damping)."""
    n, c, delta, rng = 32, 1e-4, 2e-4, np.random.default_rng(5)
    p = _filter_params()
    x = np.full((n, n), c)
    i0, j0 = 11, 17
    x[i0, j0] += delta
    el = _lattice(n, rng, x, np.full((n, n), c))
    px0, py0 = el['px'].copy(), el['py'].copy()
    out = O.noise_filter(p, el, {})
    inh = np.zeros((n, n))
    inh[i0, j0] = delta
    for a, b in ((i0 - 1, j0), (i0 + 1, j0), (i0, j0 - 1), (i0, j0 + 1)):
        inh[a, b] = -delta / 4
    m = np.arange(n)
    g = sum(np.exp(-((m + k * n) ** 2) / (2 * 9.0)) for k in range(-3, 4)) / math.sqrt(2 * math.pi * 9)
    G = np.array([[g[(a - b) % n] for b in range(n)] for a in range(n)])   # circulant
    smooth = G @ inh @ G.T
    chaotic = inh - smooth
    assert np.allclose(out['dx_chaotic'], chaotic, rtol=0, atol=1e-17)
    assert np.array_equal(out['dy_chaotic'], np.zeros((n, n)))
    lim2 = p.declustering.limit ** 2
    dl = np.where(chaotic ** 2 < lim2, 1 - chaotic ** 2 / lim2, 0)
    f = -p.declustering.force * chaotic * dl
    want = px0.copy()
    xi = p.xi_step
    for a in range(n):
        for b in range(n):
            k = 0.0
            if 1 <= a <= n - 2:
                k += xi * f[a, b] / 2
            if a >= 2:
                k -= xi * f[a - 1, b] / 4
            if a <= n - 3:
                k -= xi * f[a + 1, b] / 4
            if 1 <= b <= n - 2:
                k += xi * f[a, b] / 2
            if b >= 2:
                k -= xi * f[a, b - 1] / 4
            if b <= n - 3:
                k -= xi * f[a, b + 1] / 4
            want[a, b] += k
    assert np.allclose(el['px'], want, rtol=0, atol=1e-18)
    assert np.array_equal(el['py'], py0)
    kick = el['px'] - px0
    assert np.max(np.abs(kick)) > 1e-9
    assert abs(kick.sum()) < 1e-8 * np.abs(kick).sum()   # rounding of px + kick only


def test_noise_filter_damping_acts_on_the_change_of_chaotic_displacement():
    n, rng = 20, np.random.default_rng(6)
    x = 1e-4 * (1 + 0.1 * rng.standard_normal((n, n)))
    y = 1e-4 * (1 + 0.1 * rng.standard_normal((n, n)))
    a = _lattice(n, rng, x, y)
    b = {k: v.copy() for k, v in a.items()}
    first = O.noise_filter(_filter_params(damping=0.0), a, {})
    # the previous chaotic displacements equal to the new ones: damping adds nothing
    O.noise_filter(_filter_params(damping=0.9), b, {'dx_chaotic': first['dx_chaotic'],
                                                    'dy_chaotic': first['dy_chaotic']})
    assert np.array_equal(a['px'], b['px']) and np.array_equal(a['py'], b['py'])


def test_declustering_damping_is_read_from_the_misspelt_key():
    c = Config({'geometry': '3d', 'declustering-enabled': True})
    assert c.getfloat('declustering-damping') == 0.1       # the default's spelling
    assert O.Declustering.from_config(c).damping == 0.0   # what noise_filter.py:89 reads
    c.set('damping-declustering', 0.25)
    assert O.Declustering.from_config(c).damping == 0.25


def test_gaussian_taps_are_scipys_fourier_gaussian():
    """The product's convolution taps against scipy.ndimage.fourier_gaussian
    (the reference's smoothing) on random lattices, for widths below and
    above the point where the taps are cut."""
    from lcodepy_amd.plasma3d.noise_filter import gaussian_taps
    rng = np.random.default_rng(7)
    for n, sigma in ((964, 3), (37, 5.5), (500, 1), (500, 2.9), (60, 40)):
        taps, t0 = gaussian_taps(n, sigma)
        a = rng.standard_normal((n, 3))
        idx = (np.arange(n)[:, None] - t0 - np.arange(len(taps))[None, :]) % n
        got = np.stack([(a[idx, q] * taps[None, :]).sum(axis=1) for q in range(3)], axis=1)
        from scipy.ndimage import fourier_gaussian
        want = np.fft.ifft(fourier_gaussian(np.fft.fft(a, axis=0), (sigma, 0)), axis=0).real
        assert np.max(np.abs(got - want)) < 2e-15 * np.max(np.abs(want)), (n, sigma)
