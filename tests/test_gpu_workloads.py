"""The benchmark workloads on the GPU against the oracle.

* 129 x 129 window, 1 particle per cell (97^2 electrons on the lattice with a
  particle on each axis), background ions, a rigid Gaussian driver given as a
  density callable, 201 xi steps: every xi step's on-axis Ez and the final
  fields, currents and particle state against the oracle's rigid-beam loop.
* 513 x 513 window, 4 particles per cell (964^2 electrons with seeded thermal
  momenta), a seeded 1e6-particle driver (bench.py's workload): the float64
  form against the oracle over the first 20 xi steps (everything the oracle
  can afford at this size), and the FP32-bulk form against the float64 GPU
  form over the whole 1001-step window (on-axis Ez, the size-independent
  property the FP32-bulk target is stated on).
* 1025 x 1025 window, 4 ppc in a parabolic channel n = 1 + 0.1 min(r, 10)^2,
  an evolving 4e6-particle bunch; and 769 x 769, 9 ppc, mobile ions (mass
  1836), a driver and a witness bunch (1.5e6 particles): both forms against
  the oracle over the first 8 xi steps.

Errors are max |gpu - ref| / max |ref| per array.  The float64 bounds sit
about two orders of magnitude above the measured errors (which come from
summation order and fused multiply-adds); the FP32-bulk bound is the target
1e-4 on on-axis Ez."""
import math

import numpy as np
import pytest

from oracle import lcode_oracle as O

pytestmark = pytest.mark.gpu
torch = pytest.importorskip('torch')

FIELDS = ('Ex', 'Ey', 'Ez', 'Bx', 'By', 'Bz', 'Phi')
CURRENTS = ('ro', 'jx', 'jy', 'jz')
PARTICLE = ('x_offt', 'y_offt', 'px', 'py', 'pz', 'q')


def relerr(a, b):
    scale = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / scale) if scale > 0 else float(np.max(np.abs(a)))


def host(t):
    return t.detach().double().cpu().numpy()


def compare_state(gpu_state, oracle_state):
    """{name: error} over fields, currents (summed by column as well as per
    column) and every sort's particle arrays."""
    parts, fields, cur = gpu_state
    parts_o, fields_o, cur_o = oracle_state
    err = {}
    for k in FIELDS:
        err[k] = relerr(host(getattr(fields, k)), fields_o[k])
    for k in CURRENTS:
        err[k] = relerr(host(getattr(cur, k)), cur_o[k])
    for sort, arrs in parts_o.items():
        for k in PARTICLE:
            err[f'{sort}.{k}'] = relerr(host(getattr(parts[sort], k)), arrs[k])
    return err


# -- 129 x 129, 1 ppc, rigid driver ------------------------------------------------
C1_XI0 = -3.0


def c1_density(xp, xi, x, y):
    """Rigid electron driver: rho = -0.05 exp(-r^2 / 2) exp(-(xi - xi0)^2 / 2)
    for |xi - xi0| <= 3, xi0 = -3 (sigma_r = sigma_xi = 1)."""
    if abs(xi - C1_XI0) > 3.0:
        return 0.0 * x
    return -0.05 * xp.exp(-(x ** 2 + y ** 2) / 2) * math.exp(-(xi - C1_XI0) ** 2 / 2)


def c1_config():
    return {'geometry': '3d', 'processing-unit-type': 'mi355x', 'precision': 'fp64',
            'window-width': 6.45, 'transverse-step': 0.1,
            'window-length': 10.0, 'xi-step': 0.05, 'time-limit': 1.0, 'time-step': 1.0,
            'ion-model': 'background', 'plasma-particles-per-cell': 1,
            'rigid-beam': True, 'field-solver-subtraction-coefficient': 1.0}


def test_rigid_driver_129_1ppc_matches_oracle():
    from lcodepy_amd.config import Config
    from lcodepy_amd.simulation import Simulation
    cfg = Config(c1_config())
    cfg._adjust_config_values_3d()
    p = O.Params.from_config(cfg)
    assert p.N == 129 and p.fineness == 1 and p.xi_steps == 200
    print('rigid driver N=129 1ppc: oracle', flush=True)
    parts_o, fields_o, cur_o, ez_o = O.simulate_rigid(p, c1_density)
    assert parts_o['electrons']['x_init'].size == 97 ** 2
    print('rigid driver N=129 1ppc: gpu', flush=True)
    sim = Simulation(config=c1_config(), beam_parameters=c1_density, runas_filename='')
    sim.push_solver.verbose = False
    sim.step()
    ez_g = np.array([float(v) for v in sim.push_solver.ez_axis])
    assert ez_g.shape == ez_o.shape == (201,)
    err = compare_state(sim.push_solver._plasmastate, (parts_o, fields_o, cur_o))
    err['ez_axis'] = relerr(ez_g, ez_o)
    print('rigid driver N=129 1ppc, 201 xi steps:', {k: f'{v:.1e}' for k, v in err.items()})
    assert np.max(np.abs(ez_o)) > 1e-3           # a wake was driven
    assert err['ez_axis'] < 1e-11
    for k, v in err.items():
        assert v < 1e-10, (k, v)


# -- 513 x 513, 4 ppc, 1e6-particle driver (bench.py's workload) ----------------------
def _c2_oracle_first_steps(cfg, n_steps, n_beam):
    import bench
    p = O.Params.from_config(cfg)
    fields, particles, currents, const, xi0 = O.init_plasma(p)
    e = particles['electrons']
    e.update(bench.thermal_momenta(bench.electron_mass(cfg), e['m'].shape))
    source = O.beam_sorted(O.beam_from_records(bench.driver_beam(n_beam, p.xi_step)))
    parts, fields, cur, _, ez = O.step_dt(p, fields, particles, currents, const, xi0, source,
                                          n_steps=n_steps)
    return parts, fields, cur, ez


@pytest.mark.parametrize('precision,bound', [('fp64', 1e-11), ('fp32-bulk', 1e-4)])
def test_513_4ppc_first_steps_match_oracle(precision, bound):
    """The benchmark workload, first 20 xi steps (the driver's head enters
    the window): every field, current and particle array against the
    float64 oracle on the same seeded inputs."""
    import bench
    n_steps, n_beam = 20, 1_000_000
    print(f'513x513 4ppc {precision}: gpu', flush=True)
    run = bench.Run(precision, n_beam)
    run.steps(n_steps)
    torch.cuda.synchronize()
    run.solver.plasma_solver.ctx.raise_on_errors()
    ez_g = run.ez()
    state_g = run.solver._plasmastate
    print(f'513x513 4ppc {precision}: oracle', flush=True)
    parts_o, fields_o, cur_o, ez_o = _c2_oracle_first_steps(run.cfg, n_steps, n_beam)
    err = compare_state(state_g, (parts_o, fields_o, cur_o))
    err['ez_axis'] = relerr(ez_g, ez_o)
    print(f'513x513 4ppc {precision}, {n_steps} xi steps vs oracle:',
          {k: f'{v:.1e}' for k, v in err.items()})
    assert ez_g.shape == ez_o.shape == (n_steps,)
    assert np.max(np.abs(ez_o)) > 0
    for k, v in err.items():
        # Phi solves rho - jz of the plasma alone (fields.py:142-159).  Over
        # the first steps the electron and background-ion densities (each of
        # magnitude 1) still cancel to a few 1e-3, so the float32 rounding of
        # the electron density (6e-8 absolute) is amplified by the Poisson
        # solve relative to the small Phi: measured 7.1e-3 of max |Phi|.  Phi
        # is an output only; nothing in the step reads it.
        lim = 2e-2 if (k == 'Phi' and precision == 'fp32-bulk') else bound
        assert v < lim, (k, v)


def test_513_4ppc_fp32_bulk_tracks_fp64_over_the_window():
    """The whole 1001-step window in both forms on the GPU: on-axis Ez of
    the FP32-bulk form within 1e-4 of the float64 form's at every xi step
    (normalised by max |Ez|); the final Ex, rho and electron offsets are
    reported alongside."""
    import bench
    runs = {}
    for form in ('fp64', 'fp32-bulk'):
        print(f'513x513 4ppc {form}: whole window', flush=True)
        run = bench.Run(form, 1_000_000)
        run.steps(bench.WINDOW_STEPS)
        torch.cuda.synchronize()
        run.solver.plasma_solver.ctx.raise_on_errors()
        parts, fields, cur = run.solver._plasmastate
        runs[form] = dict(ez=run.ez(), Ex=host(fields.Ex), ro=host(cur.ro).sum(axis=0),
                          x_offt=host(parts['electrons'].x_offt))
        del run, parts, fields, cur
        torch.cuda.empty_cache()
    a, b = runs['fp32-bulk'], runs['fp64']
    assert a['ez'].shape == b['ez'].shape == (bench.WINDOW_STEPS,)
    err = {k: relerr(a[k], b[k]) for k in ('ez', 'Ex', 'ro', 'x_offt')}
    print('513x513 4ppc fp32-bulk vs fp64 over 1001 xi steps:',
          {k: f'{v:.1e}' for k, v in err.items()})
    assert err['ez'] < 1e-4


# -- 1025 x 1025 parabolic channel and 769 x 769 mobile ions, two bunches ------------
def gaussian_bunch(n, xi_step, seed, sig_r, sig_xi, n_b, xi0, gamma, emittance, first_id=1):
    """Seeded Gaussian electron bunch as particle_dtype3d records, |xi - xi0|
    within 3 sigma_xi, transverse momenta of spread emittance / sigma_r
    (normalised emittance), gamma spread 1%."""
    from lcodepy_amd.beam_generator import particle_dtype3d
    rng = np.random.RandomState(seed)
    xi = np.empty(0)
    while xi.size < n:
        d = rng.normal(0.0, sig_xi, 2 * n)
        xi = np.concatenate([xi, xi0 + d[np.abs(d) <= 3 * sig_xi]])
    b = np.zeros(n, dtype=particle_dtype3d)
    b['xi'] = xi[:n]
    b['x'] = rng.normal(0.0, sig_r, n)
    b['y'] = rng.normal(0.0, sig_r, n)
    b['px'] = rng.normal(0.0, emittance / sig_r, n)
    b['py'] = rng.normal(0.0, emittance / sig_r, n)
    b['pz'] = rng.normal(gamma, 0.01 * gamma, n)
    b['q_m'] = -1.0
    b['q_norm'] = -n_b * 2 * math.pi * sig_r ** 2 * math.sqrt(2 * math.pi) * sig_xi / (n * xi_step)
    b['id'] = np.arange(first_id, first_id + n)
    return b


def c3_density(current_time, x, y):
    """Parabolic channel n(r) = n0 (1 + 0.1 r^2), r capped at 10."""
    return 1.0 + 0.1 * np.minimum(x ** 2 + y ** 2, 100.0)


def c3_setup(precision):
    """1025 x 1025 (step 0.025), 4 ppc in the parabolic channel, background
    ions, an evolving 4e6-particle bunch (n_b 0.3, sigma_r 1, sigma_xi 1.5,
    gamma 2000 +- 1%, normalised emittance 1), xi_step 0.02.  The field
    solver's subtraction coefficient is the channel's peak density, 11: with
    the default (1 + xi_step) the Helmholtz-type solves of the predictor-
    corrector do not converge where the plasma is eleven times denser, and
    the oracle's on-axis Ez grows about 25-fold per xi step (497 after
    eight)."""
    cfg = {'geometry': '3d', 'processing-unit-type': 'mi355x', 'precision': precision,
           'window-width': 12.8, 'transverse-step': 0.025, 'window-length': 40.0,
           'xi-step': 0.02, 'time-limit': 1.0, 'time-step': 1.0, 'ion-model': 'background',
           'plasma-particles-per-cell': 4, 'plasma-shape': c3_density,
           'field-solver-subtraction-coefficient': 11.0}
    beam = gaussian_bunch(4_000_000, 0.02, 3, 1.0, 1.5, 0.3, -4.5, 2000.0, 1.0)
    return cfg, beam[np.argsort(-beam['xi'], kind='stable')], lambda x, y: c3_density(0.0, x, y)


def c5_setup(precision):
    """769 x 769 (step 0.04), 9 ppc, mobile ions of mass 1836, a driver
    (n_b 0.5, sigma_r 0.7, sigma_xi 1; 1e6 particles) and a witness 6 behind
    it (n_b 0.2, sigma_r 0.3, sigma_xi 0.3; 5e5 particles), xi_step 0.02."""
    cfg = {'geometry': '3d', 'processing-unit-type': 'mi355x', 'precision': precision,
           'window-width': 15.36, 'transverse-step': 0.04, 'window-length': 30.0,
           'xi-step': 0.02, 'time-limit': 1.0, 'time-step': 1.0, 'ion-model': 'mobile',
           'ion-mass': 1836.0, 'plasma-particles-per-cell': 9}
    drv = gaussian_bunch(1_000_000, 0.02, 5, 0.7, 1.0, 0.5, -3.0, 1000.0, 1.0)
    wit = gaussian_bunch(500_000, 0.02, 6, 0.3, 0.3, 0.2, -9.0, 1000.0, 1.0, first_id=1_000_001)
    beam = np.concatenate([drv, wit])
    return cfg, beam[np.argsort(-beam['xi'], kind='stable')], None


def _gpu_first_steps(cfg_dict, beam, n_steps):
    from lcodepy_amd.beam3d.beam_io import BeamDrain, BeamSource
    from lcodepy_amd.config import Config
    from lcodepy_amd.plasma3d.initialization import init_plasma
    from lcodepy_amd.push_solver import PusherAndSolver3D
    cfg = Config(cfg_dict)
    cfg._adjust_config_values_3d()
    fields, particles, currents, const, xi0 = init_plasma(cfg)
    solver = PusherAndSolver3D(cfg)
    solver.verbose = False
    it = solver.iter_dxi(fields, particles, currents, const, xi0, BeamSource(cfg, beam),
                         BeamDrain(cfg), 1.0)
    for _ in range(n_steps):
        next(it)
    torch.cuda.synchronize()
    solver.plasma_solver.ctx.raise_on_errors()
    ez = np.array([float(v) for v in solver.ez_axis])
    return cfg, solver._plasmastate, ez


_ORACLE_RUNS = {}


def _oracle_first_steps(name, cfg, beam, density, n_steps):
    if name not in _ORACLE_RUNS:
        p = O.Params.from_config(cfg)
        el = O.uniform_plasma(p, density=density)
        fields, particles, currents, const, xi0 = O.init_plasma(p, electrons=el)
        source = O.beam_sorted(O.beam_from_records(beam))
        parts, fields, cur, _, ez = O.step_dt(p, fields, particles, currents, const, xi0,
                                              source, n_steps=n_steps)
        _ORACLE_RUNS[name] = (parts, fields, cur, np.asarray(ez))
    return _ORACLE_RUNS[name]


# c3 float64: the channel's plasma, up to 11 times denser, carries 11 times the
# charge and current into the same fields, and the summation-order rounding of
# the deposit grows with it (measured up to 3.6e-11 in jz and pz after 8 steps).
@pytest.mark.parametrize('name,N,precision,bound', [
    ('c3', 1025, 'fp64', 1e-10), ('c3', 1025, 'fp32-bulk', 1e-4),
    ('c5', 769, 'fp64', 1e-11), ('c5', 769, 'fp32-bulk', 1e-4)])
def test_large_workloads_first_steps_match_oracle(name, N, precision, bound):
    """The 1025 x 1025 parabolic-channel workload (4e6-particle evolving
    bunch) and the 769 x 769 workload with 9 ppc, mobile ions and a
    driver-witness pair (1.5e6 beam particles), first 8 xi steps (the
    bunches' heads entering): every field, current and particle array of
    every sort against the float64 oracle on the same seeded inputs."""
    n_steps = 8
    cfg_dict, beam, density = (c3_setup if name == 'c3' else c5_setup)(precision)
    print(f'{name} {N} {precision}: gpu', flush=True)
    cfg, state_g, ez_g = _gpu_first_steps(cfg_dict, beam, n_steps)
    assert cfg.getint('window-width-steps') == N
    print(f'{name} {N}: oracle', flush=True)
    parts_o, fields_o, cur_o, ez_o = _oracle_first_steps(name, cfg, beam, density, n_steps)
    err = compare_state(state_g, (parts_o, fields_o, cur_o))
    err['ez_axis'] = relerr(ez_g, ez_o)
    # The bunches are round and on the axis, so Bz is zero up to rounding
    # (max |Bz| / max |Bx, By| = 5e-4 and 3e-7 here): its error is measured
    # against the transverse magnetic field.
    fields_g = state_g[1]
    b_scale = max(np.max(np.abs(fields_o[k])) for k in ('Bx', 'By', 'Bz'))
    err['Bz'] = float(np.max(np.abs(host(fields_g.Bz) - fields_o['Bz'])) / b_scale)
    print(f'{name} {N} {precision}, {n_steps} xi steps vs oracle:',
          {k: f'{v:.1e}' for k, v in err.items()})
    assert ez_g.shape == ez_o.shape == (n_steps,)
    assert np.max(np.abs(ez_o)) > 0
    if name == 'c5':
        assert 'ions' in parts_o and 'ions' in state_g[0]
    for k, v in err.items():
        lim = bound
        if k == 'Phi':
            # Phi solves rho - jz of the plasma (fields.py:142-159): electron
            # and ion densities of up to 11 (the channel's edge) cancel to
            # about 1e-5 over the first steps, so the rounding of the density
            # (1e-16 relative in float64, 6e-8 in float32) is amplified about
            # 1e6-fold relative to max |Phi| (measured 2.6e-9 and 4.8e-2 in the
            # channel).  Phi is an output only; nothing in the step reads it.
            lim = 1e-1 if precision == 'fp32-bulk' else 1e-8
        assert v < lim, (k, v)
