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
