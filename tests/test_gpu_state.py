"""Checkpoint and resume on the GPU (SaveRunState, load_plasma, beamfiles).

* A run with a rigid driver that is stopped at xi = -4 (a SaveRunState
  snapshot, after the driver has passed) and resumed from that file through
  ``Simulation.external_plasmastate`` ends in the same state, bit for bit,
  as the run that went through: the snapshot holds everything the xi loop
  carries (particles, fields, currents, xi_plasma_layer), and every kernel
  is deterministic.
* The end-of-time-step beamfile and plasma state of a run with an evolving
  beam read back (BeamParticles.load, load_plasma) to exactly the drained
  beam and the final plasma."""
import math
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip('torch')


def _density(xp, xi, x, y):
    """Rigid driver over -3 <= xi <= 0."""
    if xi < -3.0 or xi > 0.0:
        return 0.0 * x
    return -0.1 * xp.exp(-(x ** 2 + y ** 2) / 2) * math.exp(-(xi + 1.5) ** 2 / 0.5)


def _rigid_config(precision):
    return {'geometry': '3d', 'processing-unit-type': 'mi355x', 'precision': precision,
            'window-width': 3.2, 'transverse-step': 0.1, 'window-length': 6.0,
            'xi-step': 0.05, 'time-limit': 1.0, 'time-step': 1.0,
            'ion-model': 'background', 'plasma-particles-per-cell': 4, 'rigid-beam': True}


def _host_state(state):
    parts, fields, cur = state
    out = {f'f.{k}': getattr(fields, k).double().cpu().numpy()
           for k in ('Ex', 'Ey', 'Ez', 'Bx', 'By', 'Bz', 'Phi')}
    out.update({f'c.{k}': getattr(cur, k).double().cpu().numpy()
                for k in ('ro', 'jx', 'jy', 'jz')})
    for sort, a in parts.items():
        out.update({f'{sort}.{k}': getattr(a, k).double().cpu().numpy()
                    for k in ('x_offt', 'y_offt', 'px', 'py', 'pz', 'q')})
    return out


@pytest.mark.parametrize('precision', ['fp64', 'fp32-bulk'])
def test_resume_from_a_snapshot_reproduces_the_run(precision, tmp_path, monkeypatch):
    from lcodepy_amd.diagnostics import SaveRunState
    from lcodepy_amd.simulation import Simulation
    monkeypatch.chdir(tmp_path)
    print(f'resume {precision}: whole run', flush=True)
    save = SaveRunState(output_period=1, saving_xi_period=4.0, save_plasma=True,
                        unique_directory_name='_a')
    sim = Simulation(config=_rigid_config(precision), beam_parameters=_density,
                     diagnostics=[save], runas_filename='')
    sim.push_solver.verbose = False
    sim.step()
    whole = _host_state(sim.push_solver._plasmastate)
    ez_whole = np.array([float(v) for v in sim.push_solver.ez_axis])
    snap = tmp_path / 'snapshots_a' / f'plasmastate_{1.0:08.2f}_{-4.0:+09.2f}.npz'
    assert snap.exists(), sorted(os.listdir(tmp_path / 'snapshots_a'))

    print(f'resume {precision}: from xi = -4', flush=True)
    sim2 = Simulation(config=_rigid_config(precision), beam_parameters=_density,
                      runas_filename='')
    sim2.push_solver.verbose = False
    sim2.external_plasmastate = True
    sim2.path_to_plasmastate = str(snap)
    sim2.step()
    resumed = _host_state(sim2.push_solver._plasmastate)
    ez_resumed = np.array([float(v) for v in sim2.push_solver.ez_axis])
    assert len(ez_resumed) == 40 and np.array_equal(ez_resumed, ez_whole[-40:])
    for k, v in whole.items():
        assert np.array_equal(resumed[k], v), k


def test_end_of_step_beamfile_and_plasma_state_read_back(tmp_path, monkeypatch):
    from lcodepy_amd.beam3d.data import BeamParticles
    from lcodepy_amd.config import Config
    from lcodepy_amd.diagnostics import SaveRunState
    from lcodepy_amd.plasma3d.initialization import load_plasma
    from lcodepy_amd.simulation import Simulation
    monkeypatch.chdir(tmp_path)
    cfg = {'geometry': '3d', 'processing-unit-type': 'mi355x', 'window-width': 3.2,
           'transverse-step': 0.1, 'window-length': 2.0, 'xi-step': 0.05,
           'time-limit': 1.0, 'time-step': 1.0, 'ion-model': 'mobile', 'ion-mass': 100.0,
           'plasma-particles-per-cell': 4}
    print('beamfile + plasma state: run', flush=True)
    save = SaveRunState(output_period=1, save_beam=True, save_plasma=True)
    sim = Simulation(config=cfg, beam_parameters={'current': 0.5, 'particles_in_layer': 300},
                     diagnostics=[save], runas_filename='')
    sim.push_solver.verbose = False
    sim.step()
    drained = sim.beam_drain.beam_slice()
    beamfile = tmp_path / 'run_states' / f'beamfile_{1.0:08.2f}.npz'
    plasmafile = tmp_path / 'run_states' / f'plasmastate_{1.0:08.2f}.npz'
    assert beamfile.exists() and plasmafile.exists()
    b = BeamParticles()
    b.load(beamfile)
    assert drained.size > 0 and b.size == drained.size
    for k in ('xi', 'x', 'y', 'px', 'py', 'pz', 'q_m', 'q_norm', 'id'):
        assert torch.equal(getattr(b, k), getattr(drained, k)), k
    c = Config(cfg)
    c._adjust_config_values_3d()
    fields, particles, currents, _, xi = load_plasma(c, str(plasmafile))
    final = _host_state(sim.push_solver._plasmastate)
    assert xi == pytest.approx(-2.0, abs=1e-12)
    assert set(particles) == {'electrons', 'ions'}
    for k, v in _host_state((particles, fields, currents)).items():
        assert np.array_equal(v, final[k]), k
