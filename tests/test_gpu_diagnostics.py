"""The f(xi) probe diagnostic (FXiDiag) on the GPU against the values the
reference's host-side strategy would read from the same states
(diagnostics/FXi/Strategy3D.py:59-95): every quantity, at several probe
points, every xi step, from a run with mobile ions and an evolving beam.
Grid values and maxima are copies, so they must agree exactly; the energy
flux Sf is checked against the oracle's restatement of
calculate_energy_fluxes (1e-12 of the largest flux of the run)."""
import numpy as np
import pytest

from oracle import lcode_oracle as O

pytestmark = pytest.mark.gpu
torch = pytest.importorskip('torch')

PROBES = [[0.0, 0.3, -0.5, 1.0], [0.0, 0.1, 0.2, -0.7]]


class HostProbe:
    """What Strategy3D does, on host copies of each xi step's state."""

    def __init__(self, N, h):
        self.ix = (N // 2 + np.round(np.array(PROBES[0]) / h)).astype(int)
        self.iy = (N // 2 + np.round(np.array(PROBES[1]) / h)).astype(int)
        self.h = h
        self.rows = {k: [] for k in ('Ex', 'Ey', 'Ez', 'Bx', 'By', 'Bz', 'ne', 'ni',
                                     'rho_beam', 'Phi', 'Sf', 'dx_chaotic', 'dx_chaotic_perp')}

    def pull_config(self, config):
        pass

    def after_step_dxi(self, t, xi, particles, fields, currents, rho_beam):
        g = lambda a: a.double().cpu().numpy()[self.ix, self.iy]
        for k in ('Ex', 'Ey', 'Ez', 'Bx', 'By', 'Bz', 'Phi'):
            self.rows[k].append(g(getattr(fields, k)))
        self.rows['ne'].append(g(currents.ro[0]))
        self.rows['ni'].append(g(currents.ro[1]))
        self.rows['rho_beam'].append(g(rho_beam))
        el = {k: getattr(particles['electrons'], k).double().cpu().numpy()
              for k in ('m', 'px', 'py', 'pz')}
        f = {k: getattr(fields, k).cpu().numpy() for k in ('Ex', 'Ey', 'Ez', 'Bx', 'By', 'Bz')}
        self.rows['Sf'].append(O.energy_flux(self.h, el, f))
        dx = particles['electrons'].dx_chaotic.cpu().numpy()
        a1 = int(dx.shape[0] / 6)
        self.rows['dx_chaotic'].append(np.max(np.abs(dx)))
        self.rows['dx_chaotic_perp'].append(np.max(np.abs(dx[a1:5 * a1, a1:5 * a1])))

    def dump(self, *a):
        pass


@pytest.mark.parametrize('precision', ['fp64', 'fp32-bulk'])
def test_fxi_probes_match_host_reads(precision, tmp_path, monkeypatch):
    from lcodepy_amd.diagnostics import FXiDiag, FXiType
    from lcodepy_amd.simulation import Simulation
    monkeypatch.chdir(tmp_path)
    cfg = {'geometry': '3d', 'processing-unit-type': 'mi355x', 'precision': precision,
           'window-width': 3.2, 'transverse-step': 0.1, 'window-length': 1.5, 'xi-step': 0.05,
           'time-limit': 1.0, 'time-step': 1.0, 'ion-model': 'mobile', 'ion-mass': 100.0,
           'plasma-particles-per-cell': 4}
    flags = (FXiType.EB | FXiType.Phi | FXiType.n | FXiType.rho_beam | FXiType.Sf |
             FXiType.dx_chaotic | FXiType.dx_chaotic_perp)
    diag = FXiDiag(f_xi=flags, probe_lines=PROBES, output_period=1)
    host = HostProbe(65, 0.1)
    print(f'fxi {precision}', flush=True)
    sim = Simulation(config=cfg, beam_parameters={'current': 0.5, 'particles_in_layer': 300},
                     diagnostics=[host, diag], runas_filename='')
    sim.push_solver.verbose = False
    sim.step()
    out = np.load(tmp_path / 'diagnostics' / f'f_xi_{1.0:08.2f}.npz')
    n = len(host.rows['Ex'])
    assert n == 31 and out['xi'].shape == (n,)
    assert np.allclose(out['xi'], -0.05 * np.arange(n))
    for k, rows in host.rows.items():
        want = np.array(rows)
        got = out[k]
        assert got.shape == want.shape, k
        if k == 'Sf':
            # relative to the largest flux of the run: the first steps' flux
            # (1e-8) is a sum of kinetic terms sqrt(m^2 + p^2) - m whose
            # rounding (a few 1e-16 here) does not scale with it
            assert np.max(np.abs(got - want)) <= 1e-12 * np.max(np.abs(want)), k
        else:
            assert np.array_equal(got, want), k
    assert np.max(np.abs(out['Ez'])) > 0 and np.max(np.abs(out['rho_beam'])) > 0
