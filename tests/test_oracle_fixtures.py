"""The oracle against the reference's own golden end states.

Same checks and tolerance as tests/integration/test_3D_part.py:32-82 of the
reference: every 10th particle/grid point, np.allclose(rtol=5e-8,
atol=1e-125), currents summed over sorts."""
import os

import numpy as np
import pytest

from lcodepy_amd.beam_generator import generate_beam
from oracle import lcode_oracle as O
from tests.cases import adjusted, fixture_beam, fixture_config

RTOL = 5e-8


def _check(golden, particles, fields, currents):
    el = particles['electrons']
    for k in ('x_offt', 'y_offt', 'px', 'py', 'pz'):
        assert np.allclose(el[k][::10, ::10], golden[k], rtol=RTOL, atol=1e-125), k
    for k in ('Ex', 'Ey', 'Ez', 'Bx', 'By', 'Bz'):
        assert np.allclose(fields[k][::10, ::10], golden[k], rtol=RTOL, atol=1e-125), k
    for k in ('ro', 'jx', 'jy', 'jz'):
        assert np.allclose(currents[k][:, ::10, ::10].sum(axis=0), golden[k],
                           rtol=RTOL, atol=1e-125), k


@pytest.mark.slow
@pytest.mark.parametrize('which,n_time_steps', [('test1', 1), ('beam_evol', 5)])
def test_oracle_matches_reference_fixture(which, n_time_steps, golden_dir):
    c = adjusted(fixture_config(which))
    p = O.Params.from_config(c)
    assert p.N == 321 and p.fineness == 2 and p.subtraction_coeff == 1.0
    beam = generate_beam(c, fixture_beam(which))
    particles, fields, currents, _, _ = O.simulate(p, beam, n_time_steps)
    golden = np.load(os.path.join(golden_dir, f'3D_{which}.npz'))
    _check(golden, particles, fields, currents)
