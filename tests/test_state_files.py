"""On-disk state formats (host side, no GPU): the plasma state that
SaveRunState writes and load_plasma reads, in both of the reference's
layouts -- per-sort structured records (plasma3d/initialization.py:417-460)
and SaveRunState's flat keys (diagnostics/diagnostics_3d.py:666-689) -- and
the save-period rule."""
from types import SimpleNamespace

import numpy as np
import pytest

from lcodepy_amd.diagnostics import SaveRunState, absremainder
from lcodepy_amd.plasma3d.initialization import (STATE_CURRENTS, STATE_FIELDS, STATE_PARTICLE,
                                                 read_plasmastate, write_plasmastate)


def _state(rng, N=9, n=5, sorts=('electrons', 'ions')):
    parts = {s: SimpleNamespace(**{k: rng.normal(size=(n, n)) for k in STATE_PARTICLE})
             for s in sorts}
    fields = SimpleNamespace(**{k: rng.normal(size=(N, N)) for k in STATE_FIELDS})
    cur = SimpleNamespace(**{k: rng.normal(size=(2, N, N)) for k in STATE_CURRENTS})
    return parts, fields, cur


def test_plasmastate_round_trip_every_sort(tmp_path):
    rng = np.random.default_rng(1)
    parts, fields, cur = _state(rng)
    ro0 = rng.normal(size=(9, 9))
    path = tmp_path / 'state.npz'
    write_plasmastate(path, parts, fields, cur, -1.25, ro0)
    f, p, c, rho, xi = read_plasmastate(path, {'electrons': 0, 'ions': 1})
    assert xi == -1.25
    assert np.array_equal(rho, ro0)
    for k in STATE_FIELDS:
        assert np.array_equal(f[k], getattr(fields, k))
    for k in STATE_CURRENTS:
        assert np.array_equal(c[k], getattr(cur, k))
    for s in ('electrons', 'ions'):
        for k in STATE_PARTICLE:
            assert np.array_equal(p[s][k], getattr(parts[s], k)), (s, k)
    with np.load(path) as z:               # the flat SaveRunState keys are the electrons'
        assert z['electrons'].dtype.names == STATE_PARTICLE
        for k in STATE_PARTICLE:
            assert np.array_equal(z[k], getattr(parts['electrons'], k))


def test_plasmastate_without_optional_keys_and_missing_quantities(tmp_path):
    rng = np.random.default_rng(2)
    parts, fields, cur = _state(rng, sorts=('electrons',))
    del parts['electrons'].dx_chaotic_perp          # written as zeros
    path = tmp_path / 's.npz'
    write_plasmastate(path, parts, fields, cur, 0.05)
    _, p, _, rho, xi = read_plasmastate(path, {'electrons': 0})
    assert rho is None and xi == 0.05
    assert not np.any(p['electrons']['dx_chaotic_perp'])
    with pytest.raises(KeyError):
        read_plasmastate(path, {'electrons': 0, 'ions': 1})


def test_reads_the_flat_layout_the_reference_writes(tmp_path):
    """SaveRunState's own file: flat particle keys, no per-sort record, an
    (1,) xi_plasma_layer array."""
    rng = np.random.default_rng(3)
    data = {k: rng.normal(size=(4, 4)) for k in STATE_PARTICLE}
    data.update({k: rng.normal(size=(7, 7)) for k in STATE_FIELDS})
    data.update({k: rng.normal(size=(2, 7, 7)) for k in STATE_CURRENTS})
    data['xi_plasma_layer'] = np.array([-3.5])
    path = tmp_path / 'ref.npz'
    np.savez_compressed(path, **data)
    f, p, c, rho, xi = read_plasmastate(path, {'electrons': 0})
    assert xi == -3.5 and rho is None
    for k in STATE_PARTICLE:
        assert np.array_equal(p['electrons'][k], data[k])
    assert np.array_equal(c['jz'], data['jz']) and np.array_equal(f['Phi'], data['Phi'])


def test_save_period_rule():
    """Saved when within half a step of a multiple of the period
    (diagnostics_3d.py:709-714), also for decimal steps where % misfires."""
    assert 0.15 % 0.05 > 0.04 and absremainder(0.15, 0.05) < 1e-16
    d = SaveRunState(output_period=2, saving_xi_period=0.5, save_plasma=True)
    d.pull_config(SimpleNamespace(getfloat=lambda k: {'time-step': 1.0, 'xi-step': 0.05}[k],
                                  getbool=lambda k: False))
    assert d.conditions_check(4.0, -1.5) and d.conditions_check(4.0, -1.52)
    assert not d.conditions_check(4.0, -1.45) and not d.conditions_check(3.0, -1.5)
    tiny = SaveRunState(output_period=0.1, saving_xi_period=0.001)
    tiny.pull_config(SimpleNamespace(getfloat=lambda k: {'time-step': 1.0, 'xi-step': 0.05}[k],
                                     getbool=lambda k: False))
    assert tiny.conditions_check(7.0, -0.35)         # periods raised to the steps
    with pytest.raises(ValueError):
        SaveRunState(save_beam=True).pull_config(
            SimpleNamespace(getfloat=lambda k: 1.0, getbool=lambda k: k == 'rigid-beam'))
