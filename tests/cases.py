"""Configurations shared by the tests: the reference's fixture runs
(tests/integration/test_3D_part.py:9-82) and small synthetic cases."""
from lcodepy_amd.config import Config


def fixture_config(which: str) -> dict:
    """Config dicts of test_3D_part.py:9-26 and :59-64."""
    cfg = {'geometry': '3d', 'processing-unit-type': 'mi355x',
           'window-width': 8, 'transverse-step': 0.05,
           'window-length': 10, 'xi-step': 0.05,
           'time-limit': 1, 'time-step': 1,
           'ion-model': 'background', 'plasma-particles-per-cell': 4,
           'field-solver-subtraction-coefficient': 1.0}
    if which == 'beam_evol':
        cfg['window-length'] = 5
        cfg['time-limit'] = 5
    return cfg


def fixture_beam(which: str) -> dict:
    if which == 'test1':
        return {'current': 0.05, 'particles_in_layer': 5000,
                'default': {'length': 5.013256548}}
    return {'current': 0.5, 'particles_in_layer': 1000}


def adjusted(cfg: dict) -> Config:
    c = Config(cfg)
    c._adjust_config_values_3d()
    return c


def small_config(width=3.2, step=0.1, ppc=4, xi_step=0.05, length=2.0, ion_model='background',
                 **extra) -> dict:
    cfg = {'geometry': '3d', 'processing-unit-type': 'mi355x',
           'window-width': width, 'transverse-step': step,
           'window-length': length, 'xi-step': xi_step,
           'time-limit': 1, 'time-step': 1, 'ion-model': ion_model,
           'plasma-particles-per-cell': ppc,
           'field-solver-subtraction-coefficient': 1.0}
    cfg.update(extra)
    return cfg
