"""compute_rhoj (plasma3d/rhoj.py:7-22): deposit plus background ions."""
from __future__ import annotations

from ..config import Config
from .data import Arrays
from .weights import get_deposit_plasma


def get_rhoj_computer(config: Config):
    deposit = get_deposit_plasma(config)
    ion_model = config.get('ion-model')

    def compute_rhoj(particles: dict, const_arrays: Arrays) -> Arrays:
        ro, jx, jy, jz = deposit(particles, const_arrays)
        if ion_model == 'background':
            ro[1, :, :] = const_arrays.ro_initial
        return Arrays.wrap(ro=ro, jx=jx, jy=jy, jz=jz)

    return compute_rhoj
