"""Beam macro-particles as device tensors (beam3d/data.py:21-169).

Structure of arrays: xi x y px py pz q_m q_norm dt (float64), id and
remaining_steps (int64).  The on-disk / exchange record is the reference's
``particle_dtype3d`` (beam3d/data.py:5-7)."""
from __future__ import annotations

import numpy as np

from ..beam_generator import particle_dtype3d

FLOAT_KEYS = ('xi', 'x', 'y', 'px', 'py', 'pz', 'q_m', 'q_norm', 'dt')
KEYS = ('xi', 'x', 'y', 'px', 'py', 'pz', 'q_m', 'q_norm', 'id', 'dt', 'remaining_steps')


class BeamParticles:
    def __init__(self, xp=None):
        import torch
        from ..runtime import device
        self.xp = torch
        dev = device()
        for k in KEYS:
            dt = torch.float64 if k in FLOAT_KEYS else torch.int64
            setattr(self, k, torch.zeros(0, dtype=dt, device=dev))

    @property
    def size(self) -> int:
        return int(self.id.numel())

    @property
    def particles(self) -> np.ndarray:
        rec = np.zeros(self.size, dtype=particle_dtype3d)
        for k in particle_dtype3d.names:
            rec[k] = getattr(self, k).cpu().numpy()
        return rec

    def init_generated(self, beam_array: np.ndarray):
        import torch
        from ..runtime import device
        dev = device()
        for k in particle_dtype3d.names:
            dt = torch.int64 if k == 'id' else torch.float64
            setattr(self, k, torch.as_tensor(np.ascontiguousarray(beam_array[k])).to(dev, dt))
        self.dt = torch.zeros_like(self.q_norm)
        self.remaining_steps = torch.zeros_like(self.id)

    def load(self, *args, **kwargs):
        with np.load(*args, **kwargs) as loaded:
            rec = np.zeros(len(loaded['id']), dtype=particle_dtype3d)
            for k in particle_dtype3d.names:
                rec[k] = loaded[k]
        self.init_generated(rec)

    def save(self, *args, **kwargs):
        np.savez_compressed(*args, **kwargs,
                            **{k: getattr(self, k).cpu().numpy() for k in particle_dtype3d.names})

    def _set(self, src: dict):
        for k in KEYS:
            setattr(self, k, src[k])

    def xi_sorted(self):
        """Sort by descending xi (data.py:88-104)."""
        import torch
        idx = torch.argsort(-self.xi, stable=True)
        self._set({k: getattr(self, k)[idx] for k in KEYS})

    def get_layer(self, indexes_arr) -> 'BeamParticles':
        out = BeamParticles()
        out._set({k: getattr(self, k)[indexes_arr].contiguous() for k in KEYS})
        return out

    def cut_beam_layer(self, layer_length: int):
        layer = BeamParticles()
        layer._set({k: getattr(self, k)[:layer_length] for k in KEYS})
        self._set({k: getattr(self, k)[layer_length:] for k in KEYS})
        return layer, self

    def append(self, other: 'BeamParticles') -> 'BeamParticles':
        import torch
        self._set({k: torch.cat((getattr(self, k), getattr(other, k))) for k in KEYS})
        return self
