"""Beam macro-particles as device tensors (beam3d/data.py:21-169).

Structure of arrays, packed: the nine float64 quantities (xi x y px py pz q_m
q_norm dt) are the rows of one (9, n) tensor and the two int64 ones (id,
remaining_steps) the rows of one (2, n) tensor, so that cutting, gathering
and appending a layer is one or two tensor operations rather than eleven.
Each quantity is a contiguous row view (``beam.x`` ...).  The on-disk /
exchange record is the reference's ``particle_dtype3d`` (beam3d/data.py:5-7)."""
from __future__ import annotations

import numpy as np

from ..beam_generator import particle_dtype3d

FLOAT_KEYS = ('xi', 'x', 'y', 'px', 'py', 'pz', 'q_m', 'q_norm', 'dt')
INT_KEYS = ('id', 'remaining_steps')
KEYS = FLOAT_KEYS + INT_KEYS


def empty_rounded(shape, dtype, device):
    """An uninitialised contiguous tensor of ``shape`` at the start of a
    buffer whose element count is rounded up to a power of two: the layers
    of a beam grow and shrink step by step, and with few distinct buffer
    sizes the caching allocator reuses its blocks instead of reserving new
    device memory (a hipMalloc in the middle of the xi loop) as they grow."""
    import torch
    n = 1
    for s in shape:
        n *= int(s)
    cap = 1 << (n - 1).bit_length() if n > 1 else n
    return torch.empty(cap, dtype=dtype, device=device)[:n].view(*shape)


def _row_property(store: str, row: int):
    def get(self):
        return getattr(self, store)[row]

    def set_(self, value):
        getattr(self, store)[row].copy_(value)

    return property(get, set_)


class BeamParticles:
    def __init__(self, xp=None, f=None, i=None):
        import torch
        from ..runtime import device
        self.xp = torch
        if f is None:
            dev = device()
            f = torch.zeros(len(FLOAT_KEYS), 0, dtype=torch.float64, device=dev)
            i = torch.zeros(len(INT_KEYS), 0, dtype=torch.int64, device=dev)
        self.f, self.i = f, i

    @property
    def size(self) -> int:
        return int(self.f.shape[1])

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
        n = len(beam_array)
        f = np.zeros((len(FLOAT_KEYS), n))
        for r, k in enumerate(FLOAT_KEYS):
            if k in particle_dtype3d.names:
                f[r] = beam_array[k]
        i = np.zeros((len(INT_KEYS), n), dtype=np.int64)
        i[0] = beam_array['id']
        self.f = torch.as_tensor(f).to(dev)
        self.i = torch.as_tensor(i).to(dev)

    def load(self, *args, **kwargs):
        with np.load(*args, **kwargs) as loaded:
            rec = np.zeros(len(loaded['id']), dtype=particle_dtype3d)
            for k in particle_dtype3d.names:
                rec[k] = loaded[k]
        self.init_generated(rec)

    def save(self, *args, **kwargs):
        np.savez_compressed(*args, **kwargs,
                            **{k: getattr(self, k).cpu().numpy() for k in particle_dtype3d.names})

    def contiguous(self) -> 'BeamParticles':
        """Rows of a column slice are strided in the packed tensors; the
        kernels take contiguous rows."""
        if not self.f.is_contiguous():
            self.f = empty_rounded(self.f.shape, self.f.dtype, self.f.device).copy_(self.f)
        if not self.i.is_contiguous():
            self.i = empty_rounded(self.i.shape, self.i.dtype, self.i.device).copy_(self.i)
        return self

    def xi_sorted(self):
        """Sort by descending xi, stably (data.py:88-104)."""
        import torch
        idx = torch.argsort(-self.xi, stable=True)
        self.f, self.i = self.f[:, idx], self.i[:, idx]

    def get_layer(self, indexes_arr) -> 'BeamParticles':
        return BeamParticles(f=self.f[:, indexes_arr], i=self.i[:, indexes_arr])

    def slice(self, start: int, stop: int) -> 'BeamParticles':
        return BeamParticles(f=self.f[:, start:stop], i=self.i[:, start:stop])

    def cut_beam_layer(self, layer_length: int):
        layer = self.slice(0, layer_length)
        rest = self.slice(layer_length, self.size)
        self.f, self.i = rest.f, rest.i
        return layer, self

    def append(self, other: 'BeamParticles') -> 'BeamParticles':
        import torch
        if other.size:
            if self.size:
                n = self.size + other.size
                f = empty_rounded((self.f.shape[0], n), self.f.dtype, self.f.device)
                i = empty_rounded((self.i.shape[0], n), self.i.dtype, self.i.device)
                self.f = torch.cat((self.f, other.f), dim=1, out=f)
                self.i = torch.cat((self.i, other.i), dim=1, out=i)
            else:
                self.f, self.i = other.f, other.i
        return self


for _r, _k in enumerate(FLOAT_KEYS):
    setattr(BeamParticles, _k, _row_property('f', _r))
for _r, _k in enumerate(INT_KEYS):
    setattr(BeamParticles, _k, _row_property('i', _r))
