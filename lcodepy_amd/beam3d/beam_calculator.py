"""Beam layer calculators (beam3d/beam_calculator.py:60-181)."""
from __future__ import annotations

from .. import runtime
from .._native import ptr, stream_ptr
from ..config import Config
from .data import BeamParticles, empty_rounded
from .move import get_move_beam_particles
from .weights import get_deposit_beam


class BeamCalculator:
    def __init__(self, config: Config):
        self.config = config
        self.ctx = runtime.context(config)
        self.grid_step_size = config.getfloat('transverse-step')
        self.grid_steps = config.getint('window-width-steps')
        self.time_step = config.getfloat('time-step')
        self.substep_energy = config.getfloat('beam-substepping-energy')
        self.deposit = get_deposit_beam(config)
        self.move_particles = get_move_beam_particles(config)

    def _zeros(self):
        import torch
        return torch.zeros(self.grid_steps, self.grid_steps, dtype=torch.float64,
                           device=runtime.device())

    def start_time_step(self):
        self.rho_layout = self._zeros()

    def layout_beam_layer(self, beam_layer: BeamParticles, plasma_layer_idx):
        """Deposits the layer, returns the finished density of this layer
        (divided by h^2) and keeps the part that belongs to the next one
        (beam_calculator.py:82-93), in one library call: the running buffer
        becomes the finished density, a fresh one receives the next layer's
        share."""
        import torch
        rho_next = torch.empty(self.grid_steps, self.grid_steps, dtype=torch.float64,
                               device=runtime.device())
        n = beam_layer.size
        if n:
            # Each row of a column slice of the packed tensors is itself
            # contiguous: the rows are passed without copying the layer.
            x, y, xi, q = beam_layer.x, beam_layer.y, beam_layer.xi, beam_layer.q_norm
        else:
            x = y = xi = q = None
        self.ctx.check(self.ctx.lib.lc3d_layout_beam_layer(
            self.ctx.handle, n, int(plasma_layer_idx), ptr(x), ptr(y), ptr(xi), ptr(q),
            ptr(self.rho_layout), ptr(rho_next), stream_ptr()), 'lc3d_layout_beam_layer')
        finished, self.rho_layout = self.rho_layout, rho_next
        return finished

    def start_moving_layer(self, beam_layer: BeamParticles, fell_size):
        """Substep multipliers for the particles that did not fall from the
        previous layer (beam_calculator.py:97-114)."""
        size = beam_layer.size - fell_size
        if size <= 0:
            return
        self.ctx.check(self.ctx.lib.lc3d_beam_substep(
            self.ctx.handle, size, ptr(beam_layer.q_m), ptr(beam_layer.pz),
            float(self.substep_energy), float(self.time_step), ptr(beam_layer.dt),
            ptr(beam_layer.remaining_steps), stream_ptr()), 'lc3d_beam_substep')

    def move_beam_layer(self, beam_layer: BeamParticles, fell_size, pl_layer_idx,
                        fields_after_layer, fields_before_layer):
        """Moves the layer and splits it into (lost, moved, fell)
        (beam_calculator.py:116-151)."""
        return self.move_beam_layer_async(beam_layer, fell_size, pl_layer_idx,
                                          fields_after_layer, fields_before_layer).result()

    def move_beam_layer_async(self, beam_layer: BeamParticles, fell_size, pl_layer_idx,
                              fields_after_layer, fields_before_layer) -> 'PendingMove':
        """Launches the move and the order-preserving split of the layer on
        the stream without waiting for them; ``result()`` of the returned
        object waits for these kernels only (not for later work on the
        stream) and gives (lost, moved, fell)."""
        import torch
        n, dev = beam_layer.size, runtime.device()
        if n == 0:
            return PendingMove(None, None, None, None)
        beam_layer.contiguous()
        # lost, moved, fell: every entry is written by the move
        flags = empty_rounded((3, n), torch.uint8, dev)
        self.start_moving_layer(beam_layer, fell_size)
        self.move_particles(pl_layer_idx - 1, beam_layer, fields_after_layer,
                            fields_before_layer, flags[0], flags[1], flags[2])
        nf, ni = beam_layer.f.shape[0], beam_layer.i.shape[0]
        out_f = empty_rounded((3, nf, n), beam_layer.f.dtype, dev)
        out_i = empty_rounded((3, ni, n), beam_layer.i.dtype, dev)
        counts_dev = torch.empty(3, dtype=torch.int64, device=dev)
        self.ctx.check(self.ctx.lib.lc3d_beam_split(
            self.ctx.handle, n, nf, ni, ptr(beam_layer.f), ptr(beam_layer.i), ptr(flags),
            ptr(out_f), ptr(out_i), ptr(counts_dev), stream_ptr()), 'lc3d_beam_split')
        counts = torch.empty(3, dtype=torch.int64, pin_memory=True)
        counts.copy_(counts_dev, non_blocking=True)
        done = torch.cuda.Event()
        done.record()
        return PendingMove(out_f, out_i, counts, done)

    def create_next_layer(self, beam_layer_to_layout, fell_to_next_layer, ro_beam_full):
        beam_layer_to_move = BeamParticles().append(beam_layer_to_layout).append(fell_to_next_layer)
        return beam_layer_to_move, fell_to_next_layer.size, ro_beam_full.clone()


class PendingMove:
    """(lost, moved, fell) of a launched beam move, resolved on demand."""

    def __init__(self, out_f, out_i, counts, event):
        self.out_f, self.out_i, self.counts, self.event = out_f, out_i, counts, event

    def result(self):
        if self.out_f is None:
            return BeamParticles(), BeamParticles(), BeamParticles()
        self.event.synchronize()
        sizes = self.counts.tolist()
        return tuple(BeamParticles(f=self.out_f[g][:, :sizes[g]], i=self.out_i[g][:, :sizes[g]])
                     for g in range(3))


class RigidBeamCalculator:
    """A beam given as a charge-density callable f(xp, xi, x, y) on the
    transverse grid, meshgrid 'xy' (beam_calculator.py:155-181)."""

    def __init__(self, config: Config):
        import torch
        self.xi_step_size = config.getfloat('xi-step')
        N = config.getint('window-width-steps')
        h = config.getfloat('transverse-step')
        grid = (torch.arange(N, dtype=torch.float64, device=runtime.device()) - N // 2) * h
        self.x_grid, self.y_grid = torch.meshgrid(grid, grid, indexing='xy')

    def start_time_step(self):
        pass

    def layout_beam_layer(self, beam_charge_distribution_function, plasma_layer_idx):
        import numpy as np
        import torch
        xi = -plasma_layer_idx * self.xi_step_size
        rho = beam_charge_distribution_function(torch, xi, self.x_grid, self.y_grid)
        if isinstance(rho, np.ndarray):
            rho = torch.as_tensor(rho, device=self.x_grid.device)
        return torch.as_tensor(rho, dtype=torch.float64, device=self.x_grid.device).contiguous()

    def move_beam_layer(self, beam_layer, fell_size, pl_layer_idx, fields_after_layer,
                        fields_before_layer):
        return None, None, None

    def create_next_layer(self, beam_layer_to_layout, fell_to_next_layer, ro_beam_full):
        return None, None, ro_beam_full.clone()
