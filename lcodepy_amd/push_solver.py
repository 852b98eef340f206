"""The xi loop of one time step (push_solvers/push_solver.py:189-286)."""
from __future__ import annotations

from .beam3d.beam_calculator import BeamCalculator, RigidBeamCalculator
from .beam3d.data import BeamParticles
from .config import Config
from .plasma3d.solver import Plane2d3vPlasmaSolver


class PusherAndSolver3D:
    def __init__(self, config: Config):
        self.config = config
        self.plasma_solver = Plane2d3vPlasmaSolver(config)
        self.beam_particles_class = BeamParticles
        self.rigid_beam = config.getbool('rigid-beam')
        self.beam_calculator = (RigidBeamCalculator(config) if self.rigid_beam
                                else BeamCalculator(config))
        self.xi_max = config.getfloat('window-length')
        self.xi_step_size = config.getfloat('xi-step')
        self.xi_steps = round(self.xi_max / self.xi_step_size)
        self.grid_steps = config.getint('window-width-steps')
        self._plasmastate = None
        self.ez_axis = []          # on-axis Ez per xi step, kept on the device
        self.verbose = True

    def step_dt(self, plasma_fields, plasma_particles, plasma_currents, plasma_const_arrays,
                xi_plasma_layer_start, beam_source, beam_drain, current_time,
                diagnostics_list=()):
        """xi_steps + 1 - start iterations: beam layout, plasma step, beam move
        (push_solver.py:215-286)."""
        for _ in self.iter_dxi(plasma_fields, plasma_particles, plasma_currents,
                               plasma_const_arrays, xi_plasma_layer_start, beam_source,
                               beam_drain, current_time, diagnostics_list):
            pass

    def iter_dxi(self, plasma_fields, plasma_particles, plasma_currents, plasma_const_arrays,
                 xi_plasma_layer_start, beam_source, beam_drain, current_time,
                 diagnostics_list=()):
        """The body of step_dt as a generator yielding the xi-step index after
        every xi step (for callers that time or inspect single steps).

        Same order of dependencies as push_solver.py:215-286: the beam layer
        moved at step k is layout(k-1) plus what fell from the move at step
        k-1.  The split of a moved layer into (lost, moved, fell) is read back
        one step later, after the plasma step of the next layer is queued, so
        the host never waits for the stream to drain."""
        import torch
        from . import runtime
        rigid = self.rigid_beam
        self.beam_calculator.start_time_step()
        N = self.grid_steps
        ro_beam_prev = torch.zeros(N, N, dtype=torch.float64, device=runtime.device())
        self.ez_axis = []
        pending, prev_layout, fell_size = None, None, 0
        start = round(-xi_plasma_layer_start / self.xi_step_size) + 1
        for xi_i in range(start, self.xi_steps + 1):
            beam_layer_to_layout = beam_source.get_beam_layer_to_layout(xi_i)
            ro_beam_full = self.beam_calculator.layout_beam_layer(beam_layer_to_layout, xi_i)
            prev_plasma_fields = plasma_fields   # step_dxi returns fresh objects
            plasma_particles, plasma_fields, plasma_currents = self.plasma_solver.step_dxi(
                plasma_particles, plasma_fields, plasma_currents, plasma_const_arrays,
                ro_beam_full, ro_beam_prev)
            if not rigid:
                beam_layer_to_move = self.beam_particles_class()
                if pending is not None:
                    _lost, moved, fell = pending.result()
                    beam_drain.push_beam_layer(moved)
                    beam_layer_to_move.append(prev_layout).append(fell)
                    fell_size = fell.size
                pending = self.beam_calculator.move_beam_layer_async(
                    beam_layer_to_move, fell_size, xi_i, prev_plasma_fields, plasma_fields)
                prev_layout = beam_layer_to_layout
            # The beam calculator hands over a fresh density each step; a rigid
            # beam's callable may return its own buffer.
            ro_beam_prev = ro_beam_full.clone() if rigid else ro_beam_full
            xi_plasma_layer = -self.xi_step_size * xi_i
            for diagnostic in diagnostics_list:
                diagnostic.after_step_dxi(current_time, xi_plasma_layer, plasma_particles,
                                          plasma_fields, plasma_currents, ro_beam_full)
            self.ez_axis.append(plasma_fields.Ez[N // 2, N // 2].clone())
            if self.verbose and xi_i % 10 == 0:
                print(f't={current_time:+.4f}, xi={-xi_i * self.xi_step_size:+.4f} '
                      f'Ez={float(plasma_fields.Ez[N // 2, N // 2]):+.4e}', flush=True)
            self._plasmastate = (plasma_particles, plasma_fields, plasma_currents)
            yield xi_i
        if pending is not None:
            _lost, moved, _fell = pending.result()
            beam_drain.push_beam_layer(moved)
        xi_plasma_layer = -self.xi_step_size * self.xi_steps
        for diagnostic in diagnostics_list:
            diagnostic.dump(current_time, xi_plasma_layer, plasma_particles, plasma_fields,
                            plasma_currents, beam_drain)
        self.plasma_solver.ctx.raise_on_errors()
        self._plasmastate = (plasma_particles, plasma_fields, plasma_currents)
