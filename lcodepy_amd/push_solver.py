"""The xi loop of one time step (push_solvers/push_solver.py:189-286)."""
from __future__ import annotations

from .beam3d.beam_calculator import BeamCalculator, RigidBeamCalculator
from .beam3d.data import BeamParticles
from .config import Config
from .plasma3d.solver import Plane2d3vPlasmaSolver


def _read_on(layer, stream):
    """Marks a beam layer made on the beam stream as read by ``stream`` (the
    drain joins its layers there)."""
    if layer.size:
        layer.f.record_stream(stream)
        layer.i.record_stream(stream)


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
        the host never waits for the stream to drain.

        With a particle beam, the beam work of the layers that hold particles
        runs on a second stream, beside the plasma steps: the deposit of layer
        k+1 is queued there as soon as the plasma step of layer k is, and the
        move of step k waits (by an event) for that step's fields only, so
        both overlap the plasma step of layer k+1, which waits for its layer's
        deposit.  Empty layers are finished on the main stream (a
        cross-stream wait costs more than their one small kernel).  Tensors
        that one stream makes and the other reads are recorded on the reading
        stream, so the allocator does not reuse them early."""
        import torch
        from . import runtime
        rigid = self.rigid_beam
        self.beam_calculator.start_time_step()
        N = self.grid_steps
        main = torch.cuda.current_stream()
        side = None if rigid else torch.cuda.Stream(device=runtime.device())
        ro_beam_prev = torch.zeros(N, N, dtype=torch.float64, device=runtime.device())
        self.ez_axis = []
        pending, prev_layout, fell_size = None, None, 0
        start = round(-xi_plasma_layer_start / self.xi_step_size) + 1

        on_side = [False]   # whether the last deposit ran on the beam stream

        def layout_ahead(xi_i):
            """Deposits layer xi_i: on the beam stream if it holds particles
            (after whatever the main stream queued since the last deposit
            there: the zeroed running density, the sorted beam, an empty
            layer's deposit), else on the main stream."""
            layer = beam_source.get_beam_layer_to_layout(xi_i)
            ready = None
            if layer.size == 0:
                ro = self.beam_calculator.layout_beam_layer(layer, xi_i)
                on_side[0] = False
            else:
                if not on_side[0]:
                    side.wait_stream(main)
                with torch.cuda.stream(side):
                    ro = self.beam_calculator.layout_beam_layer(layer, xi_i)
                    ready = torch.cuda.Event()
                    ready.record(side)
                on_side[0] = True
            ro.record_stream(main)
            return layer, ro, ready

        next_layout = None if rigid or start > self.xi_steps else layout_ahead(start)
        for xi_i in range(start, self.xi_steps + 1):
            if rigid:
                beam_layer_to_layout = beam_source.get_beam_layer_to_layout(xi_i)
                ro_beam_full = self.beam_calculator.layout_beam_layer(beam_layer_to_layout, xi_i)
            else:
                beam_layer_to_layout, ro_beam_full, ready = next_layout
                if ready is not None:
                    main.wait_event(ready)
            prev_plasma_fields = plasma_fields   # step_dxi returns fresh objects
            plasma_particles, plasma_fields, plasma_currents = self.plasma_solver.step_dxi(
                plasma_particles, plasma_fields, plasma_currents, plasma_const_arrays,
                ro_beam_full, ro_beam_prev)
            if not rigid:
                next_layout = layout_ahead(xi_i + 1) if xi_i < self.xi_steps else None
                beam_layer_to_move = self.beam_particles_class()
                fell = None
                if pending is not None:
                    _lost, moved, fell = pending.result()
                    _read_on(moved, main)
                    beam_drain.push_beam_layer(moved)
                    fell_size = fell.size
                if fell is not None and prev_layout.size + fell.size > 0:
                    fields_ready = torch.cuda.Event()
                    fields_ready.record(main)
                    side.wait_event(fields_ready)
                    for fields in (prev_plasma_fields, plasma_fields):
                        for name in fields.names():
                            t = getattr(fields, name)
                            if isinstance(t, torch.Tensor):
                                t.record_stream(side)
                    with torch.cuda.stream(side):
                        beam_layer_to_move.append(prev_layout).append(fell)
                        pending = self.beam_calculator.move_beam_layer_async(
                            beam_layer_to_move, fell_size, xi_i, prev_plasma_fields,
                            plasma_fields)
                else:
                    # nothing to move: no kernel is launched
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
            _read_on(moved, main)
            beam_drain.push_beam_layer(moved)
        xi_plasma_layer = -self.xi_step_size * self.xi_steps
        for diagnostic in diagnostics_list:
            diagnostic.dump(current_time, xi_plasma_layer, plasma_particles, plasma_fields,
                            plasma_currents, beam_drain)
        self.plasma_solver.ctx.raise_on_errors()
        self._plasmastate = (plasma_particles, plasma_fields, plasma_currents)
