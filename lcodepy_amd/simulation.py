"""Top-level 3D simulation (simulation.py:27-269), single process.

The reference pipelines time steps over MPI ranks (mpi/beam_io.py); on one
GPU the drained beam of each time step becomes the next one's source, as the
reference does with one rank (mpi/beam_io.py:81-119)."""
from __future__ import annotations

import copy

from .beam3d.beam_io import BeamDrain, BeamSource, RigidBeamDrain, RigidBeamSource
from .beam3d.data import BeamParticles
from .beam_generator import generate_beam, particle_dtype3d
from .config import DEFAULT_CONFIG_VALUES, Config
from .plasma3d.initialization import init_plasma, load_plasma
from .push_solver import PusherAndSolver3D


class Simulation:
    def __init__(self, config=DEFAULT_CONFIG_VALUES, beam_parameters=None, diagnostics=(),
                 runas_filename='runas.py'):
        self.config = copy.copy(config)
        self.beam_parameters = copy.copy(beam_parameters if beam_parameters is not None else {})
        self.diagnostics_list = list(diagnostics)
        self.runas_filename = runas_filename
        self.current_time = 0.
        self.beam_source = None
        self.beam_drain = None
        self._beam_particles = None
        self.external_plasmastate = False
        self.path_to_plasmastate = 'plasmastate.npz'
        self._geometry = None
        self._pull_config()

    def _pull_config(self):
        self._config = Config(self.config)
        self._time_limit = self._config.getfloat('time-limit')
        self._time_step_size = self._config.getfloat('time-step')
        self._rigid_beam = self._config.getbool('rigid-beam')
        geometry = self._config.get('geometry').lower()
        if self._geometry is None:
            self._geometry = geometry
        elif self._geometry != geometry:
            raise Exception('Sorry, update geometry does not support now.')
        if self._geometry != '3d':
            raise Exception('lcodepy_amd implements the 3d geometry only.')
        self._config.check_hip_backend()
        self.particle_dtype = particle_dtype3d
        self._config._adjust_config_values_3d()
        self._push_solver = PusherAndSolver3D(config=self._config)
        if self._rigid_beam:
            self.BeamSource, self.BeamDrain = RigidBeamSource, RigidBeamDrain
        else:
            self.BeamSource, self.BeamDrain = BeamSource, BeamDrain
        for diagnostic in self.diagnostics_list:
            diagnostic.pull_config(config=self._config)

    @property
    def push_solver(self) -> PusherAndSolver3D:
        return self._push_solver

    def load_beamfile(self, path_to_beamfile='beamfile.npz'):
        if self._rigid_beam:
            raise Exception('We cannot load a beam in the case of a rigid beam.')
        beam = BeamParticles()
        beam.load(path_to_beamfile)
        self._beam_particles = beam

    def _plasmastate(self, current_time):
        if self.external_plasmastate:
            return self._loaded_plasmastate
        return init_plasma(self._config, current_time)

    def step(self, N_steps=None):
        self._pull_config()
        if self.external_plasmastate:
            self._loaded_plasmastate = load_plasma(self._config, self.path_to_plasmastate)
        if self._rigid_beam:
            N_steps = 1
        elif N_steps is None:
            N_steps = int(self._time_limit / self._time_step_size)
        else:
            self._time_limit = N_steps * self._time_step_size + self.current_time
        if self._rigid_beam:
            self.beam_source = self.BeamSource(self._config, self.beam_parameters)
            self.beam_drain = self.BeamDrain(self._config)
        if self.beam_source is None:
            beam = self._beam_particles
            if beam is None:
                beam = generate_beam(self._config, self.beam_parameters)
            self.beam_source = self.BeamSource(self._config, beam)
            self.beam_drain = self.BeamDrain(self._config)
        self.current_time = self._time_step_size
        for t_i in range(N_steps):
            plasmastate = self._plasmastate(self.current_time)
            self._push_solver.step_dt(*plasmastate, self.beam_source, self.beam_drain,
                                      self.current_time, self.diagnostics_list)
            if t_i != N_steps - 1 and not self._rigid_beam:
                self.beam_source = self.BeamSource(self._config, self.beam_drain.beam_slice())
                self.beam_drain = self.BeamDrain(self._config)
            self.current_time = self.current_time + self._time_step_size
        print('The work is done!')
