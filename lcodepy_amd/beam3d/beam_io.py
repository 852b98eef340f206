"""Beam layer source and drain (beam3d/beam_io.py:7-107)."""
from __future__ import annotations

import numpy as np

from ..config import Config
from .data import BeamParticles


class BeamSource:
    """Cuts the beam into xi layers, head first."""

    def __init__(self, config: Config, beam_particles):
        self.xi_step_size = config.getfloat('xi-step')
        if isinstance(beam_particles, np.ndarray):
            beam = BeamParticles()
            beam.init_generated(beam_particles)
        else:
            beam = beam_particles
        beam.xi_sorted()
        self.beam = beam

    def get_beam_layer_to_layout(self, plasma_layer_idx):
        """Particles with xi_{i+1} <= xi < xi_i, counted over the whole
        remaining beam as in beam_io.py:29-47."""
        xi_min = -self.xi_step_size * plasma_layer_idx
        xi_max = -self.xi_step_size * (plasma_layer_idx + 1)
        xi = self.beam.xi
        n = int(((xi_max <= xi) & (xi < xi_min)).sum().item()) if xi.numel() else 0
        layer, self.beam = self.beam.cut_beam_layer(n)
        return layer


class BeamDrain:
    def __init__(self, config: Config):
        self.beam_buffer = BeamParticles()
        self.lost_buffer = BeamParticles()

    def push_beam_layer(self, beam_layer: BeamParticles):
        if beam_layer.size > 0:
            self.beam_buffer.append(beam_layer)

    def push_beam_lost(self, lost_layer: BeamParticles):
        if lost_layer.size > 0:
            self.lost_buffer.append(lost_layer)

    def beam_slice(self):
        return self.beam_buffer

    def save(self, *args, **kwargs):
        self.beam_buffer.save(*args, **kwargs)


class RigidBeamSource:
    def __init__(self, config: Config, beam_charge_distribution_function):
        self.beam_charge_distribution_function = beam_charge_distribution_function

    def get_beam_layer_to_layout(self, plasma_layer_idx):
        return self.beam_charge_distribution_function


class RigidBeamDrain:
    def __init__(self, config: Config):
        pass

    def push_beam_layer(self, beam_layer):
        pass

    def push_beam_lost(self, lost_layer):
        pass

    def beam_slice(self):
        return None
