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
        # -xi of the sorted beam in ascending order, on the host once, so
        # that layer sizes need no device round trip.
        self._neg_xi = -beam.xi.cpu().numpy()
        self._offset = 0

    def get_beam_layer_to_layout(self, plasma_layer_idx):
        """Particles with xi_{i+1} <= xi < xi_i, counted over the whole
        remaining beam and cut from its front, as in beam_io.py:29-47."""
        xi_min = -self.xi_step_size * plasma_layer_idx
        xi_max = -self.xi_step_size * (plasma_layer_idx + 1)
        rest = self._neg_xi[self._offset:]
        n = int(np.searchsorted(rest, -xi_max, side='right')
                - np.searchsorted(rest, -xi_min, side='right'))
        n = max(n, 0)
        layer = self.beam.slice(self._offset, self._offset + n)
        self._offset += n
        return layer


class BeamDrain:
    """Collects the moved layers and concatenates them in one operation when
    read, or once JOIN_EVERY of them are waiting: a moved layer is a view
    into its split's output (lost, moved and fell particles of the layer),
    which it would otherwise keep allocated."""

    JOIN_EVERY = 32

    def __init__(self, config: Config):
        self._layers, self._lost = [], []

    @staticmethod
    def _joined(layers):
        import torch
        if len(layers) > 1:
            f = torch.cat([layer.f for layer in layers], dim=1)
            i = torch.cat([layer.i for layer in layers], dim=1)
            layers[:] = [BeamParticles(f=f, i=i)]
        return layers[0] if layers else BeamParticles()

    @property
    def beam_buffer(self) -> BeamParticles:
        return self._joined(self._layers)

    @property
    def lost_buffer(self) -> BeamParticles:
        return self._joined(self._lost)

    def push_beam_layer(self, beam_layer: BeamParticles):
        if beam_layer.size > 0:
            self._layers.append(beam_layer)
            if len(self._layers) > self.JOIN_EVERY:
                self._joined(self._layers)

    def push_beam_lost(self, lost_layer: BeamParticles):
        if lost_layer.size > 0:
            self._lost.append(lost_layer)
            if len(self._lost) > self.JOIN_EVERY:
                self._joined(self._lost)

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
