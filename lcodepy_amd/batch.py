"""Many independent wakefield runs advanced together, each xi step a HIP
graph replay per run on its own stream.

A run here is one time step's xi loop (push_solvers/push_solver.py:215-286)
with a rigid driver (beam3d/beam_calculator.py:148-181): a Gaussian bunch
whose charge density is tabulated per xi step.  :class:`StaticRun` keeps all
of a run's state in fixed device buffers (particle, field and current sets
used alternately as layer k and k+1, the half-step fields, two beam-density
grids, the on-axis Ez of every step) so that one xi step -- the beam density
of the new layer, the whole plasma step (lc3d_plasma_step) and the on-axis
Ez record -- is a fixed sequence of launches on fixed addresses.  It is
captured once per parity of the step as a HIP graph (torch.cuda.CUDAGraph)
and replayed; a device step counter selects the density and the Ez slot.
Each run has a context of its own, and :func:`main` replays every run's
graph for each xi step on the run's own stream, so runs of different grid
sizes share the GPU.

``python -m lcodepy_amd.batch`` times a seeded parameter scan (48 runs, grids
drawn from 129/257/385/513, 1 or 4 particles per cell, driver density and
width drawn from uniform ranges, xi_step 0.05, 500 xi steps) batched and as
the plain float64 form run one after another.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import math
import sys
import time

import numpy as np

from . import _native, runtime
from .config import Config


class RigidDriver:
    """Charge density of a rigid Gaussian driver,
    -n_b exp(-r^2 / 2 sigma_r^2) exp(-(xi - xi0)^2 / 2 sigma_xi^2) for
    |xi - xi0| <= 3 sigma_xi, at xi = -k xi_step (k = the xi-step index):
    a transverse profile and a longitudinal table."""

    def __init__(self, N: int, h: float, xi_step: float, n_steps: int, n_b: float,
                 sigma_r: float, sigma_xi: float = 1.0, xi0: float = -3.0):
        # layers 0 .. n_steps (the xi loop's n_steps + 1 iterations)
        g = (np.arange(N) - N // 2) * h
        self.profile = -n_b * np.exp(-(g[:, None] ** 2 + g[None, :] ** 2) / (2 * sigma_r ** 2))
        d = -xi_step * np.arange(n_steps + 1) - xi0
        self.long = np.where(np.abs(d) <= 3 * sigma_xi, np.exp(-d ** 2 / (2 * sigma_xi ** 2)), 0.0)
        self.xi_step = xi_step

    def callable(self):
        """The density as a reference-style rigid-beam callable
        f(xp, xi, x_grid, y_grid) (radially symmetric, so the 'xy' meshgrid
        orientation does not matter)."""
        import torch
        prof = {}

        def f(xp, xi, x_grid, y_grid):
            dev = x_grid.device
            if dev not in prof:
                prof[dev] = torch.as_tensor(self.profile, device=dev)
            k = int(round(-xi / self.xi_step))
            return prof[dev] * torch.as_tensor(self.long[k:k + 1], device=dev)[0]

        return f


def run_config(N: int, ppc: int, precision: str, n_steps: int, xi_step: float = 0.05,
               h: float = 0.05) -> dict:
    return {'geometry': '3d', 'processing-unit-type': 'mi355x', 'precision': precision,
            'window-width': (N - 1) * h / 2 + h / 2, 'transverse-step': h,
            'window-length': n_steps * xi_step, 'xi-step': xi_step, 'time-limit': 1.0,
            'time-step': 1.0, 'ion-model': 'background', 'plasma-particles-per-cell': ppc,
            'rigid-beam': True}


class StaticRun:
    """One run's xi loop on fixed buffers, steppable eagerly or by graphs."""

    def __init__(self, cfg: dict, driver: RigidDriver):
        import torch
        from .plasma3d import _bind
        from .plasma3d.fields import new_fields
        from .plasma3d.initialization import init_plasma
        config = Config(cfg)
        config._adjust_config_values_3d()
        self.config = config
        N = self.N = config.getint('window-width-steps')
        if N != driver.profile.shape[0]:
            raise ValueError(f'the driver is tabulated for N={driver.profile.shape[0]}, the run has {N}')
        # xi steps of the window: xi_steps + 1, layers 0 .. xi_steps
        # (push_solver.py:234-236)
        self.n_steps = round(config.getfloat('window-length') / config.getfloat('xi-step')) + 1
        self.ctx = runtime.context(config, private=True)
        dt = self.dt = runtime.Dtypes(config.precision())
        dev = runtime.device()
        fields, particles, currents, const, _ = init_plasma(config)
        el = particles['electrons']
        self.n_particles = int(el.m.numel())
        other = el.copy()
        # The constant arrays are shared: the library keeps a deposit's cell
        # order for the sort with the same x_init, which both layers then have.
        other.x_init, other.y_init, other.m = el.x_init, el.y_init, el.m
        self.P = [el, other]
        self.F = [fields, new_fields(N, dt.full)]
        for k in ('Ex', 'Ey', 'Ez', 'Bx', 'By', 'Bz', 'Phi'):
            getattr(self.F[1], k).zero_()
        self.H = new_fields(N, dt.half)
        self.C = [currents, currents.copy()]
        self.RB = [torch.zeros(N, N, dtype=torch.float64, device=dev) for _ in range(2)]
        self.bg = const.ro_initial.to(dt.current).contiguous()
        self.profile = torch.as_tensor(driver.profile, device=dev)
        self.long = torch.as_tensor(driver.long, device=dev)
        self.k = torch.zeros(1, dtype=torch.int64, device=dev)     # layer of the next step
        self.ez = torch.zeros(self.n_steps, dtype=torch.float64, device=dev)
        self.sorts, self.args = [], []
        for a in range(2):
            b = 1 - a
            s = (_native.Sort * 1)()
            s[0] = _bind.sort_struct(self.P[a], self.P[b], 0)
            self.sorts.append(s)
            self.args.append((ctypes.byref(_bind.fields7(self.F[a])),
                              ctypes.byref(_bind.currents(self.C[a])),
                              ctypes.byref(_bind.fields7(self.F[b])),
                              ctypes.byref(_bind.fields7(self.H)),
                              ctypes.byref(_bind.currents(self.C[b]))))
        self.graphs = None
        self.stream = None

    def step_launches(self, a: int) -> None:
        """The launches of the xi step of layer k (k % 2 == a) on the current
        stream: beam density of the layer, plasma step, on-axis Ez."""
        import torch
        b = 1 - a
        N = self.N
        torch.mul(self.profile, torch.index_select(self.long, 0, self.k), out=self.RB[b])
        fp, cp, ff, fh, cf = self.args[a]
        self.ctx.check(self.ctx.lib.lc3d_plasma_step(
            self.ctx.handle, 1, self.sorts[a], fp, cp, self.RB[b].data_ptr(),
            self.RB[a].data_ptr(), self.bg.data_ptr(), ff, fh, cf,
            _native.stream_ptr()), 'lc3d_plasma_step')
        self.ez.index_copy_(0, self.k, self.F[b].Ez[N // 2, N // 2:N // 2 + 1])
        self.k.add_(1)

    def capture(self, stream=None) -> None:
        import torch
        self.stream = stream or torch.cuda.Stream()
        self.graphs = []
        for a in range(2):
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g, stream=self.stream):
                self.step_launches(a)
            self.graphs.append(g)

    def step(self, i: int) -> None:
        """The xi step of layer i (from 0) by graph replay on the run's stream."""
        import torch
        with torch.cuda.stream(self.stream):
            self.graphs[i % 2].replay()

    def on_axis_ez(self) -> np.ndarray:
        return self.ez[:self.n_steps].cpu().numpy()


def scan_parameters(n_runs: int, seed: int = 1):
    """(N, ppc, n_b, sigma_r) of a seeded parameter scan: N from {129, 257,
    385, 513}, ppc from {1, 4}, n_b ~ U[0.05, 1], sigma_r ~ U[0.3, 1.5]."""
    rng = np.random.default_rng(seed)
    return [(int(rng.choice([129, 257, 385, 513])), int(rng.choice([1, 4])),
             float(rng.uniform(0.05, 1.0)), float(rng.uniform(0.3, 1.5)))
            for _ in range(n_runs)]


def _eager_xi_loop(cfg: dict, driver: RigidDriver) -> tuple:
    """The plain form: Simulation's xi loop (PusherAndSolver3D) with the
    driver as a rigid-beam callable.  Returns (seconds, on-axis Ez)."""
    import torch
    from .beam3d.beam_io import BeamDrain, RigidBeamSource
    from .plasma3d.initialization import init_plasma
    from .push_solver import PusherAndSolver3D
    config = Config(cfg)
    config._adjust_config_values_3d()
    fields, particles, currents, const, xi0 = init_plasma(config)
    solver = PusherAndSolver3D(config)
    solver.verbose = False
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    solver.step_dt(fields, particles, currents, const, xi0,
                   RigidBeamSource(config, driver.callable()), BeamDrain(config), 1.0)
    torch.cuda.synchronize()
    return time.perf_counter() - t0, np.array([float(v) for v in solver.ez_axis])


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description='batched parameter scan (HIP graphs per run)')
    ap.add_argument('--runs', type=int, default=48)
    ap.add_argument('--steps', type=int, default=500)
    ap.add_argument('--precision', default='fp32-bulk', choices=('fp32-bulk', 'fp64'))
    ap.add_argument('--seed', type=int, default=1)
    ap.add_argument('--no-sequential', action='store_true')
    args = ap.parse_args(argv)
    import torch
    torch.cuda.set_device(0)
    params = scan_parameters(args.runs, args.seed)
    xi_step = 0.05
    n_iter = args.steps + 1          # xi_steps + 1 iterations per run
    steps_total = args.runs * n_iter
    out = {'workload': f'{args.runs} runs, N from 129/257/385/513, 1 or 4 ppc, rigid Gaussian '
                       f'driver n_b ~ U[0.05, 1], sigma_r ~ U[0.3, 1.5], sigma_xi 1, xi_step '
                       f'{xi_step}, {args.steps} xi steps each ({n_iter} iterations of the xi '
                       f'loop, seed {args.seed})',
           'grids': [p[0] for p in params], 'ppc': [p[1] for p in params]}
    print(f'[batch] building {args.runs} runs ({args.precision})', file=sys.stderr, flush=True)
    runs = []
    for N, ppc, n_b, sr in params:
        cfg = run_config(N, ppc, args.precision, args.steps, xi_step)
        drv = RigidDriver(N, 0.05, xi_step, args.steps, n_b, sr)
        r = StaticRun(cfg, drv)
        r.capture()
        runs.append(r)
    torch.cuda.synchronize()
    print('[batch] stepping', file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    for i in range(n_iter):
        for r in runs:
            r.step(i)
    torch.cuda.synchronize()
    dt_batch = time.perf_counter() - t0
    for r in runs:
        r.ctx.raise_on_errors()
    out['batched'] = {'precision': args.precision, 'seconds': dt_batch,
                      'xi_steps_per_s': steps_total / dt_batch}
    print(f'[batch] batched {args.precision}: {steps_total / dt_batch:.0f} xi-steps/s',
          file=sys.stderr, flush=True)
    if not args.no_sequential:
        total = 0.0
        worst = 0.0
        for (N, ppc, n_b, sr), r in zip(params, runs):
            cfg = run_config(N, ppc, 'fp64', args.steps, xi_step)
            drv = RigidDriver(N, 0.05, xi_step, args.steps, n_b, sr)
            sec, ez = _eager_xi_loop(cfg, drv)
            total += sec
            ezb = r.on_axis_ez()
            k = min(len(ez), len(ezb))
            scale = np.max(np.abs(ez[:k]))
            if scale > 0:
                worst = max(worst, float(np.max(np.abs(ezb[:k] - ez[:k])) / scale))
            torch.cuda.empty_cache()
        out['sequential_fp64'] = {'seconds': total, 'xi_steps_per_s': steps_total / total}
        out['batched_over_sequential'] = (steps_total / dt_batch) / (steps_total / total)
        out['on_axis_ez_rel_err_batched_vs_sequential_fp64'] = worst
        print(f'[batch] sequential fp64: {steps_total / total:.0f} xi-steps/s',
              file=sys.stderr, flush=True)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == '__main__':
    sys.exit(main())
