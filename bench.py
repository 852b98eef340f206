#!/usr/bin/env python
"""xi-steps/s of lcodePy's quasistatic 3D xi step on one MI355X.

Workload (the 513 x 513, 4 ppc configuration): transverse window 25.6
(step 0.05, N = 513), 4 plasma macro-particles per cell (964^2 = 929,296)
with seeded thermal momenta p / (m c) ~ N(0, 1e-3) per component (a macro-
particle stores p scaled by its mass m = h^2 / ppc, gamma = sqrt(m^2 + p^2) /
m, so the stored momenta are m times the draws), background ions, xi_step 0.02 over
a window of length 20 (1001 xi steps); a 1e6-particle seeded electron
driver, n_b/n0 = 1, sigma_r = 0.5, sigma_xi = 1 (truncated at +-3 sigma),
gamma ~ N(1000, 10).  One "step" is one xi step of the whole path: beam layer
deposit, the plasma step (free stream, 3 pushes, 3 deposits, 2 field solves)
and the beam push, with all inputs resident in HBM.  The FP32-bulk form is
the headline value; the float64 form runs on the same inputs for the speed
ratio and the on-axis Ez error over the same xi steps.  After the timed
steps, the remaining steps of the window run with the library's per-launch
event timing on (push, deposit gather, solve GEMM) for the roofline.
Synthetic data (seeded), no files read except the traffic summary under
profiles/ when present.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = ('xi-steps/s at 513² 4ppc, FP32-bulk vs FP64 @Ez err≤1e-4; HBM TB/s')
HBM_PEAK = 8.0e12
MFMA_PEAK = {'fp32-bulk': 157.3e12, 'fp64': 78.6e12}   # dense matrix-core FLOP/s
WINDOW_STEPS = 1001                                     # xi steps of one time step
TRAFFIC_FILE = os.path.join(ROOT, 'profiles', 'xi_step_513_4ppc_hbm_traffic.md')
# Launches of each timed class per xi step (one plasma sort): 3 smart
# pushes, 3 deposits, and per solve 4 transform GEMMs (matrix) or 3 FFT
# kernels (rfft), 2 solves.
LAUNCHES_PER_STEP = {'push': 3, 'deposit': 3, 'solve': {'matrix': 8, 'rfft': 6}}
# Algorithmic words per particle per xi step (46: 9 + 2 x 11 reads, 3 x 5
# writes) and per grid point (166: 3 pushes x 10 N^2, 2 solves x 68 N^2).
WORDS_PER_PARTICLE = 46
WORDS_PER_GRID_POINT = 166
SOLVE_WORDS_PER_GRID_POINT = 68


def workload_config(precision: str, field_solver: str = 'auto') -> dict:
    return {'geometry': '3d', 'processing-unit-type': 'mi355x', 'precision': precision,
            'window-width': 12.8, 'transverse-step': 0.05,
            'window-length': 20.0, 'xi-step': 0.02, 'time-limit': 1.0, 'time-step': 1.0,
            'ion-model': 'background', 'plasma-particles-per-cell': 4,
            'field-solver': field_solver}


def driver_beam(n_particles: int, xi_step: float, seed: int = 1) -> np.ndarray:
    """Seeded Gaussian driver, head at xi = 0 (particle_dtype3d records)."""
    from lcodepy_amd.beam_generator import particle_dtype3d
    rng = np.random.RandomState(seed)
    sig_r, sig_xi, n_b, xi0 = 0.5, 1.0, 1.0, -3.0
    xi = np.empty(0)
    while xi.size < n_particles:
        d = rng.normal(0.0, sig_xi, 2 * n_particles)
        xi = np.concatenate([xi, xi0 + d[np.abs(d) <= 3 * sig_xi]])
    xi = xi[:n_particles]
    b = np.zeros(n_particles, dtype=particle_dtype3d)
    b['xi'] = xi
    b['x'] = rng.normal(0.0, sig_r, n_particles)
    b['y'] = rng.normal(0.0, sig_r, n_particles)
    b['px'] = rng.normal(0.0, 1e-2, n_particles)
    b['py'] = rng.normal(0.0, 1e-2, n_particles)
    b['pz'] = rng.normal(1000.0, 10.0, n_particles)
    b['q_m'] = -1.0
    b['q_norm'] = -n_b * 2 * math.pi * sig_r ** 2 * math.sqrt(2 * math.pi) * sig_xi / (
        n_particles * xi_step)
    b['id'] = np.arange(1, n_particles + 1)
    return b[np.argsort(-b['xi'])]


def electron_mass(cfg) -> float:
    """Mass of a plasma electron macro-particle: h^2 / ppc (electron mass 1)."""
    return cfg.getfloat('transverse-step') ** 2 / cfg.getfloat('plasma-particles-per-cell')


def thermal_momenta(m: float, shape, seed: int = 2, sigma: float = 1e-3) -> dict:
    """Seeded thermal momenta of macro-particles of mass m: p = m u with
    u ~ N(0, sigma) per component (sigma in units of m c)."""
    rng = np.random.default_rng(seed)
    return {k: m * rng.normal(0.0, sigma, shape) for k in ('px', 'py', 'pz')}


def thermalize(particles: dict, m: float, seed: int = 2, sigma: float = 1e-3) -> None:
    """Draws the electrons' momenta (the only mobile sort of the workload)."""
    import torch
    arrs = particles['electrons']
    for k, p in thermal_momenta(m, tuple(arrs.m.shape), seed, sigma).items():
        t = getattr(arrs, k)
        t.copy_(torch.as_tensor(p, dtype=t.dtype, device=t.device))


class Run:
    """One precision form of the workload, stepped one xi step at a time."""

    def __init__(self, precision: str, n_beam: int, field_solver: str = 'auto'):
        from lcodepy_amd.beam3d.beam_io import BeamDrain, BeamSource
        from lcodepy_amd.config import Config
        from lcodepy_amd.plasma3d.initialization import init_plasma
        from lcodepy_amd.push_solver import PusherAndSolver3D
        cfg = Config(workload_config(precision, field_solver))
        cfg._adjust_config_values_3d()
        self.cfg = cfg
        self.N = cfg.getint('window-width-steps')
        fields, particles, currents, const, xi0 = init_plasma(cfg)
        thermalize(particles, electron_mass(cfg))
        self.n_plasma = sum(int(a.m.numel()) for a in particles.values())
        self.solver = PusherAndSolver3D(cfg)
        self.solver.verbose = False
        beam = driver_beam(n_beam, cfg.getfloat('xi-step'))
        self.it = self.solver.iter_dxi(fields, particles, currents, const, xi0,
                                       BeamSource(cfg, beam), BeamDrain(cfg), 1.0)

    def steps(self, k: int) -> None:
        for _ in range(k):
            next(self.it)

    def ez(self) -> np.ndarray:
        return np.array([float(v) for v in self.solver.ez_axis])


def time_steps(run: Run, warmup: int, steps: int) -> float:
    import torch
    run.steps(warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    run.steps(steps)
    torch.cuda.synchronize()
    return time.perf_counter() - t0


def live_kernel_times(run: Run, steps: int) -> dict:
    """Average device seconds per launch of the push, the deposit gather and
    the solve GEMM over `steps` further xi steps, from HIP events the library
    records on the launch stream around those kernels."""
    ctx = run.solver.plasma_solver.ctx
    ctx.kernel_timing(True)
    run.steps(steps)
    times = ctx.kernel_times()
    ctx.kernel_timing(False)
    return times


def step_bytes(n: int, N: int, precision: str) -> int:
    """Algorithmic HBM bytes of one xi step, B_step = Np b_p + 166 N^2 b_g:
    46 words per particle, 368 B in float64 or 272 B in FP32-bulk (x_offt,
    y_offt, x_init, y_init and pz float64, the rest float32), and 166 words
    per grid point of 8 or 4 B."""
    if precision == 'fp64':
        return n * WORDS_PER_PARTICLE * 8 + WORDS_PER_GRID_POINT * N * N * 8
    return n * 272 + WORDS_PER_GRID_POINT * N * N * 4


def solve_bytes(N: int, precision: str) -> int:
    """Algorithmic bytes of one field solve: 33 N^2 words of inputs and
    outputs and 7 fields x 5 N^2 of transform passes (68 N^2 words)."""
    return SOLVE_WORDS_PER_GRID_POINT * N * N * (8 if precision == 'fp64' else 4)


def gemm_flops(N: int) -> float:
    """Flops of one solve-GEMM launch: the seven N x N products (one per
    field) of a transform pass, 2 N^3 each."""
    return 7 * 2.0 * N ** 3


def hbm_traffic(precision: str, kernel_class: str):
    """HBM bytes per launch of a kernel class from the rocprofv3 FETCH_SIZE
    and WRITE_SIZE passes (scaled by a copy of known size), when summarised."""
    key = f'traffic_bytes_per_launch {precision} {kernel_class}:'
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                if line.startswith(key):
                    return float(line.split(':', 1)[1])
    except OSError:
        pass
    return None


def deposit_bytes(n: int, N: int, bulk: int) -> int:
    """Algorithmic bytes of one deposit not fused with its push
    (weights.py:35-67): the nine particle arrays read (x_init y_init x_offt
    y_offt pz float64, px py q m of the bulk type) and the four current grids
    written.  The timed deposit is the whole cell sort and gather."""
    return n * (5 * 8 + 4 * bulk) + 4 * N * N * bulk


def push_bytes(n: int, N: int, bulk: int) -> float:
    """Algorithmic bytes of one smart-push launch (plasma3d/move.py:52-148),
    averaged over the three pushes of an xi step.  Pushes 2 and 3 read
    x_init y_init x_offt_prev y_offt_prev x_offt y_offt pz_prev (float64) and
    px_prev py_prev q m (bulk type) and write x_offt y_offt pz (float64) and
    px py (bulk type), and read the six half-step fields (bulk type); push 1
    (with the free-streaming estimate) reads the previous layer only (4
    float64 + pz + 4 bulk), writes also q, and reads the float64 fields."""
    later = n * (7 * 8 + 4 * bulk + 3 * 8 + 2 * bulk) + 6 * N * N * bulk
    first = n * (5 * 8 + 4 * bulk + 3 * 8 + 3 * bulk) + 6 * N * N * 8
    return (first + 2 * later) / 3


def _oracle_sample(p, cfg, steps: int, start_step: int) -> float:
    """Seconds of `steps` xi steps of the oracle's xi loop (step_dt) from
    xi step start_step + 1 on, with the driver's layers deposited and moved:
    the plasma state is the initial one with thermal momenta, the beam the
    part of the driver behind the start."""
    from oracle import lcode_oracle as O
    fields, particles, currents, const, _ = O.init_plasma(p)
    e = particles['electrons']
    e.update(thermal_momenta(electron_mass(cfg), e['m'].shape))
    beam = O.beam_sorted(O.beam_from_records(driver_beam(1_000_000, p.xi_step)))
    ahead = beam['xi'] >= -start_step * p.xi_step
    beam = {k: v[~ahead] for k, v in beam.items()}
    orig = O.compute_rhoj

    def serial_rhoj(pp, particles_, const_, nthreads=None):
        return orig(pp, particles_, const_, nthreads=1)

    O.compute_rhoj = serial_rhoj
    try:
        t0 = time.perf_counter()
        O.step_dt(p, fields, particles, currents, const, -start_step * p.xi_step, beam,
                  n_steps=steps)
        return time.perf_counter() - t0
    finally:
        O.compute_rhoj = orig


def cpu_baseline(sample_steps: int, one_core_steps: int) -> dict:
    """The oracle (float64 numpy + C, this repository's CPU restatement of
    the reference's Numba path) on the same workload: the xi loop with the
    driver's layers (deposit, substeps, push) and the plasma step, pushes on
    every host core (OpenMP, as the reference's Numba prange), deposit serial
    (as weights.py:35), scipy.fftpack DST/DCT.  Timed over xi steps around
    the driver's centre (xi = -3), on all cores and on one."""
    from lcodepy_amd.config import Config
    from oracle import lcode_oracle as O
    cfg = Config(workload_config('fp64'))
    cfg._adjust_config_values_3d()
    cores = os.cpu_count() or 1
    start = 140
    dt_all = _oracle_sample(O.Params.from_config(cfg, nthreads=cores), cfg, sample_steps, start)
    dt_one = _oracle_sample(O.Params.from_config(cfg, nthreads=1), cfg, one_core_steps, start)
    model = ''
    try:
        with open('/proc/cpuinfo') as f:
            model = next(l.split(':', 1)[1].strip() for l in f if l.startswith('model name'))
    except (OSError, StopIteration):
        pass
    return {'value': sample_steps / dt_all, 'unit': 'xi-steps/s', 'cores': cores, 'kind': 'port',
            'one_core': {'value': one_core_steps / dt_one, 'unit': 'xi-steps/s', 'cores': 1,
                         'steps': one_core_steps},
            'sample': f'{sample_steps} xi steps ({one_core_steps} on one core) of the same '
                      f'513x513, 929,296-particle float64 workload from xi step {start + 1} '
                      f'(driver centre: beam layers deposited and moved), deposit serial, '
                      f'pushes on {cores} threads; {model}'}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=900)
    ap.add_argument('--warmup', type=int, default=10)
    ap.add_argument('--beam-particles', type=int, default=1_000_000)
    ap.add_argument('--timing-steps', type=int, default=None,
                    help='xi steps with per-launch event timing after the timed steps '
                         '(default: the rest of the window, at most 91)')
    ap.add_argument('--cpu-sample-steps', type=int, default=20)
    ap.add_argument('--cpu-one-core-steps', type=int, default=4)
    ap.add_argument('--no-cpu-baseline', action='store_true')
    ap.add_argument('--fp64-only', action='store_true')
    ap.add_argument('--field-solver', default='auto', choices=('auto', 'matrix', 'rfft'),
                    help="the library's field solve: 'auto' (default) times both on this GPU "
                         'and takes the faster')
    ap.add_argument('--dump-ez', default=None,
                    help='write the on-axis Ez of every timed xi step of each form (npz)')
    ap.add_argument('--copy-calibration-mib', type=int, default=256,
                    help='size of the plain device copy run once before the timed steps, '
                         'the reference dispatch for rocprofv3 FETCH_SIZE/WRITE_SIZE')
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get('WORLD_SIZE', '1')) != 1:
        print('bench.py measures one GPU only: the xi step has no multi-GPU path here '
              '(a window is one GPU\'s work; independent time steps or runs are the unit '
              'that spreads over GPUs)', file=sys.stderr)
        return 2
    timing_steps = args.timing_steps
    if timing_steps is None:
        timing_steps = max(1, min(91, WINDOW_STEPS - args.warmup - args.steps))
    if args.warmup + args.steps + timing_steps > WINDOW_STEPS:
        print(f'--warmup + --steps + --timing-steps must fit in the {WINDOW_STEPS} xi steps of '
              'the window', file=sys.stderr)
        return 2
    import torch
    torch.cuda.set_device(0)
    if args.copy_calibration_mib > 0:
        n = args.copy_calibration_mib * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device='cuda')
        dst = torch.empty_like(src)
        torch.mul(src, 1.0, out=dst)   # reads and writes copy_calibration_mib MiB
        torch.cuda.synchronize()
        del src, dst

    results = {}
    forms = ['fp64'] if args.fp64_only else ['fp32-bulk', 'fp64']
    for form in forms:
        print(f'[bench] {form}: building the workload', file=sys.stderr, flush=True)
        run = Run(form, args.beam_particles, args.field_solver)
        elapsed = time_steps(run, args.warmup, args.steps)
        ez = run.ez()
        kt = live_kernel_times(run, timing_steps)
        ctx = run.solver.plasma_solver.ctx
        ctx.raise_on_errors()
        solve = {'field_solver': ctx.field_solver, 'fft_radix': ctx.fft_radix,
                 'rfft_cap': ctx.rfft_shape[0], 'rfft_tpb': ctx.rfft_shape[1],
                 'rfft_wave': ctx.rfft_wave}
        results[form] = dict(elapsed=elapsed, ez=ez, n=run.n_plasma, N=run.N, kt=kt, solve=solve)
        print(f'[bench] {form}: {args.steps / elapsed:.1f} xi-steps/s, per launch '
              + ', '.join(f'{k} {v[0] * 1e6:.1f} us x{v[1]}' for k, v in kt.items()),
              file=sys.stderr, flush=True)
        del run
        torch.cuda.empty_cache()

    head = forms[0]
    r = results[head]
    value = args.steps / r['elapsed']
    N, n = r['N'], r['n']
    bulk = 4 if head == 'fp32-bulk' else 8
    solver = r['solve']['field_solver']
    kt = {k: v[0] for k, v in r['kt'].items()}
    launches = {'push': LAUNCHES_PER_STEP['push'], 'deposit': LAUNCHES_PER_STEP['deposit'],
                'solve': LAUNCHES_PER_STEP['solve'][solver]}
    per_step = {k: kt[k] * launches[k] for k in kt}
    dominant = max(per_step, key=per_step.get)

    def entry(k):
        common = {'traffic': hbm_traffic(head, k), 'seconds_per_launch': kt[k],
                  'seconds_per_xi_step': per_step[k]}
        if k == 'solve' and solver == 'matrix':
            fl = gemm_flops(N)
            return {'bound': 'mfma', 'kernel': 'k_gemm_mfma (solve transform pass, 7 fields)',
                    'achieved': fl / kt[k] / 1e12, 'peak': MFMA_PEAK[head] / 1e12,
                    'unit': 'TFLOP/s', 'frac': fl / kt[k] / MFMA_PEAK[head], **common}
        if k == 'solve':
            nbytes = solve_bytes(N, head) / 3
            name = ('k_rf_rows_fwd / k_rf_cols / k_rf_rows_inv (LDS-FFT solve, 3 launches '
                    'per solve, 68 N^2 words per solve)')
        elif k == 'push':
            nbytes = push_bytes(n, N, bulk)
            name = 'k_push (smart push, 3 per xi step)'
        else:
            nbytes = deposit_bytes(n, N, bulk)
            name = 'deposit: cell sort + k_deposit_gather (3 per xi step)'
        return {'bound': 'hbm', 'kernel': name, 'achieved': nbytes / kt[k] / 1e9,
                'peak': HBM_PEAK / 1e9, 'unit': 'GB/s', 'frac': nbytes / kt[k] / HBM_PEAK,
                **common}

    b_step = step_bytes(n, N, head)
    roofline_step = {'bound': 'hbm', 'bytes_per_xi_step': b_step,
                     'achieved': b_step * value / 1e9, 'peak': HBM_PEAK / 1e9, 'unit': 'GB/s',
                     'frac': b_step * value / HBM_PEAK,
                     'model': 'B_step = Np b_p + 166 N^2 b_g (46 words per particle, 166 per '
                              'grid point)'}

    line = {
        'metric': METRIC, 'value': value, 'unit': 'xi-steps/s', 'n_gpus': 1,
        'steps': args.steps, 'warmup': args.warmup,
        'ms_per_step': 1e3 * r['elapsed'] / args.steps, 'higher_is_better': True,
        'scaling': 'weak', 'vs_baseline': None,
        'dtype': 'fp32+fp64' if head == 'fp32-bulk' else 'fp64',
        'data': 'synthetic (seeded plasma lattice with thermal momenta, seeded 1e6-particle '
                'Gaussian driver)',
        'config': {'workload': '513x513 window 25.6, 4 ppc (929,296 plasma macro-particles), '
                               'xi_step 0.02 (1001 xi steps), 1e6-particle driver n_b=1 '
                               'sigma_r=0.5 sigma_xi=1',
                   'precision': head, 'N': N, 'plasma_particles': n,
                   'beam_particles': args.beam_particles,
                   'thermal_momenta': 'normalised: p / (m c) ~ N(0, 1e-3) per component, '
                                      'stored p = m u (read as stored momenta the plasma '
                                      'is thermally chaotic and no precision tracks it)'},
        'particle_pushes_per_s': 3 * n * value,
        'kernel_avg_s': kt,
        'kernel_launches_timed': {k: v[1] for k, v in r['kt'].items()},
        'roofline': entry(dominant),
        'roofline_all': {k: entry(k) for k in kt},
        'roofline_step': roofline_step,
        'field_solve': r['solve'],
    }
    if 'fp64' in results and head != 'fp64':
        r64 = results['fp64']
        v64 = args.steps / r64['elapsed']
        k = min(len(r64['ez']), len(r['ez']))
        ez_err = float(np.max(np.abs(r['ez'][:k] - r64['ez'][:k])) / np.max(np.abs(r64['ez'][:k])))
        line['fp64_xi_steps_per_s'] = v64
        line['fp32_bulk_over_fp64'] = value / v64
        line['on_axis_ez_rel_err_vs_fp64'] = ez_err
        line['on_axis_ez_steps_compared'] = k
        scale = np.max(np.abs(r64['ez'][:k]))
        line['on_axis_ez_rel_err_first_steps'] = {
            str(m): float(np.max(np.abs(r['ez'][:m] - r64['ez'][:m])) / scale)
            for m in (100, 200, 300, 400, 500, 700) if m <= k}
    if args.dump_ez:
        np.savez(args.dump_ez, **{form.replace('-', '_'): res['ez'] for form, res in results.items()})
    if not args.no_cpu_baseline:
        print('[bench] cpu baseline (oracle)', file=sys.stderr, flush=True)
        line['cpu_baseline'] = cpu_baseline(args.cpu_sample_steps, args.cpu_one_core_steps)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == '__main__':
    sys.exit(main())
