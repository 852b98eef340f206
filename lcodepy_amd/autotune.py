"""Field-solve implementation and launch shape chosen by measurement.

``field-solver = 'auto'`` (the default) makes :func:`runtime.context` ask
:func:`solve_choice` which solve to build: ``lc3d_autotune_solve`` times one
``compute_fields`` of the run's geometry and precision for the transform-matrix
products and for the paired LDS FFT over its radix plans, LDS sizes and
transforms per block, and the fastest wins.  Choices are kept per device and
geometry for the life of the process.

With ``LCODEPY_AMD_TUNE_FILE`` set to a path, choices are also kept in that
JSON file (keyed by geometry and precision; the file belongs to one machine)
and read back by later processes, so that, for example, a profiled run
builds the same solve as the run it is compared with.

``python -m lcodepy_amd.autotune --sizes 129 257 513`` prints the decision
table (every candidate's time) as markdown.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import threading

from . import _native

_CHOICES: dict = {}
_LOCK = threading.Lock()
SOLVER_NAMES = {v: k for k, v in _native.SOLVERS.items()}


def _params(N: int, h: float, xi_step: float, subtraction_coeff: float,
            precision: str) -> _native.Params:
    p = _native.Params()
    p.grid_steps = N
    p.grid_step_size = h
    p.xi_step_size = xi_step
    p.boundary = h * (N / 2 - 10)
    p.subtraction_coeff = subtraction_coeff
    p.lost_radius = max(0.9 * h * N / 2, h * N / 2 - 1)
    p.precision = _native.LC3D_FP32_BULK if precision == 'fp32-bulk' else _native.LC3D_FP64
    p.max_particles = 1
    return p


def _as_dict(c: _native.SolveChoice) -> dict:
    solver = SOLVER_NAMES[c.field_solver]
    plan = tuple(c.fft_radix[i] for i in range(c.fft_nstages)) if solver == 'rfft' else None
    wave = int(c.rfft_wave) if solver == 'rfft' else 0
    if wave == 7:
        plan = None
    return {'field_solver': solver, 'fft_radix': plan, 'rfft_cap': int(c.rfft_cap),
            'rfft_tpb': int(c.rfft_tpb), 'rfft_wave': wave, 'seconds': float(c.seconds)}


def tune(N: int, h: float, xi_step: float, subtraction_coeff: float, precision: str,
         repeats: int = 20, table_cap: int = 256):
    """(best, table): the fastest candidate and every candidate timed, as
    dicts (field_solver, fft_radix, rfft_cap, rfft_tpb, rfft_wave, seconds)."""
    L = _native.lib()
    p = _params(N, h, xi_step, subtraction_coeff, precision)
    best = _native.SolveChoice()
    table = (_native.SolveChoice * table_cap)()
    n = ctypes.c_int32(0)
    rc = L.lc3d_autotune_solve(ctypes.byref(p), int(repeats), ctypes.byref(best), table,
                               table_cap, ctypes.byref(n))
    if rc != 0:
        raise RuntimeError(f'lc3d_autotune_solve failed ({rc}) for N={N}, {precision}')
    return _as_dict(best), [_as_dict(table[i]) for i in range(min(n.value, table_cap))]


def solve_choice(device_index: int, N: int, h: float, xi_step: float, subtraction_coeff: float,
                 precision: str) -> dict:
    """The measured fastest solve for this device and geometry (tuned once)."""
    key = (device_index, N, h, xi_step, subtraction_coeff, precision)
    with _LOCK:
        choice = _CHOICES.get(key)
        if choice is None:
            import torch
            path = os.environ.get('LCODEPY_AMD_TUNE_FILE')
            fkey = json.dumps([N, h, xi_step, subtraction_coeff, precision])
            stored = {}
            if path and os.path.exists(path):
                with open(path) as f:
                    stored = json.load(f)
            choice = stored.get(fkey)
            if choice is not None and choice.get('fft_radix') is not None:
                choice['fft_radix'] = tuple(choice['fft_radix'])
            if choice is None:
                with torch.cuda.device(device_index):
                    choice, _ = tune(N, h, xi_step, subtraction_coeff, precision)
                if path:
                    stored[fkey] = choice
                    with open(path, 'w') as f:
                        json.dump(stored, f, indent=1)
            _CHOICES[key] = choice
    return choice


def _describe(c: dict) -> str:
    if c['field_solver'] == 'matrix':
        return 'matrix (MFMA transform products)'
    wave = int(c.get('rfft_wave') or 0)
    if wave == 7:
        return 'rfft by waves (one transform per wave in registers, 4 per block)'
    plan = 'x'.join(str(r) for r in c['fft_radix'])
    text = f"rfft plan {plan}, LDS {c['rfft_cap']} complex, {c['rfft_tpb']} transforms/block"
    if wave:
        names = [k for b, k in ((1, 'rows fwd'), (2, 'cols'), (4, 'rows inv')) if wave & b]
        text += ', by waves: ' + ' + '.join(names)
    return text


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split('\n\n')[0])
    ap.add_argument('--sizes', type=int, nargs='+', default=[129, 257, 385, 513, 769, 1025])
    ap.add_argument('--precisions', nargs='+', default=['fp32-bulk', 'fp64'])
    ap.add_argument('--h', type=float, default=0.05)
    ap.add_argument('--repeats', type=int, default=20)
    args = ap.parse_args(argv)
    import torch
    torch.cuda.set_device(0)
    print(f'# Field-solve autotuning on {torch.cuda.get_device_name(0)}\n')
    print('Device time of one compute_fields (7 fields), each candidate timed over '
          f'{args.repeats} calls after 2 warm-up calls.\n')
    print('| N | precision | chosen | chosen us | matrix us | best rfft us | candidates |')
    print('|---|---|---|---|---|---|---|')
    details = []
    for N in args.sizes:
        for prec in args.precisions:
            best, table = tune(N, args.h, 0.02, 1.0, prec, args.repeats)
            mat = [c['seconds'] for c in table if c['field_solver'] == 'matrix' and c['seconds'] > 0]
            rf = [c['seconds'] for c in table if c['field_solver'] == 'rfft' and c['seconds'] > 0]
            print(f"| {N} | {prec} | {_describe(best)} | {best['seconds'] * 1e6:.1f} | "
                  f"{mat[0] * 1e6 if mat else float('nan'):.1f} | "
                  f"{min(rf) * 1e6 if rf else float('nan'):.1f} | {len(table)} |", flush=True)
            details.append((N, prec, table))
    for N, prec, table in details:
        print(f'\n## N = {N}, {prec}\n')
        print('| candidate | us |')
        print('|---|---|')
        for c in sorted(table, key=lambda c: c['seconds'] if c['seconds'] > 0 else 1e9):
            us = f"{c['seconds'] * 1e6:.1f}" if c['seconds'] > 0 else 'refused'
            print(f'| {_describe(c)} | {us} |')
    return 0


if __name__ == '__main__':
    sys.exit(main())
