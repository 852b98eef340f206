"""ctypes binding of ``liblcodepy_amd.so`` (the C ABI in include/lcodepy_amd.h).

The library is built in-tree by ``__graft_entry__.build()`` (or
``make -C lcodepy_amd/csrc``).  There is no fallback: if the library or a GPU
is missing, :func:`lib` raises.
"""
from __future__ import annotations

import ctypes
import os
import threading

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, 'liblcodepy_amd.so')

LC3D_FP64 = 0
LC3D_FP32_BULK = 1
SOLVERS = {'matrix': 0, 'fft': 1, 'rfft': 2}

ERR_BITS = {
    1: 'plasma particle stencil left the grid (not deposited)',
    2: 'non-finite plasma particle position (not deposited)',
    4: 'beam particle stencil partly off the grid or substep cap reached',
    8: 'more than 16384 plasma particles in one cell (summed in arrival order, so that '
       'cell is reproducible only to rounding)',
    16: 'beam particle charge contribution of 2^40 or more (not deposited)',
}
# Bits that mean results were lost; the others are reported as warnings.
FATAL_BITS = 1 | 2 | 4 | 16

_vp = ctypes.c_void_p
_dp = ctypes.c_void_p   # device pointers are passed as integers


class Params(ctypes.Structure):
    _fields_ = [('grid_steps', ctypes.c_int32),
                ('grid_step_size', ctypes.c_double),
                ('xi_step_size', ctypes.c_double),
                ('boundary', ctypes.c_double),
                ('subtraction_coeff', ctypes.c_double),
                ('lost_radius', ctypes.c_double),
                ('precision', ctypes.c_int32),
                ('max_particles', ctypes.c_int64),
                ('fft_nstages', ctypes.c_int32),
                ('fft_radix', ctypes.c_int32 * 12),
                ('eigenvalues', ctypes.POINTER(ctypes.c_double)),
                ('field_solver', ctypes.c_int32),
                ('rfft_cap', ctypes.c_int32),
                ('rfft_tpb', ctypes.c_int32),
                ('declustering_side', ctypes.c_int32),
                ('declustering_ntaps', ctypes.c_int32),
                ('declustering_tap0', ctypes.c_int32),
                ('declustering_taps', ctypes.POINTER(ctypes.c_double)),
                ('declustering_force', ctypes.c_double),
                ('declustering_damping', ctypes.c_double),
                ('declustering_limit', ctypes.c_double),
                ('rfft_wave', ctypes.c_int32)]


class Fields6(ctypes.Structure):
    _fields_ = [(k, _vp) for k in ('Ex', 'Ey', 'Ez', 'Bx', 'By', 'Bz')]


class Fields(ctypes.Structure):
    _fields_ = [(k, _vp) for k in ('Ex', 'Ey', 'Ez', 'Bx', 'By', 'Bz', 'Phi')]


class Currents(ctypes.Structure):
    _fields_ = [(k, _vp) for k in ('ro', 'jx', 'jy', 'jz')]


class SolveChoice(ctypes.Structure):
    _fields_ = [('field_solver', ctypes.c_int32),
                ('fft_nstages', ctypes.c_int32),
                ('fft_radix', ctypes.c_int32 * 12),
                ('rfft_cap', ctypes.c_int32),
                ('rfft_tpb', ctypes.c_int32),
                ('seconds', ctypes.c_double),
                ('rfft_wave', ctypes.c_int32)]


PROBE_MAX_SOURCES = 16


class ProbeSources(ctypes.Structure):
    _fields_ = [('n', ctypes.c_int32),
                ('src', _vp * PROBE_MAX_SOURCES),
                ('is_f32', ctypes.c_int32 * PROBE_MAX_SOURCES)]


class Sort(ctypes.Structure):
    _fields_ = [('n', ctypes.c_int64),
                ('x_init', _vp), ('y_init', _vp), ('m', _vp),
                ('x_offt_prev', _vp), ('y_offt_prev', _vp),
                ('px_prev', _vp), ('py_prev', _vp), ('q_prev', _vp), ('pz_prev', _vp),
                ('x_offt', _vp), ('y_offt', _vp),
                ('px', _vp), ('py', _vp), ('q', _vp), ('pz', _vp),
                ('column', ctypes.c_int32),
                ('dx_chaotic_prev', _vp), ('dy_chaotic_prev', _vp),
                ('dx_chaotic', _vp), ('dy_chaotic', _vp)]


SYMBOLS = ('lc3d_create', 'lc3d_destroy', 'lc3d_last_error', 'lc3d_precision',
           'lc3d_error_flags', 'lc3d_kernel_timing', 'lc3d_kernel_times', 'lc3d_move_wo_fields', 'lc3d_move_smart', 'lc3d_deposit',
           'lc3d_compute_fields', 'lc3d_plasma_step', 'lc3d_deposit_beam',
           'lc3d_layout_beam_layer',
           'lc3d_beam_substep', 'lc3d_move_beam', 'lc3d_energy_flux', 'lc3d_probe_sample',
           'lc3d_abs_max', 'lc3d_energy_flux_noalloc', 'lc3d_autotune_solve',
           'lc3d_noise_filter', 'lc3d_beam_split')

_LIB = None
_LOCK = threading.Lock()


def load_library(path: str = LIB_PATH) -> ctypes.CDLL:
    """dlopen the library and declare the signatures (no GPU needed)."""
    if not os.path.exists(path):
        raise RuntimeError(f'{path} is missing: build it with __graft_entry__.build() '
                           'or make -C lcodepy_amd/csrc (there is no CPU fallback)')
    L = ctypes.CDLL(path)
    i32, i64, dbl = ctypes.c_int32, ctypes.c_int64, ctypes.c_double
    P = ctypes.POINTER
    L.lc3d_create.argtypes = [P(Params), P(_vp)]
    L.lc3d_create.restype = ctypes.c_int
    L.lc3d_destroy.argtypes = [_vp]
    L.lc3d_destroy.restype = None
    L.lc3d_last_error.argtypes = [_vp]
    L.lc3d_last_error.restype = ctypes.c_char_p
    L.lc3d_precision.argtypes = [_vp]
    L.lc3d_error_flags.argtypes = [_vp, P(ctypes.c_uint32), _vp]
    L.lc3d_kernel_timing.argtypes = [_vp, i32]
    L.lc3d_kernel_times.argtypes = [_vp, P(dbl), P(i64)]
    L.lc3d_move_wo_fields.argtypes = [_vp, i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp]
    L.lc3d_move_smart.argtypes = [_vp, P(Sort), P(Fields6), i32, _vp]
    L.lc3d_deposit.argtypes = [_vp, P(Sort), _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]
    L.lc3d_compute_fields.argtypes = [_vp, P(Fields), i32, P(Fields), _vp, _vp, P(Currents),
                                      P(Currents), P(Fields), P(Fields), _vp]
    L.lc3d_plasma_step.argtypes = [_vp, i32, P(Sort), P(Fields), P(Currents), _vp, _vp, _vp,
                                   P(Fields), P(Fields), P(Currents), _vp]
    L.lc3d_deposit_beam.argtypes = [_vp, i64, i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp]
    L.lc3d_layout_beam_layer.argtypes = [_vp, i64, i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp]
    L.lc3d_beam_substep.argtypes = [_vp, i64, _vp, _vp, dbl, dbl, _vp, _vp, _vp]
    L.lc3d_energy_flux.argtypes = [_vp, P(Fields6), i64, _vp, _vp, _vp, _vp, _vp, _vp]
    L.lc3d_energy_flux_noalloc.argtypes = [i32, dbl, i32, P(Fields6), i64, _vp, _vp, _vp, _vp,
                                           _vp, _vp]
    L.lc3d_autotune_solve.argtypes = [P(Params), i32, P(SolveChoice), P(SolveChoice), i32,
                                      P(i32)]
    L.lc3d_noise_filter.argtypes = [_vp, P(Sort), _vp]
    L.lc3d_beam_split.argtypes = [_vp, i64, i32, i32, _vp, _vp, _vp, _vp, _vp, _vp, _vp]
    L.lc3d_probe_sample.argtypes = [_vp, P(ProbeSources), i32, _vp, _vp, _vp, _vp]
    L.lc3d_abs_max.argtypes = [_vp, _vp, i64, i64, i64, i64, i64, i64, _vp, _vp]
    L.lc3d_move_beam.argtypes = [_vp, i64, i64, _vp, _vp, P(Fields6), P(Fields6), _vp, _vp,
                                 _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]
    for name in SYMBOLS:
        if name not in ('lc3d_destroy', 'lc3d_last_error'):
            getattr(L, name).restype = ctypes.c_int
    return L


def lib() -> ctypes.CDLL:
    global _LIB
    with _LOCK:
        if _LIB is None:
            _LIB = load_library()
    return _LIB


def ptr(t) -> int | None:
    """Device address of a torch tensor (None for None)."""
    if t is None:
        return None
    if not t.is_cuda:
        raise ValueError('lcodepy_amd kernels need device tensors')
    if not t.is_contiguous():
        raise ValueError('lcodepy_amd kernels need C-contiguous tensors')
    return t.data_ptr()


def stream_ptr(stream=None) -> int:
    """The HIP stream handle of ``stream`` or of torch's current stream.  The
    current stream is read with torch's raw-stream query (one C call, where
    ``torch.cuda.current_stream()`` re-checks the device on every call, a
    sizeable part of an xi step's host time)."""
    import torch
    if stream is not None:
        return stream.cuda_stream
    raw = getattr(torch._C, '_cuda_getCurrentRawStream', None)
    if raw is None:
        return torch.cuda.current_stream().cuda_stream
    return raw(torch._C._cuda_getDevice())


class Context:
    """One lc3d context: geometry, precision and workspaces on one device."""

    def __init__(self, N: int, h: float, xi_step: float, boundary: float,
                 subtraction_coeff: float, lost_radius: float, precision: str,
                 max_particles: int, eigenvalues: np.ndarray | None = None,
                 fft_radix: tuple | None = None, field_solver: str = 'matrix',
                 rfft_cap: int = 0, rfft_tpb: int = 0, declustering: dict | None = None,
                 rfft_wave: int = 0):
        import torch
        if not torch.cuda.is_available():
            raise RuntimeError('lcodepy_amd needs a ROCm GPU (torch.cuda.is_available() is False)')
        self.lib = lib()
        self.N = N
        self.h = float(h)
        self.precision = precision
        self.max_particles = int(max_particles)
        self.field_solver = field_solver
        self.fft_radix = tuple(fft_radix) if fft_radix else None
        self.rfft_shape = (int(rfft_cap), int(rfft_tpb))
        # FFT kernels by waves: 1 rows forward, 2 columns, 4 rows inverse
        self.rfft_wave = int(rfft_wave)
        p = Params()
        p.grid_steps = N
        p.grid_step_size = h
        p.xi_step_size = xi_step
        p.boundary = boundary
        p.subtraction_coeff = subtraction_coeff
        p.lost_radius = lost_radius
        p.precision = LC3D_FP32_BULK if precision == 'fp32-bulk' else LC3D_FP64
        p.max_particles = self.max_particles
        if field_solver not in SOLVERS:
            raise ValueError(f'unknown field solver {field_solver!r} (one of {sorted(SOLVERS)})')
        p.field_solver = SOLVERS[field_solver]
        p.rfft_cap = int(rfft_cap)
        p.rfft_tpb = int(rfft_tpb)
        p.rfft_wave = self.rfft_wave
        self.declustering_side = 0
        self._dc_taps = None
        if declustering is not None:
            self._dc_taps = np.ascontiguousarray(declustering['taps'], dtype=np.float64)
            p.declustering_side = self.declustering_side = int(declustering['side'])
            p.declustering_ntaps = len(self._dc_taps)
            p.declustering_tap0 = int(declustering['tap0'])
            p.declustering_taps = self._dc_taps.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
            p.declustering_force = float(declustering['force'])
            p.declustering_damping = float(declustering['damping'])
            p.declustering_limit = float(declustering['limit'])
        if fft_radix:
            p.fft_nstages = len(fft_radix)
            for i, r in enumerate(fft_radix):
                p.fft_radix[i] = r
        self._lam = None
        if eigenvalues is not None:
            self._lam = np.ascontiguousarray(eigenvalues, dtype=np.float64)
            p.eigenvalues = self._lam.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        h_ = _vp()
        rc = self.lib.lc3d_create(ctypes.byref(p), ctypes.byref(h_))
        if rc != 0:
            raise RuntimeError(f'lc3d_create failed with code {rc} (N={N}, precision={precision}, '
                               f'field solver {field_solver}, FFT plan {fft_radix}, rfft shape '
                               f'{rfft_cap}/{rfft_tpb}, kernels by waves {int(rfft_wave)}; the "fft" '
                               'solver is not built into the '
                               'GPU library)')
        self.handle = h_

    def __del__(self):
        h = getattr(self, 'handle', None)
        if h:
            try:
                self.lib.lc3d_destroy(h)
            except Exception:
                pass
            self.handle = None

    def check(self, rc: int, what: str) -> None:
        if rc != 0:
            msg = self.lib.lc3d_last_error(self.handle)
            raise RuntimeError(f'{what} failed ({rc}): {msg.decode() if msg else ""}')

    def error_flags(self, stream=None) -> int:
        v = ctypes.c_uint32(0)
        self.check(self.lib.lc3d_error_flags(self.handle, ctypes.byref(v), stream_ptr(stream)),
                   'lc3d_error_flags')
        return int(v.value)

    KERNEL_CLASSES = ('push', 'deposit', 'solve')

    def kernel_timing(self, enable: bool) -> None:
        """Start (and clear) or stop the event timing of the smart push, the
        whole deposit (cell sort and gather) and the solve's transform
        launches."""
        self.check(self.lib.lc3d_kernel_timing(self.handle, int(bool(enable))),
                   'lc3d_kernel_timing')

    def kernel_times(self) -> dict:
        """{class: (seconds per launch, launches)} since timing was enabled."""
        sec = (ctypes.c_double * 3)()
        cnt = (ctypes.c_int64 * 3)()
        self.check(self.lib.lc3d_kernel_times(self.handle, sec, cnt), 'lc3d_kernel_times')
        return {k: (sec[i], cnt[i]) for i, k in enumerate(self.KERNEL_CLASSES)}

    def raise_on_errors(self, stream=None) -> None:
        flags = self.error_flags(stream)
        if flags & ~FATAL_BITS:
            import warnings
            warnings.warn('lcodepy_amd kernels reported: ' + '; '.join(
                m for b, m in ERR_BITS.items() if flags & b & ~FATAL_BITS))
        if flags & FATAL_BITS:
            msgs = [m for b, m in ERR_BITS.items() if flags & b & FATAL_BITS]
            raise FloatingPointError('lcodepy_amd kernels reported: ' + '; '.join(msgs))
