"""lc3d_energy_flux_noalloc, the allocation-free form of lc3d_energy_flux
(partial sums in the library's static device array instead of a context
workspace), through the C ABI: two calls in a row, the second on a smaller
window and fewer particles, each equal bit for bit to lc3d_energy_flux on the
same inputs; sizes beyond the limits are refused with a message."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip('torch')


def _inputs(rng, N, n, f32):
    dev = lambda a, dt=torch.float64: torch.as_tensor(a).to(device='cuda', dtype=dt)
    fields = {k: dev(0.1 * rng.standard_normal((N, N))) for k in ('Ex', 'Ey', 'Ez', 'Bx', 'By', 'Bz')}
    bulk = torch.float32 if f32 else torch.float64
    parts = {'m': dev(np.full(n, 0.0025), bulk), 'px': dev(1e-3 * rng.standard_normal(n), bulk),
             'py': dev(1e-3 * rng.standard_normal(n), bulk), 'pz': dev(1e-3 * rng.standard_normal(n))}
    return fields, parts


def _f6(fields):
    from lcodepy_amd import _native
    return _native.Fields6(*(fields[k].data_ptr() for k in ('Ex', 'Ey', 'Ez', 'Bx', 'By', 'Bz')))


@pytest.mark.parametrize('precision', ['fp64', 'fp32-bulk'])
def test_noalloc_energy_flux_matches_the_context_entry(precision):
    from lcodepy_amd import _native
    f32 = precision == 'fp32-bulk'
    prec = _native.LC3D_FP32_BULK if f32 else _native.LC3D_FP64
    rng = np.random.default_rng(5)
    h = 0.05
    L = _native.lib()
    for N, n in ((513, 929_296), (129, 9_409)):   # the second call smaller than the first
        print(f'energy flux noalloc {precision} N={N} n={n}', flush=True)
        ctx = _native.Context(N=N, h=h, xi_step=0.02, boundary=h * (N / 2 - 10),
                              subtraction_coeff=1.0, lost_radius=1.0, precision=precision,
                              max_particles=n)
        fields, parts = _inputs(rng, N, n, f32)
        f6 = _f6(fields)
        out = torch.full((2,), float('nan'), dtype=torch.float64, device='cuda')
        args = (n, parts['m'].data_ptr(), parts['px'].data_ptr(), parts['py'].data_ptr(),
                parts['pz'].data_ptr())
        stream = torch.cuda.current_stream().cuda_stream
        ctx.check(L.lc3d_energy_flux(ctx.handle, ctypes.byref(f6), *args, out[0:1].data_ptr(),
                                     stream), 'lc3d_energy_flux')
        rc = L.lc3d_energy_flux_noalloc(N, h, prec, ctypes.byref(f6), *args, out[1:2].data_ptr(),
                                        stream)
        assert rc == 0, L.lc3d_last_error(None)
        torch.cuda.synchronize()
        ref, got = out.cpu().numpy()
        print(f'  context entry {ref:.17g}, allocation-free {got:.17g}')
        assert np.isfinite(ref) and ref > 0
        assert got == ref
        del ctx


def test_noalloc_energy_flux_refuses_sizes_beyond_its_limits():
    from lcodepy_amd import _native
    L = _native.lib()
    x = torch.zeros((9, 9), dtype=torch.float64, device='cuda')
    f6 = _native.Fields6(*([x.data_ptr()] * 6))
    psi = torch.zeros(1, dtype=torch.float64, device='cuda')
    for N, n in ((1026, 0), (0, 0), (9, -1), (9, 1 << 31)):
        rc = L.lc3d_energy_flux_noalloc(N, 0.1, _native.LC3D_FP64, ctypes.byref(f6), n, None, None,
                                        None, None, psi.data_ptr(), None)
        assert rc == 1
        assert b'energy_flux_noalloc' in L.lc3d_last_error(None)
