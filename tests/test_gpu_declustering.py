"""The declustering noise filter (plasma3d/noise_filter.py:82-238) on the GPU
against the oracle's restatement on identical seeded inputs: the factory's
noise_filter alone (kicked px and py, the new chaotic displacements), with
and without damping, in both precision forms, and the plasma step with the
filter after its last push (plasma3d/solver.py:69-73) over 20 xi steps.
Tolerances: float64 1e-12 of each array's largest value for one filter call
(the GPU smooths by a direct circular convolution, the reference by FFTs),
1e-11 after 20 steps; FP32-bulk px, py (float32) 1e-6."""
import numpy as np
import pytest

from oracle import lcode_oracle as O
from tests.test_gpu_kernels import _cfg, _gauss_beam, _params, dev, host, relerr

pytestmark = pytest.mark.gpu
torch = pytest.importorskip('torch')


def _decl_cfg(N, precision, damping):
    cfg = _cfg(N, precision=precision)
    cfg.set('declustering-enabled', True)
    cfg.set('declustering-averaging', 3)
    cfg.set('declustering-force', 0.3)
    cfg.set('declustering-limit', 1e-3)
    if damping:
        cfg.set('damping-declustering', damping)
    return cfg


@pytest.mark.parametrize('precision,damping', [('fp64', 0.0), ('fp64', 0.6), ('fp32-bulk', 0.6)])
def test_noise_filter_matches_oracle(precision, damping):
    from lcodepy_amd.plasma3d.data import Arrays
    from lcodepy_amd.plasma3d.noise_filter import get_noise_filter
    N = 65
    print(f'noise filter N={N} {precision} damping {damping}', flush=True)
    cfg = _decl_cfg(N, precision, damping)
    p = _params(N)
    p.declustering = O.Declustering(averaging=3, force=0.3, damping=damping, limit=1e-3)
    el = O.uniform_plasma(p)
    n = el['x_offt'].shape[0]
    rng = np.random.default_rng(11)
    el['x_offt'] = 3e-4 * rng.standard_normal((n, n))
    el['y_offt'] = 3e-4 * rng.standard_normal((n, n))
    el['x_offt'][rng.random((n, n)) < 0.02] = 0.0      # a few zero offsets switch terms off
    f32 = precision == 'fp32-bulk'
    r32 = (lambda a: a.astype(np.float32).astype(np.float64)) if f32 else (lambda a: a)
    for k in ('px', 'py', 'pz'):
        el[k] = r32(1e-3 * rng.standard_normal((n, n)))
    prev = {'dx_chaotic': 1e-4 * rng.standard_normal((n, n)),
            'dy_chaotic': 1e-4 * rng.standard_normal((n, n))}
    ref = O.noise_filter(p, {k: v.copy() for k, v in el.items()}, prev)
    before = {k: el[k].copy() for k in ('px', 'py')}

    bulk = torch.float32 if f32 else torch.float64
    parts = Arrays.wrap(**{k: dev(v, bulk if k in ('px', 'py', 'q', 'm') else torch.float64)
                           for k, v in el.items()})
    parts_prev = Arrays.wrap(dx_chaotic=dev(prev['dx_chaotic']), dy_chaotic=dev(prev['dy_chaotic']))
    out = get_noise_filter(cfg)(parts, parts_prev)
    torch.cuda.synchronize()
    err = {k: relerr(host(getattr(out, k)), ref[k]) for k in ('px', 'py', 'dx_chaotic', 'dy_chaotic')}
    kick = {k: relerr(host(getattr(out, k)) - before[k], ref[k] - before[k]) for k in ('px', 'py')}
    print('  errors', {k: f'{v:.1e}' for k, v in err.items()},
          'kicks', {k: f'{v:.1e}' for k, v in kick.items()})
    assert np.max(np.abs(ref['px'] - before['px'])) > 1e-7    # the filter did act
    assert err['dx_chaotic'] < 1e-12 and err['dy_chaotic'] < 1e-12
    bound = 1e-6 if f32 else 1e-12
    assert err['px'] < bound and err['py'] < bound
    if not f32:
        assert kick['px'] < 1e-10 and kick['py'] < 1e-10


@pytest.mark.parametrize('precision', ['fp64', 'fp32-bulk'])
def test_plasma_step_with_declustering_matches_oracle(precision):
    from lcodepy_amd.plasma3d.initialization import init_plasma
    from lcodepy_amd.plasma3d.solver import Plane2d3vPlasmaSolver
    N = 65
    print(f'plasma step with declustering N={N} {precision}', flush=True)
    cfg = _decl_cfg(N, precision, 0.4)
    p = _params(N)
    p.declustering = O.Declustering(averaging=3, force=0.3, damping=0.4, limit=1e-3)
    fields_o, parts_o, cur_o, const_o, _ = O.init_plasma(p)
    fields_g, parts_g, cur_g, const_g, _ = init_plasma(cfg)
    solver = Plane2d3vPlasmaSolver(cfg)
    rb_prev = np.zeros((N, N))
    for step in range(20):
        rb_full = _gauss_beam(N, p.h, -(step + 1) * p.xi_step, amp=-0.5)
        parts_o, fields_o, cur_o = O.step_dxi(p, parts_o, fields_o, cur_o, const_o, rb_full, rb_prev)
        parts_g, fields_g, cur_g = solver.step_dxi(parts_g, fields_g, cur_g, const_g,
                                                   dev(rb_full), dev(rb_prev))
        rb_prev = rb_full
    torch.cuda.synchronize()
    solver.ctx.raise_on_errors()
    el_o, el_g = parts_o['electrons'], parts_g['electrons']
    err = {k: relerr(host(getattr(fields_g, k)), fields_o[k]) for k in ('Ex', 'Ez', 'Bx', 'Bz')}
    err.update({k: relerr(host(getattr(el_g, k)), el_o[k])
                for k in ('x_offt', 'y_offt', 'px', 'py', 'pz', 'dx_chaotic', 'dy_chaotic')})
    print('  after 20 steps', {k: f'{v:.1e}' for k, v in err.items()})
    assert np.max(np.abs(el_o['dx_chaotic'])) > 0
    bound = 1e-4 if precision == 'fp32-bulk' else 1e-11
    for k, v in err.items():
        assert v < bound, (k, v)
