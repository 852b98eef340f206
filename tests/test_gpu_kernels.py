"""HIP kernels against the float64 oracle on identical seeded inputs.

Each test calls the product through its Python factories, which reach the
kernels through the C ABI (liblcodepy_amd.so).  Tolerances (float64 form):
max |gpu - oracle| / max |oracle| per array below 1e-12 for a single solve,
push or deposit (the GPU fuses multiply-adds and sums in another order), and
below 1e-9 after a few tens of xi steps."""
import math

import numpy as np
import pytest

from oracle import lcode_oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip('torch')


def _cfg(N, h=0.1, xi_step=0.05, precision='fp64', ion_model='background', fineness=2, c=1.0):
    from lcodepy_amd.config import Config
    cfg = Config({'geometry': '3d', 'processing-unit-type': 'mi355x', 'transverse-step': h,
                  'xi-step': xi_step, 'ion-model': ion_model, 'precision': precision,
                  'window-length': 1.0, 'time-step': 1.0,
                  'field-solver-subtraction-coefficient': c})
    cfg.set('window-width-steps', N)
    cfg.set('window-width', N * h)
    cfg.set('plasma-fineness', fineness)
    cfg.set('plasma-particles-per-cell', fineness ** 2)
    return cfg


def _params(N, h=0.1, xi_step=0.05, c=1.0, fineness=2, ion_model='background'):
    return O.Params(N=N, h=h, xi_step=xi_step, subtraction_coeff=c, fineness=fineness,
                    ion_model=ion_model, nthreads=0)


def dev(a, dtype=None):
    t = torch.as_tensor(np.ascontiguousarray(a))
    return t.to(device='cuda', dtype=dtype or t.dtype)


def host(t):
    return t.detach().double().cpu().numpy()


def relerr(a, b):
    scale = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / scale) if scale > 0 else float(np.max(np.abs(a)))


def _smooth(rng, N, k=3):
    x = np.linspace(0, 1, N)
    f = np.zeros((N, N))
    for _ in range(k):
        a, b, ph = rng.uniform(1, 4), rng.uniform(1, 4), rng.uniform(0, 6)
        f += rng.normal() * np.sin(a * np.pi * x + ph)[:, None] * np.cos(b * np.pi * x - ph)[None, :]
    return f


def _sources(rng, N):
    cur = {}
    for k in ('ro', 'jx', 'jy', 'jz'):
        a = np.zeros((2, N, N))
        a[0, 3:-3, 3:-3] = 0.1 * rng.standard_normal((N - 6, N - 6))
        a[1, 3:-3, 3:-3] = 0.01 * rng.standard_normal((N - 6, N - 6))
        cur[k] = a
    return cur


@pytest.mark.skip(reason='the LDS-FFT field solve raised GPU memory faults at N=65 and N=129 '
                         '(it passed at N=33); no cause was found in its code, IR or ISA and its '
                         'CPU/ASan build is clean, so it is not built into the GPU library and '
                         'its GPU tests stay off until a cause is shown')
def test_fft_field_solver_on_gpu():
    pass


FP32_BOUND = 1e-5   # FP32-bulk kernels against the float64 oracle


_SOLVE_CASES = ([(n, 'fp64', 'matrix') for n in (33, 65, 97, 113, 129, 321, 353, 417, 513)]
                + [(n, 'fp32-bulk', 'matrix') for n in (65, 129, 513)]
                + [(n, 'fp64', 'rfft') for n in (33, 65, 97, 113, 129, 321, 353, 417, 513, 769,
                                                 1025)]
                + [(n, 'fp32-bulk', 'rfft') for n in (65, 129, 513, 1025)])


@pytest.mark.parametrize('N,precision,solver', _SOLVE_CASES)
def test_compute_fields_matches_oracle(N, precision, solver, extra=None):
    """One field solve with random smooth predictor, previous fields, beam
    densities and currents of both columns, by the transform-matrix products
    ('matrix') or the LDS FFTs ('rfft'; 2(N-1) = 192, 224, 640, 704, 832 and
    1536 take the radix-3, 7, 5, 11, 13 and 3 stages).  FP32-bulk (currents,
    predictor and solve in float32; N - 1 a multiple of 64 at N = 65, 129
    and 513, so the transform products take their last-row/column edge sums)
    against the float64 oracle on the same float32-rounded inputs."""
    from lcodepy_amd.plasma3d.data import Arrays
    from lcodepy_amd.plasma3d.fields import get_field_computer
    print(f'compute_fields N={N} {precision} {solver} {extra or ""}', flush=True)
    rng = np.random.default_rng(N)
    cfg = _cfg(N, precision=precision)
    cfg.set('field-solver', solver)
    for k, v in (extra or {}).items():
        cfg.set(k, v)
    p = _params(N)
    f32 = precision == 'fp32-bulk'
    const = O.Const(ro_initial=np.zeros((N, N)), dirichlet=O.dirichlet_matrix(N, p.h),
                    mixed=O.mixed_matrix(N, p.h, p.subtraction_coeff),
                    neumann=O.neumann_matrix(N, p.h))
    fk = ('Ex', 'Ey', 'Ez', 'Bx', 'By', 'Bz', 'Phi')
    r32 = (lambda a: a.astype(np.float32).astype(np.float64)) if f32 else (lambda a: a)
    pred = {k: r32(0.01 * _smooth(rng, N)) for k in fk}
    prev = {k: 0.01 * _smooth(rng, N) for k in fk}
    cp = {k: r32(v) for k, v in _sources(rng, N).items()}
    cf = {k: r32(v) for k, v in _sources(rng, N).items()}
    rbf, rbp = 0.05 * _smooth(rng, N), 0.05 * _smooth(rng, N)
    full_o, half_o = O.compute_fields(p, pred, prev, const, rbf, rbp, cp, cf)

    compute = get_field_computer(cfg)
    low = torch.float32 if f32 else torch.float64
    wrap = lambda d, dt=torch.float64: Arrays.wrap(**{k: dev(v, dt) for k, v in d.items()})
    full_g, half_g = compute(wrap(pred, low), wrap(prev), None, dev(rbf), dev(rbp),
                             wrap(cp, low), wrap(cf, low))
    torch.cuda.synchronize()
    # float32 rounding of the length-2048 transforms and the Neumann solve
    # of Bz reaches 1.06e-5 at N = 1025
    bound = (FP32_BOUND if N <= 513 else 3e-5) if f32 else 1e-12
    worst = 0.0
    for k in fk:
        ef = relerr(host(getattr(full_g, k)), full_o[k])
        eh = relerr(host(getattr(half_g, k)), half_o[k])
        worst = max(worst, ef, eh)
        assert ef < bound, ('full', k, ef)
        assert eh < bound, ('half', k, eh)
    print(f'compute_fields N={N} {precision} {solver}: worst {worst:.1e}')


@pytest.mark.parametrize('N,precision,plan,cap,tpb', [
    (513, 'fp64', (2,) * 10, 2048, 1),
    (513, 'fp64', (4, 4, 4, 4, 4), 4096, 1),
    (513, 'fp32-bulk', (8, 8, 16), 4096, 1),
    (513, 'fp32-bulk', (16, 4, 16), 8192, 3),
    (65, 'fp64', (2, 16, 4), 1024, 3),
    (65, 'fp32-bulk', (8, 16), 2048, 7),
    (97, 'fp64', (3, 4, 16), 4096, 5),
    (321, 'fp32-bulk', (5, 2, 4, 16), 8192, 6),
])
def test_rfft_plans_and_launch_shapes(N, precision, plan, cap, tpb):
    """The LDS-FFT solve with explicit radix plans (every radix in several
    stage orders), LDS sizes and transforms per block, including a last
    block with fewer transforms than the others, against the oracle."""
    test_compute_fields_matches_oracle(N, precision, 'rfft',
                                       extra={'fft-radix-plan': list(plan), 'rfft-cap': cap,
                                              'rfft-tpb': tpb})


@pytest.mark.parametrize('N,precision', [(257, 'fp64'), (257, 'fp32-bulk'), (513, 'fp64'),
                                         (513, 'fp32-bulk'), (1025, 'fp64'),
                                         (1025, 'fp32-bulk')])
def test_rfft_by_waves(N, precision):
    """The LDS-FFT solve with each transform done by one wave in registers
    (lengths 2(N-1) = 512, 1024, 2048: radix 8 x 8 x 8, 16 x 16 x 4,
    32 x 32 x 2), against the oracle."""
    test_compute_fields_matches_oracle(N, precision, 'rfft', extra={'rfft-wave': True})


@pytest.mark.parametrize('N,precision,mask,plan,cap,tpb', [
    (513, 'fp32-bulk', 2, (16, 16, 4), 4096, 2),
    (513, 'fp64', 5, (8, 8, 8, 2), 2048, 1),
    (257, 'fp32-bulk', 3, (8, 8, 8), 2048, 1),
    (257, 'fp64', 6, (4, 4, 4, 8), 4096, 3),
])
def test_rfft_kernels_mixed(N, precision, mask, plan, cap, tpb):
    """Some of the three FFT kernels by waves (mask: 1 rows forward,
    2 columns, 4 rows inverse), the others by the Stockham plan, against the
    oracle: the kernels meet only in the natural-order work array."""
    test_compute_fields_matches_oracle(N, precision, 'rfft',
                                       extra={'rfft-wave': mask, 'fft-radix-plan': list(plan),
                                              'rfft-cap': cap, 'rfft-tpb': tpb})


def _random_plasma(rng, p, spread=0.3, pmax=0.05):
    a = O.uniform_plasma(p)
    n = a['x_init'].shape
    a['x_offt'] = rng.uniform(-spread, spread, n) * p.h
    a['y_offt'] = rng.uniform(-spread, spread, n) * p.h
    for k in ('px', 'py', 'pz'):
        a[k] = rng.normal(0, pmax, n)
    a['q'][rng.uniform(size=n) < 0.01] = 0.0
    return a


def _sort_arrays(a, dtype=torch.float64):
    from lcodepy_amd.plasma3d.data import Arrays
    pos = ('x_init', 'y_init', 'x_offt', 'y_offt', 'pz')
    return Arrays.wrap(**{k: dev(v, torch.float64 if k in pos else dtype) for k, v in a.items()})


def _round_bulk(a):
    """The FP32-bulk form's inputs: px py q m rounded to float32."""
    b = dict(a)
    for k in ('px', 'py', 'q', 'm'):
        b[k] = a[k].astype(np.float32).astype(np.float64)
    return b


@pytest.mark.parametrize('N,precision', [(65, 'fp64'), (129, 'fp64'), (65, 'fp32-bulk'),
                                         (129, 'fp32-bulk')])
def test_deposit_matches_oracle_and_is_deterministic(N, precision):
    from lcodepy_amd.plasma3d.data import Arrays
    from lcodepy_amd.plasma3d.weights import get_deposit_plasma
    print(f'deposit N={N} {precision}', flush=True)
    f32 = precision == 'fp32-bulk'
    rng = np.random.default_rng(10 + N)
    p = _params(N)
    a = _random_plasma(rng, p)
    if f32:
        a = _round_bulk(a)
    g = [np.zeros((N, N)) for _ in range(4)]
    O.deposit_sort(p, a, *g)
    deposit = get_deposit_plasma(_cfg(N, precision=precision))
    const = Arrays.wrap()
    const.sorts = {'electrons': 0}
    parts = {'electrons': _sort_arrays(a, torch.float32 if f32 else torch.float64)}
    out1 = deposit(parts, const)
    out2 = deposit(parts, const)
    torch.cuda.synchronize()
    # float32 rounding of the length-2048 transforms and the Neumann solve
    # of Bz reaches 1.06e-5 at N = 1025
    bound = (FP32_BOUND if N <= 513 else 3e-5) if f32 else 1e-12
    for u, v, w in zip(out1, out2, g):
        assert torch.equal(u, v)                       # bit-reproducible
        e = relerr(host(u[0]), w)                      # cell-order vs particle-order sums
        print(f'deposit N={N} {precision}: {e:.1e}')
        assert e < bound
        assert not torch.any(u[1])


@pytest.mark.parametrize('N,frac', [(65, 0.005), (65, 0.05), (65, 0.3), (129, 0.45)])
def test_deposit_of_collapsed_plasma(N, frac):
    """A fraction of the particles collapsed into the cell on the axis
    (about 28, 230 or 1400 at N=65: the workgroup sort of dense cell lists,
    short and long; over 16384 in one cell at N=129: arrival-order summation,
    reported as a warning bit).  Against the oracle within 1e-12; bitwise reproducible
    unless a cell overflows."""
    import warnings
    from lcodepy_amd.plasma3d.data import Arrays
    from lcodepy_amd.plasma3d.weights import get_deposit_plasma
    from lcodepy_amd import runtime
    print(f'collapsed deposit N={N} frac={frac}', flush=True)
    rng = np.random.default_rng(30 + N)
    p = _params(N)
    a = _random_plasma(rng, p, spread=0.3)
    pick = rng.uniform(size=a['x_init'].shape) < frac
    a['x_offt'][pick] = -a['x_init'][pick] + rng.uniform(-0.45, 0.45, pick.sum()) * p.h
    a['y_offt'][pick] = -a['y_init'][pick] + rng.uniform(-0.45, 0.45, pick.sum()) * p.h
    g = [np.zeros((N, N)) for _ in range(4)]
    O.deposit_sort(p, a, *g)
    cfg = _cfg(N)
    deposit = get_deposit_plasma(cfg)
    const = Arrays.wrap()
    const.sorts = {'electrons': 0}
    parts = {'electrons': _sort_arrays(a)}
    out1 = deposit(parts, const)
    out2 = deposit(parts, const)
    torch.cuda.synchronize()
    ctx = runtime.context(cfg)
    flags = ctx.error_flags()
    overflow = bool(flags & 8)
    assert overflow == (N == 129), flags
    assert not flags & ~8, flags
    for u, v, w in zip(out1, out2, g):
        if not overflow:
            assert torch.equal(u, v)
        assert relerr(host(u[0]), w) < 1e-12


@pytest.mark.parametrize('N,precision', [(65, 'fp64'), (129, 'fp64'), (65, 'fp32-bulk'),
                                         (129, 'fp32-bulk')])
def test_push_matches_oracle(N, precision):
    """Free stream and one smart push in random smooth fields.  FP32-bulk:
    float32 momenta and half-step fields, float64 positions and pz."""
    from lcodepy_amd.plasma3d.data import Arrays
    from lcodepy_amd.plasma3d.move import get_plasma_particles_mover
    print(f'push N={N} {precision}', flush=True)
    f32 = precision == 'fp32-bulk'
    rng = np.random.default_rng(20 + N)
    p = _params(N)
    prev = _random_plasma(rng, p)
    fields = {k: 0.2 * _smooth(rng, N) for k in ('Ex', 'Ey', 'Ez', 'Bx', 'By', 'Bz', 'Phi')}
    if f32:
        prev = _round_bulk(prev)
        fields = {k: v.astype(np.float32).astype(np.float64) for k, v in fields.items()}
    full_o = O.move_wo_fields(p, {'electrons': prev})
    full_o = O.move_smart(p, fields, {'electrons': prev}, full_o)['electrons']

    wo, smart = get_plasma_particles_mover(_cfg(N, precision=precision))
    const = Arrays.wrap()
    const.sorts = {'electrons': 0}
    bulk = torch.float32 if f32 else torch.float64
    prev_g = {'electrons': _sort_arrays(prev, bulk)}
    full_g = wo(prev_g, const)
    full_g = smart(Arrays.wrap(**{k: dev(v, bulk) for k, v in fields.items()}), prev_g, full_g,
                   const)['electrons']
    torch.cuda.synchronize()
    # float32 rounding of the length-2048 transforms and the Neumann solve
    # of Bz reaches 1.06e-5 at N = 1025
    bound = (FP32_BOUND if N <= 513 else 3e-5) if f32 else 1e-12
    for k in ('x_offt', 'y_offt', 'px', 'py', 'pz', 'q'):
        e = relerr(host(getattr(full_g, k)), full_o[k])
        assert e < bound, (k, e)


def _wall_plasma(rng, p):
    """A lattice whose outer particles are set up to cross the reflection
    boundary b = h (N/2 - 10).  With zero offsets and momenta
    (px, pz) = (a m, s m) the position update moves a particle by
    D xi_step, D = a / (sqrt(1 + a^2 + s^2) - s); for s = 2 D that gives
    a = (4 - sqrt(12)) / (2 (1 - 1 / D^2)) - so the outer particles land
    beyond b (reflected, momenta zeroed) or beyond 3 b (still outside after
    the reflection: absorbed, q = 0).  Some particles get a full-step
    estimate outside b, so the midpoint used for the field gather is
    reflected as well."""
    a = O.uniform_plasma(p)
    n = a['x_init'].shape
    m = a['m']
    for k in ('x_offt', 'y_offt', 'px', 'py', 'pz'):
        a[k] = np.zeros(n)
    b = p.boundary
    edge = np.abs(a['x_init']) > b - 0.6
    sign = np.sign(a['x_init'])
    kind = rng.integers(0, 3, n)
    for kk, frac in {0: 0.3, 1: 1.2, 2: 3.2}.items():   # x displacement / b
        sel = edge & (kind == kk)
        D = frac * b / p.xi_step
        s = 2 * D
        aa = (2 * s / D - math.sqrt(4 * s * s / (D * D) - 4 * (1 - 1 / (D * D)))) / (2 * (1 - 1 / (D * D)))
        a['px'][sel] = sign[sel] * aa * m[sel]
        a['pz'][sel] = s * m[sel]
    est = {k: v.copy() for k, v in a.items()}
    far = edge & (rng.uniform(size=n) < 0.3)
    est['x_offt'][far] = sign[far] * 2 * (b - np.abs(a['x_init'][far]) + 0.1)
    return a, est


def _edge_quiet_fields(rng, N, h):
    """Smooth random fields damped towards the window edge (the wall-bound
    particles are ultra-relativistic, so full-strength fields would swamp
    their prescribed motion)."""
    g = (np.arange(N) - N // 2) * h
    mask = np.exp(-2 * (g[:, None] ** 2 + g[None, :] ** 2))
    return {k: 1e-3 * mask * _smooth(rng, N) for k in ('Ex', 'Ey', 'Ez', 'Bx', 'By', 'Bz', 'Phi')}


def test_push_reflects_and_absorbs_at_the_wall():
    """Particles driven across the wall: reflected ones (momentum zeroed)
    and absorbed ones (q = 0, offsets and momenta zeroed), also for the
    midpoint reflection of the field gather, against the oracle."""
    from lcodepy_amd.plasma3d.data import Arrays
    from lcodepy_amd.plasma3d.move import get_plasma_particles_mover
    N, xi_step = 65, 0.5
    print(f'wall push N={N}', flush=True)
    rng = np.random.default_rng(77)
    p = _params(N, xi_step=xi_step)
    prev, est = _wall_plasma(rng, p)
    fields = _edge_quiet_fields(rng, N, p.h)
    mid = prev['x_init'] + (prev['x_offt'] + est['x_offt']) / 2
    assert np.sum(np.abs(mid) > p.boundary) > 10 and np.max(np.abs(mid)) < 2 * p.boundary
    full_est = {k: v.copy() for k, v in est.items()}
    full_o = O.move_smart(p, fields, {'electrons': prev}, {'electrons': full_est})['electrons']
    absorbed = (full_o['q'] == 0) & (prev['q'] != 0)
    x_new = prev['x_init'] + full_o['x_offt']
    reflected = (~absorbed) & (full_o['px'] == 0) & (prev['px'] != 0)
    print(f'wall push: {int(absorbed.sum())} absorbed, {int(reflected.sum())} reflected')
    assert absorbed.sum() > 10 and reflected.sum() > 10
    assert np.all(np.abs(x_new[~absorbed]) < p.boundary)

    _, smart = get_plasma_particles_mover(_cfg(N, xi_step=xi_step))
    const = Arrays.wrap()
    const.sorts = {'electrons': 0}
    prev_g = {'electrons': _sort_arrays(prev)}
    full_g = {'electrons': _sort_arrays(est)}
    out = smart(Arrays.wrap(**{k: dev(v) for k, v in fields.items()}), prev_g, full_g,
                const)['electrons']
    torch.cuda.synchronize()
    from lcodepy_amd import runtime
    assert runtime.context(_cfg(N, xi_step=xi_step)).error_flags() == 0
    for k in ('x_offt', 'y_offt', 'px', 'py', 'pz', 'q'):
        e = relerr(host(getattr(out, k)), full_o[k])
        # ultra-relativistic particles and a 0.5 xi step amplify the
        # rounding of the fused multiply-adds the GPU contracts (measured
        # 1.2e-12 for x_offt)
        assert e < 1e-11, (k, e)
    assert np.array_equal(host(out.q) == 0, full_o['q'] == 0)


def _gauss_beam(N, h, xi, amp=-0.05):
    grid = (np.arange(N) - N // 2) * h
    r2 = grid[:, None] ** 2 + grid[None, :] ** 2
    return amp * np.exp(-r2 / 2) * np.exp(-(xi + 1.0) ** 2 / 2)


@pytest.mark.parametrize('N,ion_model', [(33, 'background'), (33, 'mobile'), (65, 'background'),
                                         (129, 'background'), (129, 'mobile')])
def test_plasma_step_matches_oracle(N, ion_model):
    from lcodepy_amd.plasma3d.initialization import init_plasma
    from lcodepy_amd.plasma3d.solver import Plane2d3vPlasmaSolver
    print(f'plasma step N={N} {ion_model}', flush=True)
    cfg = _cfg(N, ion_model=ion_model)
    cfg.set('ion-mass', 100.0)
    p = _params(N, ion_model=ion_model)
    p.ion_mass = 100.0
    fields_o, parts_o, cur_o, const_o, _ = O.init_plasma(p)
    fields_g, parts_g, cur_g, const_g, _ = init_plasma(cfg)
    assert relerr(host(const_g.ro_initial), const_o.ro_initial) < 1e-14
    solver = Plane2d3vPlasmaSolver(cfg)
    rb_prev = np.zeros((N, N))
    worst = 0.0
    for step in range(30):
        rb_full = _gauss_beam(N, p.h, -(step + 1) * p.xi_step, amp=-0.5)
        parts_o, fields_o, cur_o = O.step_dxi(p, parts_o, fields_o, cur_o, const_o, rb_full, rb_prev)
        parts_g, fields_g, cur_g = solver.step_dxi(parts_g, fields_g, cur_g, const_g,
                                                   dev(rb_full), dev(rb_prev))
        rb_prev = rb_full
    torch.cuda.synchronize()
    # measured at most 3.3e-14 (N = 129, mobile ions)
    for k in ('Ex', 'Ey', 'Ez', 'Bx', 'By', 'Bz', 'Phi'):
        e = relerr(host(getattr(fields_g, k)), fields_o[k])
        worst = max(worst, e)
        assert e < 1e-12, (k, e)
    for sort in parts_o:
        for k in ('x_offt', 'y_offt', 'px', 'py', 'pz'):
            e = relerr(host(getattr(parts_g[sort], k)), parts_o[sort][k])
            assert e < 1e-12, (sort, k, e)
    for k in ('ro', 'jx', 'jy', 'jz'):
        assert relerr(host(getattr(cur_g, k)), cur_o[k]) < 1e-12, k
    print(f'worst field error after 30 steps: {worst:.2e}')


@pytest.mark.parametrize('precision', ['fp64', 'fp32-bulk'])
def test_step_currents_do_not_depend_on_the_visit_order(precision):
    """The currents a plasma step leaves come from its last push, which bins
    the particles in the cell order of the deposit before (particles that
    changed cells land out of order in their new cell's list); a standalone
    deposit of the same particles bins them in particle order.  Both sum each
    cell's records in particle-id order, so they agree bit for bit.  N = 33,
    10 steps with a beam and random transverse momenta up to 0.4, so that
    particles cross cells; the particle arrays are first shuffled (a seeded
    permutation of the ids), so that a cell's ids do not follow the lattice
    and a particle arriving from a neighbour lands out of id order whatever
    the waves' timing."""
    from lcodepy_amd.plasma3d.data import Arrays
    from lcodepy_amd.plasma3d.initialization import init_plasma
    from lcodepy_amd.plasma3d.solver import Plane2d3vPlasmaSolver
    from lcodepy_amd.plasma3d.weights import get_deposit_plasma
    N = 33
    print(f'visit order N={N} {precision}', flush=True)
    cfg = _cfg(N, precision=precision)
    h = cfg.getfloat('transverse-step')
    fields, parts, cur, const, _ = init_plasma(cfg)
    el = parts['electrons']
    n = el.x_init.shape[0]
    perm = torch.from_numpy(np.random.default_rng(7).permutation(n)).to(el.x_init.device)
    for k in el.names():
        v = getattr(el, k)
        if isinstance(v, torch.Tensor) and v.dim() == 1 and v.shape[0] == n:
            setattr(el, k, v[perm].contiguous())
    rng = np.random.default_rng(8)
    for k in ('px', 'py'):
        v = getattr(el, k)
        v.copy_(torch.from_numpy(rng.uniform(-0.4, 0.4, n)).to(v.dtype))
    solver = Plane2d3vPlasmaSolver(cfg)
    rb_prev = np.zeros((N, N))
    for step in range(10):
        rb_full = _gauss_beam(N, h, -(step + 1) * 0.05, amp=-1.0)
        parts, fields, cur = solver.step_dxi(parts, fields, cur, const, dev(rb_full), dev(rb_prev))
        rb_prev = rb_full
    moved = torch.abs(parts['electrons'].x_offt) > h
    assert float(moved.double().mean()) > 0.2   # crossed at least one cell
    single = Arrays.wrap()
    single.sorts = {'electrons': 0}
    out = get_deposit_plasma(cfg)(parts, single)
    torch.cuda.synchronize()
    solver.ctx.raise_on_errors()
    for k, u in zip(('ro', 'jx', 'jy', 'jz'), out):
        assert torch.any(u[0] != 0), k
        assert torch.equal(u[0], getattr(cur, k)[0]), k


@pytest.mark.parametrize('precision', ['fp64', 'fp32-bulk'])
def test_energy_flux_matches_oracle(precision):
    """calculate_energy_fluxes (Strategy3D.py:6-25) on the device against the
    oracle's restatement on the same state (10 plasma steps with a beam at
    N = 65, so fields and momenta are non-trivial): within 1e-12, the inputs
    being identical and only the summation order differing."""
    from lcodepy_amd.diagnostics import calculate_energy_fluxes
    from lcodepy_amd.plasma3d.initialization import init_plasma
    from lcodepy_amd.plasma3d.solver import Plane2d3vPlasmaSolver
    N = 65
    print(f'energy flux N={N} {precision}', flush=True)
    cfg = _cfg(N, precision=precision)
    h = cfg.getfloat('transverse-step')
    fields, parts, cur, const, _ = init_plasma(cfg)
    solver = Plane2d3vPlasmaSolver(cfg)
    rb_prev = np.zeros((N, N))
    for step in range(10):
        rb_full = _gauss_beam(N, h, -(step + 1) * 0.05, amp=-0.5)
        parts, fields, cur = solver.step_dxi(parts, fields, cur, const, dev(rb_full), dev(rb_prev))
        rb_prev = rb_full
    psi = calculate_energy_fluxes(h, parts, fields)
    torch.cuda.synchronize()
    solver.ctx.raise_on_errors()
    el = parts['electrons']
    ref = O.energy_flux(h, {k: host(getattr(el, k)) for k in ('m', 'px', 'py', 'pz')},
                        {k: host(getattr(fields, k)) for k in ('Ex', 'Ey', 'Ez', 'Bx', 'By', 'Bz')})
    got = float(psi.item())
    print(f'energy flux {precision}: {got:.15e} vs oracle {ref:.15e}')
    assert psi.dtype == torch.float64 and psi.is_cuda
    assert ref != 0.0
    assert abs(got - ref) <= 1e-12 * abs(ref)
