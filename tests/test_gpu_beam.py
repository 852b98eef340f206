"""The beam layer move on the GPU against the oracle, on branches the
workload beams (gamma ~ 1000, well inside the window) never take:

* substepping: slow particles (gamma of a few) with a high substepping
  energy get their time step halved several times
  (beam3d/beam_calculator.py:18-34, 97-114) and run 4 or 8 substeps of the
  explicit-midpoint loop (beam3d/move.py:31-127);
* loss: particles that reach the lost radius max(0.9 R, R - 1) at the half
  or the full substep are flagged lost and their id negated
  (beam3d/move.py:56-65, 113-118);
* falling: particles that slip behind the layer at the half step, and the
  CPU scheme's trailing test on x (beam3d/move.py:123), which flags every
  particle with x < xi_{k+1}.

The lost, moved and fell groups must hold the same ids, and every particle
the same state, as the oracle's (float64, 1e-12 relative)."""
import numpy as np
import pytest

from oracle import lcode_oracle as O

pytestmark = pytest.mark.gpu
torch = pytest.importorskip('torch')

N, H, XI_STEP, TIME_STEP, E_SUB = 65, 0.1, 0.05, 0.01, 20.0
LAYER = 4          # plasma layer index of the move: beam layer LAYER - 1


def _config():
    from lcodepy_amd.config import Config
    cfg = Config({'geometry': '3d', 'processing-unit-type': 'mi355x', 'transverse-step': H,
                  'xi-step': XI_STEP, 'window-length': 1.0, 'time-step': TIME_STEP,
                  'beam-substepping-energy': E_SUB, 'ion-model': 'background',
                  'field-solver-subtraction-coefficient': 1.0})
    cfg.set('window-width-steps', N)
    cfg.set('window-width', N * H)
    cfg.set('plasma-fineness', 2)
    cfg.set('plasma-particles-per-cell', 4)
    return cfg


def _params():
    return O.Params(N=N, h=H, xi_step=XI_STEP, subtraction_coeff=1.0, time_step=TIME_STEP,
                    substepping_energy=E_SUB)


def _smooth(rng, amp):
    x = np.linspace(0, 1, N)
    f = np.zeros((N, N))
    for _ in range(3):
        a, b, ph = rng.uniform(1, 4), rng.uniform(1, 4), rng.uniform(0, 6)
        f += rng.normal() * np.sin(a * np.pi * x + ph)[:, None] * np.cos(b * np.pi * x - ph)[None, :]
    return amp * f


def _layer(rng):
    p = _params()
    xi_k, xi_k1 = -(LAYER - 1) * XI_STEP, -LAYER * XI_STEP
    n_fast, n_slow, n_edge, n_back = 300, 120, 60, 40
    n = n_fast + n_slow + n_edge + n_back
    b = {k: np.zeros(n) for k in ('xi', 'x', 'y', 'px', 'py', 'pz', 'q_m', 'q_norm', 'dt')}
    b['id'] = np.arange(1, n + 1, dtype=np.int64)
    b['remaining_steps'] = np.zeros(n, dtype=np.int64)
    b['q_m'][:] = -1.0
    b['q_norm'][:] = -1e-4
    b['xi'] = rng.uniform(xi_k1 + 0.01, xi_k - 0.001, n)
    r = rng.uniform(0, 1.5, n)
    phi = rng.uniform(0, 2 * np.pi, n)
    b['x'], b['y'] = r * np.cos(phi), r * np.sin(phi)
    b['pz'] = rng.normal(1000.0, 10.0, n)
    b['px'] = rng.normal(0, 1.0, n)
    b['py'] = rng.normal(0, 1.0, n)
    s = slice(n_fast, n_fast + n_slow)                 # slow: substeps
    b['pz'][s] = rng.uniform(0.2, 3.0, n_slow)
    b['px'][s] = rng.normal(0, 0.1, n_slow)
    b['py'][s] = rng.normal(0, 0.1, n_slow)
    e = slice(n_fast + n_slow, n_fast + n_slow + n_edge)   # at the lost radius, outward
    L = p.lost_radius
    re = L - rng.uniform(0.0005, 0.004, n_edge)
    b['x'][e], b['y'][e] = re * np.cos(phi[e]), re * np.sin(phi[e])
    b['px'][e], b['py'][e] = 600.0 * np.cos(phi[e]), 600.0 * np.sin(phi[e])
    k = slice(n_fast + n_slow + n_edge, n)              # at the back of the layer, slow
    b['xi'][k] = xi_k1 + rng.uniform(1e-5, 5e-4, n_back)
    b['pz'][k] = rng.uniform(0.5, 2.0, n_back)
    return b


def test_beam_move_substeps_losses_and_falls_match_oracle():
    from lcodepy_amd.beam3d.beam_calculator import BeamCalculator
    from lcodepy_amd.beam3d.data import FLOAT_KEYS, INT_KEYS, BeamParticles
    from lcodepy_amd.plasma3d.data import Arrays
    rng = np.random.default_rng(3)
    b = _layer(rng)
    p = _params()
    fk = {k: _smooth(rng, 0.05) for k in ('Ex', 'Ey', 'Ez', 'Bx', 'By', 'Bz')}
    fk1 = {k: _smooth(rng, 0.05) for k in ('Ex', 'Ey', 'Ez', 'Bx', 'By', 'Bz')}
    # Fields are gathered only at half-step positions inside the lost radius,
    # whose 5 x 5 stencils stay on the grid (|x| < (N // 2 - 2) h).
    assert np.max(np.hypot(b['x'], b['y'])) < p.lost_radius < (N // 2 - 2) * H

    layer_o = {k: v.copy() for k, v in b.items()}
    lost_o, moved_o, fell_o = O.BeamCalc(p).move_beam_layer(layer_o, 0, LAYER, fk, fk1)
    mult = layer_o['dt'] / TIME_STEP
    print(f'beam move: {lost_o["id"].size} lost, {moved_o["id"].size} moved, '
          f'{fell_o["id"].size} fell; substep multipliers {sorted(set(mult.round(6)))}')
    assert lost_o['id'].size >= 10 and np.all(lost_o['id'] < 0)
    assert np.sum(mult <= 0.25) >= 20                 # 4 or more substeps
    assert np.sum((mult <= 0.25) & (layer_o['remaining_steps'] == 0)
                  & np.isin(layer_o['id'], moved_o['id'])) >= 10

    print('beam move: gpu', flush=True)
    cfg = _config()
    f = torch.tensor(np.stack([b[k] for k in FLOAT_KEYS]), dtype=torch.float64, device='cuda')
    i = torch.tensor(np.stack([b[k] for k in INT_KEYS]), dtype=torch.int64, device='cuda')
    layer_g = BeamParticles(f=f, i=i)
    wrap = lambda d: Arrays.wrap(**{k: torch.tensor(v, device='cuda') for k, v in d.items()})
    calc = BeamCalculator(cfg)
    lost_g, moved_g, fell_g = calc.move_beam_layer(layer_g, 0, LAYER, wrap(fk), wrap(fk1))
    torch.cuda.synchronize()
    calc.ctx.raise_on_errors()
    for name, g, o in (('lost', lost_g, lost_o), ('moved', moved_g, moved_o),
                       ('fell', fell_g, fell_o)):
        ids = g.id.cpu().numpy()
        assert np.array_equal(ids, o['id']), name          # same particles, same order
        for k in ('x', 'y', 'xi', 'px', 'py', 'pz', 'dt'):
            got = getattr(g, k).cpu().numpy()
            scale = max(np.max(np.abs(o[k])), 1e-300) if o[k].size else 1.0
            assert np.max(np.abs(got - o[k]), initial=0.0) <= 1e-12 * scale, (name, k)
        assert np.array_equal(g.remaining_steps.cpu().numpy(), o['remaining_steps']), name


@pytest.mark.parametrize('n', [1, 700, 1024, 5000, 32 * 1024, 40000])
def test_beam_split_keeps_layer_order(n):
    """lc3d_beam_split (beam_calculator.py:124-139) against numpy's boolean
    gathers, bit for bit: one chunk, several chunks (one block per chunk,
    up to 32) and more (one block walking them all); three random
    disjoint-or-not flag rows, nine float64 and two int64 rows."""
    from lcodepy_amd import runtime
    from lcodepy_amd._native import ptr, stream_ptr
    print(f'beam split n={n}', flush=True)
    ctx = runtime.context(_config())
    rng = np.random.default_rng(n)
    f = rng.normal(size=(9, n))
    iv = rng.integers(-2**40, 2**40, size=(2, n), dtype=np.int64)
    flags = (rng.uniform(size=(3, n)) < np.array([[0.02], [0.9], [0.1]])).astype(np.uint8)
    dev = runtime.device()
    fd, ivd, fld = (torch.from_numpy(a).to(dev) for a in (f, iv, flags))
    out_f = torch.full((3, 9, n), np.nan, dtype=torch.float64, device=dev)
    out_i = torch.zeros((3, 2, n), dtype=torch.int64, device=dev)
    counts = torch.full((3,), -1, dtype=torch.int64, device=dev)
    ctx.check(ctx.lib.lc3d_beam_split(ctx.handle, n, 9, 2, ptr(fd), ptr(ivd), ptr(fld),
                                      ptr(out_f), ptr(out_i), ptr(counts), stream_ptr()),
              'lc3d_beam_split')
    torch.cuda.synchronize()
    got_f, got_i, got_c = out_f.cpu().numpy(), out_i.cpu().numpy(), counts.cpu().numpy()
    for g in range(3):
        sel = flags[g] != 0
        m = int(sel.sum())
        assert got_c[g] == m, (g, got_c[g], m)
        assert np.array_equal(got_f[g][:, :m], f[:, sel]), g
        assert np.array_equal(got_i[g][:, :m], iv[:, sel]), g


@pytest.mark.parametrize('n', [3000, 200_000])
def test_beam_deposit_is_reproducible(n):
    """lc3d_deposit_beam (beam3d/weights.py:40-64, 152-163) sums each cell's
    charge exactly in 128-bit fixed point: two calls on the same layer give
    the same bits (200,000 particles in a 65 x 65 window put thousands of
    additions on a cell, whose float64 atomic sum would vary with their
    order), the result is within 3e-14 of max |rho| of the oracle's serial
    deposit (measured 6.5e-16 and 8.1e-15: the oracle's own sequential
    rounding), a call adds onto the running buffers, and the layout call
    finishes the running layer (divided by h^2) and overwrites the next."""
    from lcodepy_amd import runtime
    from lcodepy_amd.beam3d.weights import get_deposit_beam
    print(f'beam deposit n={n}', flush=True)
    cfg = _config()
    deposit = get_deposit_beam(cfg)
    rng = np.random.default_rng(n)
    layer = 4
    xi_layer = -XI_STEP * layer
    # inside the window (|x|, |y| < 2.5 of 3.2): every stencil on the grid
    x = np.clip(rng.normal(0.0, 0.8, n), -2.5, 2.5)
    y = np.clip(rng.normal(0.0, 0.8, n), -2.5, 2.5)
    xi = xi_layer - rng.uniform(0.0, 1.0, n) * XI_STEP
    q = -rng.uniform(0.5, 1.5, n) * 1e-3
    dev = runtime.device()
    xd, yd, xid, qd = (torch.from_numpy(a).to(dev) for a in (x, y, xi, q))
    out = []
    for _ in range(2):
        ro0 = torch.zeros(N, N, dtype=torch.float64, device=dev)
        ro1 = torch.zeros(N, N, dtype=torch.float64, device=dev)
        deposit(layer, xd, yd, xid, qd, ro0, ro1)
        out.append((ro0, ro1))
    torch.cuda.synchronize()
    assert torch.equal(out[0][0], out[1][0]) and torch.equal(out[0][1], out[1][1])
    r0 = np.zeros((N, N))
    r1 = np.zeros((N, N))
    O.lib().oracle_deposit_beam(O.ctypes.c_int64(n), O.ctypes.c_int(N), O.ctypes.c_double(H),
                                O.ctypes.c_double(XI_STEP), O.ctypes.c_int64(layer), O._p(x),
                                O._p(y), O._p(xi), O._p(q), O._p(r0), O._p(r1),
                                O.ctypes.c_int(0))
    g0, g1 = out[0][0].cpu().numpy(), out[0][1].cpu().numpy()
    e0 = np.max(np.abs(g0 - r0)) / np.max(np.abs(r0))
    e1 = np.max(np.abs(g1 - r1)) / np.max(np.abs(r1))
    print(f'beam deposit n={n}: {e0:.1e} {e1:.1e}')
    assert e0 < 3e-14 and e1 < 3e-14
    # onto running buffers
    base0 = torch.from_numpy(rng.normal(size=(N, N))).to(dev)
    base1 = torch.from_numpy(rng.normal(size=(N, N))).to(dev)
    ro0, ro1 = base0.clone(), base1.clone()
    deposit(layer, xd, yd, xid, qd, ro0, ro1)
    torch.cuda.synchronize()
    assert torch.equal(ro0, base0 + out[0][0]) and torch.equal(ro1, base1 + out[0][1])
    # lc3d_layout_beam_layer (beam_calculator.py:82-93): the running buffer
    # becomes the finished layer / h^2, the next one is overwritten
    from lcodepy_amd._native import ptr, stream_ptr
    ctx = runtime.context(cfg)
    run = base0.clone()
    nxt = torch.full((N, N), 7.0, dtype=torch.float64, device=dev)
    ctx.check(ctx.lib.lc3d_layout_beam_layer(ctx.handle, n, layer, ptr(xd), ptr(yd), ptr(xid),
                                             ptr(qd), ptr(run), ptr(nxt), stream_ptr()),
              'lc3d_layout_beam_layer')
    torch.cuda.synchronize()
    # true division, as the reference's rho /= h**2 (torch would multiply by the
    # reciprocal of a host scalar)
    want = (base0 + out[0][0]).cpu().numpy() / H ** 2
    assert np.array_equal(run.cpu().numpy(), want) and torch.equal(nxt, out[0][1])
