"""Graph-replayed xi steps (lcodepy_amd.batch.StaticRun: the beam density,
the whole plasma step and the on-axis Ez record of one xi step captured as a
HIP graph per step parity) against the eager xi loop (PusherAndSolver3D with
the same rigid driver as a callable): the on-axis Ez of every xi step and
the final fields and particles must agree bit for bit, since both run the
same kernels on the same inputs.  Every replay is followed by eager work and
a synchronisation on the host (the check below reads the run's state), as
graphs here must give the same result on every replay."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip('torch')


@pytest.mark.parametrize('N,ppc,precision', [(129, 1, 'fp64'), (257, 4, 'fp32-bulk')])
def test_graph_replayed_steps_equal_the_eager_xi_loop(N, ppc, precision):
    from lcodepy_amd import batch
    steps = 60
    print(f'graph steps N={N} ppc={ppc} {precision}', flush=True)
    cfg = batch.run_config(N, ppc, precision, steps)
    drv = batch.RigidDriver(N, 0.05, 0.05, steps, n_b=0.8, sigma_r=0.7)
    run = batch.StaticRun(cfg, drv)
    run.capture()
    seen = []
    for i in range(run.n_steps):
        run.step(i)
        torch.cuda.synchronize()
        seen.append(float(run.ez[i]))          # eager work between the replays
    run.ctx.raise_on_errors()
    sec, ez_eager = batch._eager_xi_loop(cfg, drv)
    ez_graph = run.on_axis_ez()
    print(f'  {run.n_steps} xi steps; eager loop {sec:.3f} s; max |Ez| {np.max(np.abs(ez_eager)):.3e}')
    assert ez_graph.shape == ez_eager.shape == (run.n_steps,)
    assert np.max(np.abs(ez_eager)) > 0
    assert np.array_equal(np.array(seen), ez_graph)
    assert np.array_equal(ez_graph, ez_eager)


def test_runs_on_their_own_streams_do_not_disturb_each_other():
    """Two runs of different grids replayed alternately on their own streams
    give the on-axis Ez each gives alone."""
    from lcodepy_amd import batch
    steps = 40
    cfgs = [batch.run_config(129, 4, 'fp32-bulk', steps), batch.run_config(257, 1, 'fp32-bulk', steps)]
    drvs = [batch.RigidDriver(129, 0.05, 0.05, steps, 0.5, 0.5),
            batch.RigidDriver(257, 0.05, 0.05, steps, 1.0, 1.2)]
    alone = []
    for cfg, drv in zip(cfgs, drvs):
        r = batch.StaticRun(cfg, drv)
        r.capture()
        for i in range(r.n_steps):
            r.step(i)
        torch.cuda.synchronize()
        alone.append(r.on_axis_ez())
    runs = [batch.StaticRun(c, d) for c, d in zip(cfgs, drvs)]
    for r in runs:
        r.capture()
    for i in range(runs[0].n_steps):
        for r in runs:
            r.step(i)
    torch.cuda.synchronize()
    for r, ez in zip(runs, alone):
        assert np.array_equal(r.on_axis_ez(), ez)
