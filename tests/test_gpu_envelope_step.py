"""GPU tests of the envelope step a wake pass takes itself (qp_wake_set_envelope_step).

A pass with a step set must leave the envelope where a pass followed by
qp_envelope_advance(dtau, 1, chi_iterations) leaves it: the step is enqueued on
the wake's stream behind the pass, with the chi that pass stored.  Covered: graph
replay and eager launches, register-resident and generic FFT grids, both
precisions, two passes in a row, a pass the stop rule ends early, and streamed
output at the same time.
"""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

L, LX, DX = 20.0, 4.0, 0.1
LASER = (1.4, 4.0, 6.0, 6.0, LX - 2.0, L / 2 + 0.5, L / 2)
K0, DT = 10.0, 2.0


def setup(qp_ctx, oracle, n, prec, use_graph=1, b_threshold=100.0, async_output=0, laser=LASER):
    from quetzalpic_amd import capi
    g = oracle.grid_from_box(LX, L, L, DX, L / n, L / n, K=5, b_threshold=b_threshold)
    y, z, w = oracle.init_lattice(g, 2, 2)
    wk = capi.Wake(qp_ctx, g.nx, g.ny, g.nz, g.dx, g.dy, g.dz, g.ly, g.lz, K=5, b_threshold=b_threshold,
                   precision=prec, use_graph=use_graph, async_output=async_output)
    wk.set_particles(y, z, w)
    env = capi.Envelope(qp_ctx, g.nx, g.ny, g.nz, g.dx, g.dy, g.dz, K0)
    env.set_gaussian(*laser)
    wk.attach_envelope(env)
    return g, wk, env


def in_pass(qp_ctx, oracle, n, prec, passes=1, m=1, **kw):
    g, wk, env = setup(qp_ctx, oracle, n, prec, **kw)
    wk.set_envelope_step(DT, m)
    sts = []
    for _ in range(passes):
        _, st, ss = wk.run(fields=("psi",)) if kw.get("async_output") else wk.run()
        sts.append((st, ss))
    a, chi = env.get(), env.get_chi()
    wk.close()
    env.close()
    return a, chi, sts


def sequential(qp_ctx, oracle, n, prec, passes=1, m=1, **kw):
    g, wk, env = setup(qp_ctx, oracle, n, prec, **kw)
    sts = []
    for _ in range(passes):
        _, st, ss = wk.run()
        sts.append((st, ss))
        env.advance(DT, steps=1, chi_iterations=m)
    a, chi = env.get(), env.get_chi()
    wk.close()
    env.close()
    return a, chi, sts


@pytest.mark.parametrize("prec,n,use_graph", [(1, 32, 1), (1, 32, 0), (1, 24, 1), (0, 32, 1)])
def test_step_in_the_pass_equals_run_then_advance(qp_ctx, oracle, prec, n, use_graph):
    a1, c1, _ = in_pass(qp_ctx, oracle, n, prec, use_graph=use_graph)
    a0, c0, _ = sequential(qp_ctx, oracle, n, prec, use_graph=use_graph)
    assert rel_err(c1, c0)[0] < 1e-9
    assert rel_err(a1, a0)[0] < 1e-11
    g = oracle.grid_from_box(LX, L, L, DX, L / n, L / n)
    from oracle import envelope as envo
    start = envo.gaussian(g.nx, g.ny, g.nz, g.dx, g.dy, g.dz, *LASER)
    assert np.abs(a1 - start).max() > 1e-3 * np.abs(start).max()   # the step moved it


def test_two_passes_and_two_chi_corrections(qp_ctx, oracle):
    a1, _, _ = in_pass(qp_ctx, oracle, 32, 1, passes=2, m=2)
    a0, _, _ = sequential(qp_ctx, oracle, 32, 1, passes=2, m=2)
    assert rel_err(a1, a0)[0] < 1e-10


def test_stopped_pass(qp_ctx, oracle):
    """THRESHOLD_B = 0.3 with a0 = 8 ends the pass early; the step still covers every slice."""
    strong = dict(b_threshold=0.3, laser=(8.0,) + LASER[1:])
    a1, c1, s1 = in_pass(qp_ctx, oracle, 32, 1, **strong)
    a0, c0, s0 = sequential(qp_ctx, oracle, 32, 1, **strong)
    assert s1 == s0 and s0[0][0] == 1
    assert rel_err(c1, c0)[0] < 1e-9
    assert rel_err(a1, a0)[0] < 1e-11


def test_with_streamed_output(qp_ctx, oracle):
    a1, _, _ = in_pass(qp_ctx, oracle, 32, 1, async_output=1)
    a0, _, _ = sequential(qp_ctx, oracle, 32, 1)
    assert rel_err(a1, a0)[0] < 1e-11
