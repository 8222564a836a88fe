"""Deposit-window statistics of the mixed path (qp_wake_window_stats): the
count of deposit blocks follows the launch structure of the pass, and a cold,
cell-sorted plasma never overflows the LDS window."""
import pytest

pytestmark = pytest.mark.gpu


def test_window_stats_cold_plasma(qp_ctx, oracle):
    from quetzalpic_amd import capi
    g = oracle.grid_from_box(2.0, 20.0, 20.0, 0.1, 20.0 / 32, 20.0 / 32, K=5)
    assert (g.nx, g.ny, g.nz) == (20, 32, 32)
    y, z, w = oracle.init_lattice(g, 2, 2)
    wk = capi.Wake(qp_ctx, g.nx, g.ny, g.nz, g.dx, g.dy, g.dz, g.ly, g.lz, K=5, precision=capi.QP_PREC_MIXED)
    with pytest.raises(capi.QpError, match="no pass"):
        wk.window_stats()
    wk.set_particles(y, z, w)
    _, st, _ = wk.run()
    assert st == 0
    ovf, tot = wk.window_stats()
    import json
    pt = json.loads(wk.describe())["tile"][0]   # threads (= particles) per block of the tuned tile
    blocks = (len(y) + pt - 1) // pt
    # slices nx-1 .. 1: A + 5 B + 5 C; slice 0: A + S
    assert tot == ((g.nx - 1) * 11 + 2) * blocks
    assert ovf == 0
    with pytest.raises(capi.QpError, match="out of range"):
        wk.window_stats(1)
    wk.close()
    f = capi.Wake(qp_ctx, g.nx, g.ny, g.nz, g.dx, g.dy, g.dz, g.ly, g.lz, K=5, precision=capi.QP_PREC_FP64)
    with pytest.raises(capi.QpError, match="mixed-precision"):
        f.window_stats()
    f.close()
