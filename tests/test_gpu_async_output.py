"""GPU tests of the streamed per-slice output (qp_wake_params.async_output, mixed path).

The planes of every slice are snapshotted on the device, copied on a second
stream into pinned host slots and handed to the callback by a writer thread
while later slices compute (the slice step stays a hipGraph).  The callback must
see what the inline callback sees: the same slices in the same order with the
same fields, the same stop slice and STOPPED view, for single wakes and
batches; a failing callback ends the run with an error and leaves the handle
usable.
"""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

L, LX, DX = 20.0, 4.0, 0.1
FIELDS = ("psi", "rho", "jx", "jy", "jz", "ex", "ey", "ez", "by", "bz", "chi")
LASER = (1.2, 4.0, 6.0, 6.0, LX - 2.0, L / 2 + 0.5, L / 2)


def make(qp_ctx, oracle, n, async_output, b_threshold=100.0):
    from quetzalpic_amd import capi
    g = oracle.grid_from_box(LX, L, L, DX, L / n, L / n, K=5, b_threshold=b_threshold)
    y, z, w = oracle.init_lattice(g, 2, 2)
    wk = capi.Wake(qp_ctx, g.nx, g.ny, g.nz, g.dx, g.dy, g.dz, g.ly, g.lz, K=5, b_threshold=b_threshold,
                   precision=capi.QP_PREC_MIXED, async_output=async_output)
    wk.set_particles(y, z, w)
    return g, wk


def run_logged(wk, laser):
    wk.set_laser_gaussian(*laser)
    order = []
    out, st, stop = wk.run(fields=FIELDS, callback=lambda v: order.append((v.slice, v.kind)) and 0)
    return out, st, stop, order


@pytest.mark.parametrize("n", [32, 24])
def test_async_equals_inline(qp_ctx, oracle, n):
    _, a = make(qp_ctx, oracle, n, 0)
    _, b = make(qp_ctx, oracle, n, 1)
    want, st0, _, order0 = run_logged(a, LASER)
    got, st1, _, order1 = run_logged(b, LASER)
    assert st0 == st1 == 0
    assert order1 == order0 and len(order0) > 10
    for f in FIELDS:
        assert rel_err(got[f], want[f])[0] < 1e-9, f
    # again on the same handle (captured slice steps and output slots reused)
    got2, _, _, order2 = run_logged(b, LASER)
    assert order2 == order0
    for f in FIELDS:
        assert rel_err(got2[f], want[f])[0] < 1e-9, f
    a.close()
    b.close()


def test_async_stop(qp_ctx, oracle):
    """THRESHOLD_B = 0.3 with a0 = 8: the same stop slice and STOPPED view."""
    strong = (8.0,) + LASER[1:]
    _, a = make(qp_ctx, oracle, 32, 0, b_threshold=0.3)
    _, b = make(qp_ctx, oracle, 32, 1, b_threshold=0.3)
    want, st0, ss0, order0 = run_logged(a, strong)
    got, st1, ss1, order1 = run_logged(b, strong)
    assert st0 == st1 == 1 and ss0 == ss1 >= 0
    assert order1 == order0
    assert (ss0, 1) in order0   # QP_SLICE_STOPPED
    for f in ("by", "bz", "chi", "psi", "ex"):
        assert rel_err(got[f], want[f])[0] < 1e-9, f
    a.close()
    b.close()


def test_async_batch(qp_ctx, oracle):
    from quetzalpic_amd import capi
    g = oracle.grid_from_box(LX, L, L, DX, L / 32, L / 32, K=5)
    y, z, w = oracle.init_lattice(g, 2, 2)
    outs = []
    for mode in (0, 1):
        bt = capi.WakeBatch(qp_ctx, 3, g.nx, g.ny, g.nz, g.dx, g.dy, g.dz, g.ly, g.lz, K=5, async_output=mode)
        for m in range(3):
            bt.set_particles(m, y, z, w * (1 + 0.2 * m))
            bt.set_laser_gaussian(m, 1.0 + 0.3 * m, *LASER[1:])
        res, st, stops = bt.run(fields=FIELDS)
        assert st == 0 and stops == [-1, -1, -1]
        outs.append(res)
        bt.close()
    for m in range(3):
        for f in FIELDS:
            assert rel_err(outs[1][m][f], outs[0][m][f])[0] < 1e-9, (m, f)


def test_async_callback_error(qp_ctx, oracle):
    from quetzalpic_amd import capi
    g, b = make(qp_ctx, oracle, 32, 1)
    b.set_laser_gaussian(*LASER)
    with pytest.raises(capi.QpError):
        b.run(fields=("psi",), callback=lambda v: 1 if v.slice == g.nx - 5 else 0)
    _, a = make(qp_ctx, oracle, 32, 0)
    want, _, _, _ = run_logged(a, LASER)
    got, st, _, _ = run_logged(b, LASER)     # the handle still works
    assert st == 0
    assert rel_err(got["psi"], want["psi"])[0] < 1e-9
    a.close()
    b.close()


def test_async_needs_the_mixed_path(qp_ctx, oracle):
    from quetzalpic_amd import capi
    g = oracle.grid_from_box(LX, L, L, DX, L / 32, L / 32, K=5)
    with pytest.raises(capi.QpError):
        capi.Wake(qp_ctx, g.nx, g.ny, g.nz, g.dx, g.dy, g.dz, g.ly, g.lz, K=5, precision=capi.QP_PREC_FP64,
                  async_output=1)
