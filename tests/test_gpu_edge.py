"""GPU edge cases of the slice loop: a wake with no plasma particles (a profile
that is zero everywhere leaves init_wake_particles, src/System_3d.cpp:724-748,
with none), driven by bunch sources alone, on both device paths against the
oracle; and a batch in which one member has no particles beside members that
have them, against single runs."""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

L, LX, DX = 20.0, 3.0, 0.1
FIELDS = ("psi", "rho", "jx", "jy", "jz", "ex", "ey", "ez", "by", "bz", "chi")


def bunch_sources(g, amp=0.05):
    x = (np.arange(g.nx) * g.dx)[:, None, None]
    y = (np.arange(g.ny) * g.dy)[None, :, None]
    z = (np.arange(g.nz) * g.dz)[None, None, :]
    rb = -amp * np.exp(-((x - 1.5) ** 2 / 0.5 + (y - L / 2) ** 2 / 4 + (z - L / 2) ** 2 / 4))
    return rb, -rb


@pytest.mark.parametrize("prec,n,tol", [(0, 32, 1e-10), (1, 32, 1e-5), (1, 24, 1e-5)])
def test_wake_without_plasma_particles(qp_ctx, oracle, prec, n, tol):
    from quetzalpic_amd import capi
    g = oracle.grid_from_box(LX, L, L, DX, L / n, L / n, K=3)
    rb, jb = bunch_sources(g)
    e = np.zeros(0)
    want, st0, _ = oracle.wake_run(g, e, e, e, np.zeros((g.nx, g.ny, g.nz)), rho_b=rb, jx_b=jb)
    wk = capi.Wake(qp_ctx, g.nx, g.ny, g.nz, g.dx, g.dy, g.dz, g.ly, g.lz, K=3, precision=prec)
    wk.set_particles(e, e, e)
    wk.set_a_sqr(np.zeros((g.nx, g.ny, g.nz)))
    wk.set_bunch_sources(rb, jb)
    got, st, _ = wk.run(fields=FIELDS)
    wk.close()
    assert st == st0 == 0
    assert np.abs(want["psi"]).max() > 1e-3 and np.abs(want["by"]).max() > 1e-4   # the bunch drives a field
    for f in ("psi", "ex", "ey", "ez", "by", "bz", "rho", "jx"):
        assert rel_err(got[f], want[f])[0] < tol, (f, rel_err(got[f], want[f]))
    for f in ("jy", "jz", "chi"):
        assert not np.any(got[f]), f


def test_batch_member_without_particles(qp_ctx, oracle):
    from quetzalpic_amd import capi
    g = oracle.grid_from_box(LX, L, L, DX, L / 32, L / 32, K=3)
    y, z, w = oracle.init_lattice(g, 2, 2)
    laser = (1.5, 4.0, 6.0, 6.0, 1.5, L / 2, L / 2)
    e = np.zeros(0)
    members = [(y, z, w), (e, e, e), (y, z, 0.5 * w)]
    bt = capi.WakeBatch(qp_ctx, 3, g.nx, g.ny, g.nz, g.dx, g.dy, g.dz, g.ly, g.lz, K=3)
    for m, (yy, zz, ww) in enumerate(members):
        bt.set_particles(m, yy, zz, ww)
        bt.set_laser_gaussian(m, *laser)
    outs, st, stops = bt.run(fields=FIELDS)
    bt.close()
    assert st == 0 and stops == [-1, -1, -1]
    for f in ("psi", "ex", "by", "rho", "chi"):
        assert not np.any(outs[1][f]), f        # no plasma, no bunch: every field stays zero
    for m in (0, 2):
        yy, zz, ww = members[m]
        wk = capi.Wake(qp_ctx, g.nx, g.ny, g.nz, g.dx, g.dy, g.dz, g.ly, g.lz, K=3, precision=capi.QP_PREC_MIXED)
        wk.set_particles(yy, zz, ww)
        wk.set_laser_gaussian(*laser)
        single, _, _ = wk.run(fields=FIELDS)
        wk.close()
        for f in ("psi", "ex", "ey", "by", "bz", "rho", "jx", "jy", "chi"):
            assert rel_err(outs[m][f], single[f])[0] < 1e-9, (m, f)
