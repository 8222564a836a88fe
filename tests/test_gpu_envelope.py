"""GPU tests of the laser envelope (qp_envelope_*, an extension with no
counterpart in the reference) and of its coupling to the wake.

The device advance is held to the envelope oracle (oracle/envelope.py, pinned
on the CPU by dense solves and closed forms) on register-resident FFT grids
(2^p sides, square and not) and generic ones; the coupling is held to the
wake with the same Gaussian laser and to the slice loop oracle followed by the
envelope oracle.  Also: the vacuum Gaussian-beam closed form on the device,
hipGraph replays against eager sweeps, argument errors, and the rebuild of the
wake's captured slice graphs when its driver inputs change.
"""
import numpy as np
import pytest

from conftest import rel_err
from oracle import envelope as envo

pytestmark = pytest.mark.gpu

L, LX, DX = 20.0, 4.0, 0.1
LASER = (1.2, 4.0, 6.0, 6.0, LX - 2.0, L / 2 + 0.5, L / 2 - 0.5)


def random_state(shape, seed):
    rng = np.random.default_rng(seed)
    a = rng.normal(size=shape) + 1j * rng.normal(size=shape)
    chi = rng.uniform(0.0, 1.5, size=shape)
    return a, chi


@pytest.mark.parametrize("n1,n2", [(32, 32), (64, 32), (16, 128), (24, 24), (20, 36), (48, 96), (96, 64),
                                   (256, 256), (512, 512), (1024, 64)])
@pytest.mark.parametrize("m", [0, 2])
def test_advance_matches_oracle(qp_ctx, n1, n2, m):
    """2^p and 3 * 2^p sides (48, 96) take the register-resident FFT kernels, the
    others the generic ones.  n1 = 256, 512 and 1024 give the column kernel its
    64-, 128- and 256-thread shapes (one transform per block)."""
    from quetzalpic_amd import capi
    nx, dx, dy, dz, k0, dt = 12, 0.3, 0.7, 0.5, 5.0, 0.8
    a, chi = random_state((nx, n1, n2), 3 + n1 + n2 + m)
    e = capi.Envelope(qp_ctx, nx, n1, n2, dx, dy, dz, k0)
    e.set(a)
    e.set_chi(chi)
    assert np.array_equal(e.get_chi(), chi)
    e.advance(dt, steps=3, chi_iterations=m)
    got = e.get()
    asq = e.get_a_sqr()
    e.close()
    want = envo.advance(a, chi, dt, k0, dx, dy, dz, steps=3, m=m)
    assert rel_err(got, want)[0] < 1e-11
    assert np.abs(asq - 0.5 * np.abs(got) ** 2).max() <= 1e-14 * asq.max()


def test_graph_replays_equal_eager_sweeps(qp_ctx):
    from quetzalpic_amd import capi
    nx, n, dx, d, k0 = 10, 32, 0.25, 0.6, 8.0
    a, chi = random_state((nx, n, n), 7)
    out = []
    for use_graph in (0, 1):
        e = capi.Envelope(qp_ctx, nx, n, n, dx, d, d, k0)
        e.set(a)
        e.set_chi(chi)
        e.advance(0.5, steps=2, chi_iterations=1, use_graph=use_graph)
        e.advance(0.9, steps=1, chi_iterations=1, use_graph=use_graph)   # new dtau: graph captured again
        e.advance(0.5, steps=2, chi_iterations=2, use_graph=use_graph)   # new chi_iterations: again
        out.append(e.get())
        e.close()
    want = envo.advance(envo.advance(envo.advance(a, chi, 0.5, k0, dx, d, d, 2, 1), chi, 0.9, k0, dx, d, d, 1, 1),
                        chi, 0.5, k0, dx, d, d, 2, 2)
    assert rel_err(out[1], out[0])[0] < 1e-14
    assert rel_err(out[1], want)[0] < 1e-11


def test_gaussian_equals_the_wake_laser(qp_ctx, oracle):
    from quetzalpic_amd import capi
    g = oracle.grid_from_box(LX, L, L, DX, L / 32, L / 32, K=5)
    e = capi.Envelope(qp_ctx, g.nx, g.ny, g.nz, g.dx, g.dy, g.dz, 10.0)
    e.set_gaussian(*LASER)
    a = e.get()
    asq = e.get_a_sqr()
    e.close()
    want = envo.gaussian(g.nx, g.ny, g.nz, g.dx, g.dy, g.dz, *LASER)
    assert np.abs(a - want).max() <= 1e-14 * np.abs(want).max()
    ref = oracle.gaussian_a_sqr(g, *LASER)
    assert np.abs(asq - ref).max() <= 1e-14 * ref.max()


def wake(qp_ctx, oracle, prec, n=32, K=5):
    from quetzalpic_amd import capi
    g = oracle.grid_from_box(LX, L, L, DX, L / n, L / n, K=K)
    y, z, w = oracle.init_lattice(g, 2, 2)
    wk = capi.Wake(qp_ctx, g.nx, g.ny, g.nz, g.dx, g.dy, g.dz, g.ly, g.lz, K=K, precision=prec)
    wk.set_particles(y, z, w)
    return g, (y, z, w), wk


FIELDS = ("psi", "ex", "ey", "jx", "by", "chi")


@pytest.mark.parametrize("prec", [0, 1])
def test_wake_reads_the_envelope(qp_ctx, oracle, prec):
    """A wake with an attached Gaussian envelope equals the wake with the same Gaussian laser."""
    from quetzalpic_amd import capi
    g, _, wk = wake(qp_ctx, oracle, prec)
    wk.set_laser_gaussian(*LASER)
    want, st0, _ = wk.run(fields=FIELDS)
    e = capi.Envelope(qp_ctx, g.nx, g.ny, g.nz, g.dx, g.dy, g.dz, 10.0)
    e.set_gaussian(*LASER)
    wk.attach_envelope(e)
    got, st1, _ = wk.run(fields=FIELDS)
    assert st0 == 0 and st1 == 0
    for f in FIELDS:
        assert rel_err(got[f], want[f])[0] < 1e-9, f
    # the chi of every slice went to the envelope
    assert np.array_equal(e.get_chi(), got["chi"])
    wk.close()
    e.close()


def test_mixed_graph_run_stores_chi(qp_ctx, oracle):
    """Without a callback the mixed path replays captured slice steps; the chi they
    store equals the chi delivered by a callback run."""
    from quetzalpic_amd import capi
    g, _, wk = wake(qp_ctx, oracle, capi.QP_PREC_MIXED)
    e = capi.Envelope(qp_ctx, g.nx, g.ny, g.nz, g.dx, g.dy, g.dz, 10.0)
    e.set_gaussian(*LASER)
    wk.attach_envelope(e)
    assert '"envelope": 1' in wk.describe()
    _, st, _ = wk.run()
    assert st == 0
    chi_graph = e.get_chi()
    out, _, _ = wk.run(fields=("chi",))
    assert np.abs(chi_graph).max() > 0
    assert rel_err(chi_graph, out["chi"])[0] < 1e-12
    wk.close()
    e.close()


@pytest.mark.parametrize("prec,tol", [(0, 1e-9), (1, 2e-5)])
def test_coupled_step_matches_the_oracles(qp_ctx, oracle, prec, tol):
    """wake pass -> chi -> two envelope steps, against the slice loop oracle and the envelope oracle."""
    from quetzalpic_amd import capi
    g, (y, z, w), wk = wake(qp_ctx, oracle, prec)
    k0, dt = 10.0, 2.0
    e = capi.Envelope(qp_ctx, g.nx, g.ny, g.nz, g.dx, g.dy, g.dz, k0)
    e.set_gaussian(*LASER)
    wk.attach_envelope(e)
    _, st, _ = wk.run(fields=("psi",))
    assert st == 0
    e.advance(dt, steps=2, chi_iterations=2)
    got = e.get()
    wk.close()
    e.close()
    a0 = envo.gaussian(g.nx, g.ny, g.nz, g.dx, g.dy, g.dz, *LASER)
    ref, _, _ = oracle.wake_run(g, y, z, w, envo.a_sqr(a0), fields=("chi",))
    assert np.abs(ref["chi"]).max() > 0
    want = envo.advance(a0, ref["chi"], dt, k0, g.dx, g.dy, g.dz, steps=2, m=2)
    assert rel_err(got, want)[0] < tol
    assert np.abs(got - a0).max() > 1e3 * tol * np.abs(a0).max()   # the step did move the envelope


def test_vacuum_gaussian_beam_on_device(qp_ctx):
    """Spot, on-axis amplitude and Gouy phase at tau = Z_R (k0 = 20, w0 = 4)."""
    from quetzalpic_amd import capi
    k0, w0, box, n, nx, dxi = 20.0, 4.0, 48.0, 128, 40, 0.25
    zr = k0 * w0 ** 2 / 2
    d = box / n
    e = capi.Envelope(qp_ctx, nx, n, n, dxi, d, d, k0)
    e.set_gaussian(1.0, 6.0, 2 * w0, 2 * w0, nx * dxi / 2, box / 2, box / 2)
    a = e.get()
    e.advance(zr / 40, steps=40, chi_iterations=0)
    b = e.get()
    e.close()
    c, mid = nx // 2, n // 2

    def spot(p):
        q = np.abs(p) ** 2
        y = np.arange(n) * d
        s = q.sum()
        yc, zc = (q.sum(1) * y).sum() / s, (q.sum(0) * y).sum() / s
        return np.sqrt(2 * (q * ((y[:, None] - yc) ** 2 + (y[None, :] - zc) ** 2)).sum() / s)

    assert abs(spot(b[c]) / spot(a[c]) - np.sqrt(2)) < 0.005 * np.sqrt(2)
    assert abs(abs(b[c, mid, mid]) / abs(a[c, mid, mid]) - 1 / np.sqrt(2)) < 0.005
    assert abs(np.angle(b[c, mid, mid]) + np.pi / 4) < 0.01


def test_argument_errors(qp_ctx, oracle):
    from quetzalpic_amd import capi
    g, _, wk = wake(qp_ctx, oracle, capi.QP_PREC_FP64)
    e = capi.Envelope(qp_ctx, g.nx, g.ny, g.nz, g.dx, g.dy, g.dz, 10.0)
    bad = capi.Envelope(qp_ctx, g.nx - 1, g.ny, g.nz, g.dx, g.dy, g.dz, 10.0)
    with pytest.raises(capi.QpError):
        wk.attach_envelope(bad)
    with pytest.raises(capi.QpError):
        e.advance(0.0)
    with pytest.raises(capi.QpError):
        e.advance(1.0, chi_iterations=17)
    wk.attach_envelope(e)
    assert capi.load().qp_envelope_destroy(e.h) == capi.QP_ERR_STATE   # still attached
    wk.attach_envelope(None)
    b = capi.WakeBatch(qp_ctx, 2, g.nx, g.ny, g.nz, g.dx, g.dy, g.dz, g.ly, g.lz, K=5)
    with pytest.raises(capi.QpError):
        capi._check(capi.load().qp_wake_attach_envelope(b.h, e.h), "attach to a batch")
    b.close()
    bad.close()
    e.close()
    wk.close()


def test_mixed_graphs_follow_driver_changes(qp_ctx, oracle):
    """The mixed path keeps its captured slice steps between passes; changing the
    laser, a_sqr or the envelope must not replay steps captured for the old inputs
    (compared through the particle state after graph-replayed passes)."""
    from quetzalpic_amd import capi
    g, (y, z, w), wk = wake(qp_ctx, oracle, capi.QP_PREC_MIXED)
    laser2 = (0.9, 4.0, 5.0, 7.0, LX - 1.8, L / 2 - 1.0, L / 2 + 1.0)
    asq2 = oracle.gaussian_a_sqr(g, *laser2)

    def state(setup):
        fresh = capi.Wake(qp_ctx, g.nx, g.ny, g.nz, g.dx, g.dy, g.dz, g.ly, g.lz, K=5,
                          precision=capi.QP_PREC_MIXED)
        fresh.set_particles(y, z, w)
        setup(fresh)
        fresh.run()
        s = fresh.get_particles()
        fresh.close()
        return s

    def same(s1, s2):
        return all(rel_err(p, q)[0] < 1e-9 for p, q in zip(s1, s2))

    e = capi.Envelope(qp_ctx, g.nx, g.ny, g.nz, g.dx, g.dy, g.dz, 10.0)
    e.set_gaussian(*laser2)
    steps = [(lambda h: h.set_laser_gaussian(*LASER)), (lambda h: h.set_a_sqr(asq2)),
             (lambda h: h.attach_envelope(e))]
    wants = [state(steps[0]), state(steps[1]), state(steps[1])]
    assert not same(wants[0], wants[1])
    for setup, want in zip(steps, wants):
        setup(wk)
        wk.set_particles(y, z, w)
        wk.run()
        assert same(wk.get_particles(), want)
    wk.close()
    e.close()
