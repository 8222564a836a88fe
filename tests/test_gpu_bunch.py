"""GPU tests of the driver bunch on the device (qp_bunch_*, qp_wake_attach_bunch).

The wake streams an attached bunch: slice i's rho_b, jx_b are deposited into
two planes right before the slice and the particles of x-cell i gather the
fields of slices i and i+1 once they are final.  Held to the oracle
(deposit_bunch_densities, deposit_wakefield_to_bunch and
advance_bunch_particles without QED, src/System_3d.cpp:843-946, pinned on the
CPU in tests/test_oracle_bunch.py):
  * the streamed wake equals the wake fed the oracle's 3D source arrays;
  * the gathered fields and chi equal the oracle's gather of the 3D field
    arrays the same pass delivers, each array with its staggered origin;
  * the push equals the oracle's push;
  * three beam time steps (deposit -> wake -> gather -> push) equal the oracle
    chain (FP64 path), and the mixed path follows the FP64 one.
"""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

L, LX, DX = 20.0, 6.0, 0.1
FS = 50.0
FIELDS3 = ("ex", "ey", "ez", "by", "bz")
WAKE_FIELDS = ("psi", "rho", "jx", "ex", "ey", "ez", "by", "bz", "chi")


def make_bunch(g, n=4000, seed=7, scale=1.0):
    """Electrons (cmr = -1) around (3.5, L/2, L/2) plus a few photons, a few
    particles outside the x range and a few on the periodic seam."""
    rng = np.random.default_rng(seed)
    b = {"x": rng.normal(3.5, 0.6, n), "y": rng.normal(L / 2, 1.0, n), "z": rng.normal(L / 2 + 0.3, 0.8, n)}
    b["x"][:20] = rng.uniform(-0.5, 0.0, 20)                       # behind the box
    b["x"][20:40] = rng.uniform((g.nx - 1) * g.dx, LX + 0.5, 20)   # last cell and beyond
    b["y"][40:60] = rng.uniform(g.ly - g.dy, g.ly, 20)              # periodic seam in y
    b["z"][60:80] = rng.uniform(-g.dz, 0.0, 20)                     # k1 = -1 wraps
    b["px"] = rng.normal(1000.0, 20.0, n)
    b["py"] = rng.normal(0.0, 0.5, n)
    b["pz"] = rng.normal(0.0, 0.5, n)
    b["photon"] = np.zeros(n, np.int32)
    b["photon"][80:120] = 1
    b["gamma"] = np.sqrt(1 + b["px"] ** 2 + b["py"] ** 2 + b["pz"] ** 2)
    ph = b["photon"] == 1
    b["gamma"][ph] = np.sqrt(b["px"] ** 2 + b["py"] ** 2 + b["pz"] ** 2)[ph]
    b["cmr"] = np.full(n, -1.0)
    b["q"] = np.full(n, -0.02 * scale)
    return b


def grid(oracle, n=32, K=5):
    return oracle.grid_from_box(LX, L, L, DX, L / n, L / n, K=K)


def plasma_wake(qp_ctx, oracle, g, prec):
    from quetzalpic_amd import capi
    y, z, w = oracle.init_lattice(g, 2, 2)
    wk = capi.Wake(qp_ctx, g.nx, g.ny, g.nz, g.dx, g.dy, g.dz, g.ly, g.lz, K=g.K, precision=prec)
    wk.set_particles(y, z, w)
    return wk, (y, z, w)


def dev_bunch(qp_ctx, g, b):
    from quetzalpic_amd import capi
    db = capi.Bunch(qp_ctx, g.nx, g.ny, g.nz, g.dx, g.dy, g.dz, FS)
    db.set_particles(b)
    return db


@pytest.mark.parametrize("prec,n,tol", [(0, 32, 1e-11), (1, 32, 1e-9), (0, 24, 1e-11)])
def test_streamed_sources_equal_3d_sources(qp_ctx, oracle, prec, n, tol):
    g = grid(oracle, n)
    b = make_bunch(g)
    rho3, jx3 = oracle.bunch_deposit(g, b)
    assert np.abs(rho3).max() > 0.1
    db = dev_bunch(qp_ctx, g, b)
    drho, djx = db.deposit()
    assert np.abs(drho - rho3).max() <= 1e-13 * np.abs(rho3).max()
    assert np.abs(djx - jx3).max() <= 1e-13 * np.abs(jx3).max()
    wa, _ = plasma_wake(qp_ctx, oracle, g, prec)
    wa.set_bunch_sources(rho3, jx3)
    want, st0, _ = wa.run(fields=WAKE_FIELDS)
    wb, _ = plasma_wake(qp_ctx, oracle, g, prec)
    wb.attach_bunch(db)
    assert '"bunch": 1' in wb.describe()
    got, st1, _ = wb.run(fields=WAKE_FIELDS)
    assert st0 == 0 and st1 == 0
    for f in WAKE_FIELDS:
        assert rel_err(got[f], want[f])[0] < tol, (f, rel_err(got[f], want[f]))
    wa.close()
    wb.close()
    db.close()


@pytest.mark.parametrize("prec", [0, 1])
def test_gathered_fields_match_the_oracle(qp_ctx, oracle, prec):
    g = grid(oracle)
    b = make_bunch(g)
    db = dev_bunch(qp_ctx, g, b)
    wk, _ = plasma_wake(qp_ctx, oracle, g, prec)
    wk.attach_bunch(db)
    out, st, _ = wk.run(fields=FIELDS3)
    assert st == 0
    got = db.get_fields()
    want = oracle.bunch_fields(g, b, out, FS)
    for f in FIELDS3 + ("chi",):
        scale = np.abs(want[f]).max()
        assert scale > 0, f
        assert np.abs(got[f] - want[f]).max() <= 1e-12 * scale, (f, np.abs(got[f] - want[f]).max() / scale)
    # no array reaches these (e_x, staggered by -dx/2, still reaches x in [-dx/2, 0))
    outside = (b["x"] < -0.5 * g.dx) | (np.floor(b["x"] / g.dx) >= g.nx - 1)
    assert outside.sum() >= 20 and not any(got[f][outside].any() for f in FIELDS3)
    assert got["ex"][(b["x"] >= -0.5 * g.dx) & (b["x"] < 0)].all()
    wk.close()
    db.close()


def test_graph_replayed_pass_gathers_the_same(qp_ctx, oracle):
    """Without a callback the mixed path replays captured slice steps (slice index
    from the control block); the gathered fields equal a callback pass."""
    from quetzalpic_amd import capi
    g = grid(oracle)
    b = make_bunch(g)
    db = dev_bunch(qp_ctx, g, b)
    wk, _ = plasma_wake(qp_ctx, oracle, g, capi.QP_PREC_MIXED)
    wk.attach_bunch(db)
    wk.run()
    f_graph = db.get_fields()
    wk.run(fields=("psi",))
    f_cb = db.get_fields()
    # (1e-9: the atomic sums of the deposits round differently from pass to pass,
    # and e_x = d psi / dx amplifies that; a stale captured pointer would be O(1))
    for f in FIELDS3:
        assert rel_err(f_graph[f], f_cb[f])[0] < 1e-9, f
    # more particles than the arrays held: the arrays move, the graphs are captured again
    b2 = make_bunch(g, n=6000, seed=8)
    db.set_particles(b2)
    wk.run()
    f2 = db.get_fields()
    fresh, _ = plasma_wake(qp_ctx, oracle, g, capi.QP_PREC_MIXED)
    db2 = dev_bunch(qp_ctx, g, b2)
    fresh.attach_bunch(db2)
    fresh.run(fields=("psi",))
    f3 = db2.get_fields()
    for f in FIELDS3:
        assert rel_err(f2[f], f3[f])[0] < 1e-9, f
    for o in (wk, fresh, db, db2):
        o.close()


def test_push_matches_the_oracle(qp_ctx, oracle):
    g = grid(oracle)
    b = make_bunch(g)
    db = dev_bunch(qp_ctx, g, b)
    wk, _ = plasma_wake(qp_ctx, oracle, g, 0)
    wk.attach_bunch(db)
    wk.run()
    fields = db.get_fields()
    db.push(0.8)
    got = db.get_particles()
    want = oracle.bunch_push(b, fields, 0.8)
    for k in db.STATE:
        assert np.abs(got[k] - want[k]).max() <= 1e-14 * max(np.abs(want[k]).max(), 1.0), k
    # the re-bucketed bunch deposits what the oracle deposits from the pushed state
    want_n = dict(want, q=b["q"], cmr=b["cmr"], photon=b["photon"])
    rho3, _ = oracle.bunch_deposit(g, want_n)
    drho, _ = db.deposit()
    assert np.abs(drho - rho3).max() <= 1e-13 * np.abs(rho3).max()
    wk.close()
    db.close()


def oracle_chain(oracle, g, b, y, z, w, steps, dt):
    a = np.zeros((g.nx, g.ny, g.nz))
    for _ in range(steps):
        rho3, jx3 = oracle.bunch_deposit(g, b)
        out, st, _ = oracle.wake_run(g, y, z, w, a, rho_b=rho3, jx_b=jx3, fields=FIELDS3)
        assert st == 0
        f = oracle.bunch_fields(g, b, out, FS)
        b = dict(oracle.bunch_push(b, f, dt))
    return b, f


def device_chain(qp_ctx, oracle, g, b, prec, steps, dt):
    db = dev_bunch(qp_ctx, g, b)
    wk, _ = plasma_wake(qp_ctx, oracle, g, prec)
    wk.attach_bunch(db)
    for _ in range(steps):
        _, st, _ = wk.run()
        assert st == 0
        f = db.get_fields()
        db.push(dt)
    s = db.get_particles()
    wk.close()
    db.close()
    return s, f


def test_time_loop_matches_the_oracle(qp_ctx, oracle):
    """Three beam steps of run() (src/System_3d.cpp:199-236) without QED."""
    g = grid(oracle, 24)
    b = make_bunch(g, n=2000, scale=2.0)
    y, z, w = oracle.init_lattice(g, 2, 2)
    dt = 2.0
    want, fw = oracle_chain(oracle, g, b, y, z, w, 3, dt)
    got, fg = device_chain(qp_ctx, oracle, g, b, 0, 3, dt)
    for f in FIELDS3:
        assert rel_err(fg[f], fw[f])[0] < 1e-9, f
    for k in ("x", "y", "z", "px", "py", "pz", "gamma"):
        assert rel_err(got[k], want[k])[0] < 1e-12, k
    assert np.abs(got["py"] - b["py"]).max() > 1e-3   # the bunch did feel the wake
    mixed, fm = device_chain(qp_ctx, oracle, g, b, 1, 3, dt)
    for f in FIELDS3:
        assert rel_err(fm[f], fg[f])[1] < 1e-5, (f, rel_err(fm[f], fg[f]))
    assert rel_err(mixed["py"], got["py"])[0] < 1e-5


def test_argument_errors(qp_ctx, oracle):
    from quetzalpic_amd import capi
    g = grid(oracle)
    wk, _ = plasma_wake(qp_ctx, oracle, g, 0)
    db = dev_bunch(qp_ctx, g, make_bunch(g, n=100))
    bad = capi.Bunch(qp_ctx, g.nx, g.ny + 1, g.nz, g.dx, g.dy, g.dz, FS)
    with pytest.raises(capi.QpError):
        wk.attach_bunch(bad)
    wk.attach_bunch(db)
    assert capi.load().qp_bunch_destroy(db.h) == capi.QP_ERR_STATE
    bt = capi.WakeBatch(qp_ctx, 2, g.nx, g.ny, g.nz, g.dx, g.dy, g.dz, g.ly, g.lz, K=5)
    with pytest.raises(capi.QpError):
        capi._check(capi.load().qp_wake_attach_bunch(bt.h, db.h), "attach to a batch")
    bt.close()
    wk.close()          # destroying the wake detaches
    assert capi.load().qp_bunch_destroy(bad.h) == capi.QP_OK
    bad.h = None
    db.close()
