"""GPU tests of batched wakes (qp_wake_create_batch, mixed precision).

Each member of a batch must give the result of the same wake run alone in its
own handle (which the mixed-path tests hold to the FP64 oracle); members differ
in laser, density, particle count and thermal momenta; members stop
independently at the slice the oracle stops at.
"""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

FIELDS = ("psi", "rho", "jx", "jy", "jz", "ex", "ey", "ez", "by", "bz", "chi")
L, LX, DX = 20.0, 4.0, 0.1


def member_inputs(oracle, g, kind):
    """Particles (y, z, w, py, pz) and laser (a0, xs, ys, zs, x0, y0, z0) of member `kind`."""
    if kind == 0:
        y, z, w = oracle.init_lattice(g, 2, 2)
        py = pz = None
        laser = (1.0, 4.0, 6.0, 6.0, LX - 2.0, L / 2, L / 2)
    elif kind == 1:
        y, z, w = oracle.init_lattice(g, 2, 2)
        rng = np.random.default_rng(11)
        py, pz = rng.normal(0.0, 0.01, len(y)), rng.normal(0.0, 0.01, len(y))
        laser = (1.5, 3.0, 5.0, 7.0, LX - 1.5, L / 2 + 2.0, L / 2 - 1.0)
    else:
        y, z, w = oracle.init_lattice(g, 1, 1)   # a quarter of the particles of the others
        w = w * 1.3
        py = pz = None
        laser = (0.8, 4.0, 6.0, 6.0, LX - 2.5, L / 2, L / 2 + 1.5)
    return y, z, w, py, pz, laser


def run_single(qp_ctx, g, inp):
    from quetzalpic_amd import capi
    y, z, w, py, pz, laser = inp
    wk = capi.Wake(qp_ctx, g.nx, g.ny, g.nz, g.dx, g.dy, g.dz, g.ly, g.lz, K=g.K, D=g.D,
                   psi_threshold=g.psi_threshold, b_threshold=g.b_threshold, precision=capi.QP_PREC_MIXED)
    wk.set_particles(y, z, w, py, pz)
    wk.set_laser_gaussian(*laser)
    out, st, ss = wk.run(fields=FIELDS)
    wk.close()
    return out, st, ss


def run_batch(qp_ctx, g, inps, use_graph=1):
    from quetzalpic_amd import capi
    b = capi.WakeBatch(qp_ctx, len(inps), g.nx, g.ny, g.nz, g.dx, g.dy, g.dz, g.ly, g.lz, K=g.K, D=g.D,
                       psi_threshold=g.psi_threshold, b_threshold=g.b_threshold, use_graph=use_graph)
    for m, (y, z, w, py, pz, laser) in enumerate(inps):
        b.set_particles(m, y, z, w, py, pz)
        b.set_laser_gaussian(m, *laser)
    outs, st, stops = b.run(fields=FIELDS)
    desc = b.describe()
    b.close()
    return outs, st, stops, desc


@pytest.mark.parametrize("n", [32, 24])
def test_batch_members_equal_single_runs(qp_ctx, oracle, n):
    """32 x 32 takes the register-resident FFT kernels, 24 x 24 the generic ones."""
    import json
    g = oracle.grid_from_box(LX, L, L, DX, L / n, L / n, K=5)
    inps = [member_inputs(oracle, g, k) for k in range(3)]
    outs, st, stops, desc = run_batch(qp_ctx, g, inps)
    assert json.loads(desc)["members"] == 3
    assert st == 0 and stops == [-1, -1, -1]
    for k, inp in enumerate(inps):
        single, st1, _ = run_single(qp_ctx, g, inp)
        assert st1 == 0
        for f in FIELDS:
            assert rel_err(outs[k][f], single[f])[0] < 1e-9, (k, f, rel_err(outs[k][f], single[f]))
    # member 0 against the FP64 oracle (same Gaussian as a host <a^2> array)
    y, z, w, _, _, laser = inps[0]
    want, _, _ = oracle.wake_run(g, y, z, w, oracle.gaussian_a_sqr(g, *laser), fields=("psi", "ex"))
    for f in ("psi", "ex"):
        mx, rms = rel_err(outs[0][f], want[f])
        assert mx < 1e-5 and rms < 1e-5, (f, mx, rms)


def test_batch_without_graph_equals_graph(qp_ctx, oracle):
    g = oracle.grid_from_box(LX, L, L, DX, L / 32, L / 32, K=5)
    inps = [member_inputs(oracle, g, k) for k in range(3)]
    a, _, _, _ = run_batch(qp_ctx, g, inps, use_graph=1)
    b, _, _, _ = run_batch(qp_ctx, g, inps, use_graph=0)
    for k in range(3):
        for f in FIELDS:
            assert rel_err(a[k][f], b[k][f])[0] < 1e-9, (k, f)


def test_batch_members_stop_independently(qp_ctx, oracle):
    """With THRESHOLD_B = 0.3 the a0 = 8 member trips the stop rule at the slice
    the oracle stops at; the others run to the end, unchanged."""
    g = oracle.grid_from_box(LX, L, L, DX, L / 32, L / 32, K=5, b_threshold=0.3)
    y, z, w = oracle.init_lattice(g, 2, 2)
    lasers = [(1.0, 4.0, 6.0, 6.0, LX - 2.0, L / 2, L / 2), (8.0, 4.0, 6.0, 6.0, LX - 2.0, L / 2, L / 2),
              (1.2, 4.0, 6.0, 6.0, LX - 2.0, L / 2 + 1.0, L / 2)]
    inps = [(y, z, w, None, None, las) for las in lasers]
    outs, st, stops, _ = run_batch(qp_ctx, g, inps)
    _, st0, ss0 = oracle.wake_run(g, y, z, w, oracle.gaussian_a_sqr(g, *lasers[1]), fields=("by",))
    assert st0 == 1
    assert st == 1 and stops[0] == -1 and stops[2] == -1 and stops[1] == ss0
    for k, inp in enumerate(inps):
        single, st1, ss1 = run_single(qp_ctx, g, inp)
        assert ss1 == stops[k]
        for f in FIELDS:
            assert rel_err(outs[k][f], single[f])[0] < 1e-9, (k, f)


def test_batch_staged_spectrum_reads(qp_ctx, oracle):
    """Eight 256 x 256 members: the batch's paired B_y / B_z inverse row pass
    launches 64 x 2 x 8 = 1024 blocks, which read the column-major spectrum
    through LDS (staged reads); a single 256 x 256 wake launches 128 blocks,
    which read it directly.  Both must give the same fields."""
    fields = ("psi", "jx", "ex", "by", "bz")
    g = oracle.grid_from_box(LX, L, L, 0.2, L / 256, L / 256, K=5)
    assert (g.ny, g.nz) == (256, 256)
    inps = [member_inputs(oracle, g, k % 3) for k in range(8)]
    from quetzalpic_amd import capi
    b = capi.WakeBatch(qp_ctx, len(inps), g.nx, g.ny, g.nz, g.dx, g.dy, g.dz, g.ly, g.lz, K=g.K, D=g.D,
                       psi_threshold=g.psi_threshold, b_threshold=g.b_threshold, use_graph=1)
    for m, (y, z, w, py, pz, laser) in enumerate(inps):
        b.set_particles(m, y, z, w, py, pz)
        b.set_laser_gaussian(m, *laser)
    outs, st, stops = b.run(fields=fields)
    b.close()
    assert st == 0 and stops == [-1] * 8
    for k in range(3):
        y, z, w, py, pz, laser = inps[k]
        wk = capi.Wake(qp_ctx, g.nx, g.ny, g.nz, g.dx, g.dy, g.dz, g.ly, g.lz, K=g.K, D=g.D,
                       psi_threshold=g.psi_threshold, b_threshold=g.b_threshold, precision=capi.QP_PREC_MIXED)
        wk.set_particles(y, z, w, py, pz)
        wk.set_laser_gaussian(*laser)
        single, st1, _ = wk.run(fields=fields)
        wk.close()
        assert st1 == 0
        for f in fields:
            assert rel_err(outs[k][f], single[f])[0] < 1e-9, (k, f, rel_err(outs[k][f], single[f]))
