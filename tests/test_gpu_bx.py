"""Longitudinal B (extension with no reference counterpart).

In the quasistatic limit the x component of curl curl B = curl j reads
lap_perp B_x = dz j_y - dy j_z.  The mixed path solves it per slice from the
slice's output currents (j_y on (dy/2, 0), j_z on (0, dz/2), B_x on the cell
corners).  Pinned by the discrete identity on the delivered planes: the
periodic 5-point Laplacian of B_x equals the staggered curl of the same
slice's j_y, j_z minus its mean (the dropped (0,0) mode), and B_x has zero
mean.  A centred driver with K = 0 has no transverse current and gives B_x = 0.
"""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

L, LX, DX = 20.0, 4.0, 0.1


def curl_source(jy, jz, dy, dz):
    return (np.roll(jy, -1, axis=1) - jy) / dz - (np.roll(jz, -1, axis=0) - jz) / dy


def lap5(b, dy, dz):
    return ((np.roll(b, -1, 0) - 2 * b + np.roll(b, 1, 0)) / dy ** 2 +
            (np.roll(b, -1, 1) - 2 * b + np.roll(b, 1, 1)) / dz ** 2)


def run(qp_ctx, oracle, n, longitudinal_b, K=5, y0=L / 2 + 2.5, z0=L / 2 - 1.5):
    from quetzalpic_amd import capi
    g = oracle.grid_from_box(LX, L, L, DX, L / n, L / n, K=K)
    y, z, w = oracle.init_lattice(g, 2, 2)
    wk = capi.Wake(qp_ctx, g.nx, g.ny, g.nz, g.dx, g.dy, g.dz, g.ly, g.lz, K=K, precision=capi.QP_PREC_MIXED,
                   longitudinal_b=longitudinal_b)
    wk.set_particles(y, z, w)
    wk.set_laser_gaussian(1.5, 4.0, 5.0, 6.0, LX - 2.0, y0, z0)
    fields = ("psi", "ex", "jy", "jz") + (("bx",) if longitudinal_b else ())
    out, st, _ = wk.run(fields=fields)
    wk.close()
    assert st == 0
    return g, out


@pytest.mark.parametrize("n", [32, 24])
def test_bx_satisfies_the_curl_identity(qp_ctx, oracle, n):
    g, out = run(qp_ctx, oracle, n, 1)
    bx, jy, jz = out["bx"], out["jy"], out["jz"]
    assert np.abs(bx).max() > 0
    for i in range(g.nx):
        s = curl_source(jy[i], jz[i], g.dy, g.dz)
        s -= s.mean()
        scale = max(np.abs(s).max(), 1e-300)
        assert np.abs(lap5(bx[i], g.dy, g.dz) - s).max() <= 1e-9 * scale + 1e-300, i
        assert abs(bx[i].mean()) <= 1e-12 * max(np.abs(bx[i]).max(), 1e-300), i


def test_bx_leaves_the_other_fields_alone(qp_ctx, oracle):
    _, a = run(qp_ctx, oracle, 32, 1)
    _, b = run(qp_ctx, oracle, 32, 0)
    for f in ("psi", "ex", "jy", "jz"):
        assert rel_err(a[f], b[f])[0] < 1e-9, f


def test_bx_vanishes_without_transverse_current(qp_ctx, oracle):
    _, out = run(qp_ctx, oracle, 32, 1, K=0, y0=L / 2, z0=L / 2)
    assert np.abs(out["jy"]).max() == 0 and np.abs(out["jz"]).max() == 0
    assert np.abs(out["bx"]).max() == 0


def test_bx_is_a_mixed_path_extension(qp_ctx, oracle):
    from quetzalpic_amd import capi
    g = oracle.grid_from_box(LX, L, L, DX, L / 32, L / 32, K=5)
    with pytest.raises(capi.QpError):
        capi.Wake(qp_ctx, g.nx, g.ny, g.nz, g.dx, g.dy, g.dz, g.ly, g.lz, K=5, precision=capi.QP_PREC_FP64,
                  longitudinal_b=1)
