"""GPU parity of the plain FP64 slice loop against the FP64 oracle (C-ABI)."""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

FIELDS = ("psi", "rho", "jx", "jy", "jz", "ex", "ey", "ez", "by", "bz", "chi")


def small_case(oracle, L=20.0, n=32, lx=4.0, dx=0.1, a0=1.0, K=5, ppc=2):
    g = oracle.grid_from_box(lx, L, L, dx, L / n, L / n, K=K)
    y, z, w = oracle.init_lattice(g, ppc, ppc)
    a = oracle.gaussian_a_sqr(g, a0, 4.0, 6.0, 6.0, lx - 2.0, L / 2, L / 2)
    return g, y, z, w, a


def run_gpu(qp_ctx, g, y, z, w, a, precision=0, py=None, pz=None, rho_b=None, jx_b=None):
    from quetzalpic_amd import capi
    wk = capi.Wake(qp_ctx, g.nx, g.ny, g.nz, g.dx, g.dy, g.dz, g.ly, g.lz, K=g.K, D=g.D,
                   psi_threshold=g.psi_threshold, b_threshold=g.b_threshold, precision=precision)
    wk.set_particles(y, z, w, py, pz)
    wk.set_a_sqr(a)
    if rho_b is not None or jx_b is not None:
        wk.set_bunch_sources(rho_b, jx_b)
    out, st, ss = wk.run(fields=FIELDS)
    wk.close()
    return out, st, ss


@pytest.mark.parametrize("K", [5, 1, 0])
def test_fp64_wake_matches_oracle(qp_ctx, oracle, K):
    g, y, z, w, a = small_case(oracle, K=K)
    want, st0, _ = oracle.wake_run(g, y, z, w, a)
    got, st1, _ = run_gpu(qp_ctx, g, y, z, w, a)
    assert st0 == st1 == 0
    for f in FIELDS:
        mx, rms = rel_err(got[f], want[f])
        assert mx < 1e-10, (f, mx, rms)


def test_fp64_wake_generic_fft_grid(qp_ctx, oracle):
    """A 24 x 24 grid is not a power of two: the solves take the generic LDS kernels."""
    g, y, z, w, a = small_case(oracle, n=24)
    want, st0, _ = oracle.wake_run(g, y, z, w, a)
    got, st1, _ = run_gpu(qp_ctx, g, y, z, w, a)
    assert st0 == st1 == 0
    for f in FIELDS:
        mx, rms = rel_err(got[f], want[f])
        assert mx < 1e-10, (f, mx, rms)
