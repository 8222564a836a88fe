"""GPU parity of the periodic 2D solve and Fourier2d transforms (C-ABI) against
the FP64 oracle and against numpy's FFT, including closed forms."""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

SIZES = [(64, 64), (96, 128), (100, 200), (256, 256), (7, 9), (2, 3), (512, 512), (30, 50)]


@pytest.mark.parametrize("n1,n2", SIZES)
def test_fourier2d_matches_numpy(qp_ctx, n1, n2):
    from quetzalpic_amd import capi
    rng = np.random.default_rng(n1 * 1000 + n2)
    a = rng.standard_normal((n1, n2))
    f = capi.Fourier2d(qp_ctx, n1, n2)
    s = f.forward_transform(a)
    ref = np.fft.rfft2(a)
    assert rel_err(s, ref)[0] < 1e-14
    back = f.backward_transform(ref)
    assert rel_err(back, np.fft.irfft2(ref, s=(n1, n2)) * n1 * n2)[0] < 1e-14
    f.close()


@pytest.mark.parametrize("n1,n2", SIZES)
@pytest.mark.parametrize("D", [0.0, 10.0])
def test_solve2d_matches_oracle(qp_ctx, oracle, n1, n2, D):
    from quetzalpic_amd import capi
    rng = np.random.default_rng(7 + n1 + n2)
    src = rng.standard_normal((n1, n2))
    d1, d2 = 20.0 / n1, 24.0 / n2
    f = capi.Fourier2d(qp_ctx, n1, n2)
    got = f.solve_poisson_equation(src, d1, d2, D)
    want = oracle.solve_poisson(src, d1, d2, D)
    assert rel_err(got, want)[0] < 1e-13
    f.close()


def test_solve2d_single_mode_closed_form(qp_ctx):
    """A single Fourier mode cos(2 pi (m j/n1 + l k/n2)) is returned divided by its eigenvalue."""
    from quetzalpic_amd import capi
    n1, n2, d1, d2 = 64, 96, 0.3, 0.25
    j = np.arange(n1)[:, None]
    k = np.arange(n2)[None, :]
    f = capi.Fourier2d(qp_ctx, n1, n2)
    for (m, l, D) in [(1, 0, 0.0), (3, 5, 0.0), (0, 7, 10.0), (31, 47, 10.0)]:
        src = np.cos(2 * np.pi * (m * j / n1 + l * k / n2))
        lam = -D + (2 * np.cos(2 * np.pi * m / n1) - 2) / d1 ** 2 + (2 * np.cos(2 * np.pi * l / n2) - 2) / d2 ** 2
        got = f.solve_poisson_equation(src, d1, d2, D) / (n1 * n2)
        assert np.abs(got - src / lam).max() < 1e-12 * np.abs(src / lam).max()
    # (0,0) mode dropped, also for D != 0
    got = f.solve_poisson_equation(np.ones((n1, n2)), d1, d2, 10.0)
    assert np.abs(got).max() < 1e-12
    f.close()


@pytest.mark.parametrize("n1,n2", [(16, 16), (32, 64), (256, 256), (128, 1024), (1024, 1024), (512, 32),
                                   (48, 48), (96, 192), (192, 192), (384, 96), (48, 384), (96, 256)])
def test_wave_and_generic_kernels_agree(qp_ctx, n1, n2, monkeypatch):
    """Grids with sides 2^p or 3 * 2^p (radix-6 first pass) use the register-
    resident kernels; QP_FFT=generic forces the generic LDS kernels on the same
    grid.  Both must give the same solve."""
    from quetzalpic_amd import capi
    rng = np.random.default_rng(n1 + 3 * n2)
    src = rng.standard_normal((n1, n2))
    d1, d2 = 24.0 / n1, 24.0 / n2
    fw = capi.Fourier2d(qp_ctx, n1, n2)
    monkeypatch.setenv("QP_FFT", "generic")
    fg = capi.Fourier2d(qp_ctx, n1, n2)
    monkeypatch.delenv("QP_FFT")
    for D in (0.0, 10.0):
        a = fw.solve_poisson_equation(src, d1, d2, D)
        b = fg.solve_poisson_equation(src, d1, d2, D)
        assert rel_err(a, b)[0] < 1e-14
    fw.close()
    fg.close()
