"""Pins of the oracle's transform and periodic solve (CPU only).

The reference delegates these to FFTW (src/Fourier2d.cpp:33-34, un-vendored);
its conventions (unnormalised r2c/c2r, n1 x (n2/2+1) half spectrum) are pinned
here against numpy.fft, single-mode closed forms and an independent 5-point
stencil residual.
"""
import numpy as np
import pytest


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 7, 8, 9, 12, 15, 16, 25, 27, 49, 64, 96, 100, 101, 128, 192, 200,
                               256, 1024])
def test_fft_matches_numpy(oracle, n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    fwd = oracle.fft(x, -1)
    inv = oracle.fft(x, +1)
    assert np.abs(fwd - np.fft.fft(x)).max() <= 1e-13 * max(1.0, np.abs(x).sum())
    assert np.abs(inv - np.fft.ifft(x) * n).max() <= 1e-13 * max(1.0, np.abs(x).sum())


@pytest.mark.parametrize("shape", [(8, 8), (6, 10), (7, 9), (64, 64), (96, 128), (100, 200), (5, 3), (1, 4)])
def test_rfft2_irfft2_match_numpy(oracle, shape):
    rng = np.random.default_rng(sum(shape))
    a = rng.standard_normal(shape)
    s = oracle.rfft2(a)
    assert np.abs(s - np.fft.rfft2(a)).max() < 1e-12 * np.abs(a).sum()
    spec = np.fft.rfft2(a)
    back = oracle.irfft2(spec, shape[1])
    assert np.abs(back - a * a.size).max() < 1e-12 * np.abs(a).sum()


def test_c2r_ignores_imaginary_dc_and_nyquist(oracle):
    """FFTW c2r semantics: the imaginary parts of the k2 = 0 and k2 = n2/2 terms do not contribute."""
    rng = np.random.default_rng(3)
    a = rng.standard_normal((8, 8))
    s = np.fft.rfft2(a)
    s2 = s.copy()
    s2[:, 0] += 1j * rng.standard_normal(8)
    s2[:, 4] += 1j * rng.standard_normal(8)
    # make the perturbation Hermitian-inconsistent only in those columns
    b1 = oracle.irfft2(s, 8)
    b2 = oracle.irfft2(s2, 8)
    ref = np.fft.irfft2(s2, s=(8, 8)) * 64
    assert np.abs(b2 - ref).max() < 1e-11
    assert np.abs(b1 - a * 64).max() < 1e-11


def eig(m, l, n1, n2, d1, d2, D):
    return -D + (2 * np.cos(2 * np.pi * m / n1) - 2) / d1 ** 2 + (2 * np.cos(2 * np.pi * l / n2) - 2) / d2 ** 2


@pytest.mark.parametrize("m,l,D", [(1, 0, 0.0), (0, 1, 0.0), (3, 5, 0.0), (7, 2, 10.0), (31, 47, 10.0), (0, 0, 10.0)])
def test_solve_single_mode_closed_form(oracle, m, l, D):
    n1, n2, d1, d2 = 64, 96, 0.3125, 0.25
    j = np.arange(n1)[:, None]
    k = np.arange(n2)[None, :]
    for phase in (np.cos, np.sin):
        src = phase(2 * np.pi * (m * j / n1 + l * k / n2))
        got = oracle.solve_poisson(src, d1, d2, D) / (n1 * n2)
        if m == 0 and l == 0:
            want = np.zeros_like(src)  # the (0,0) mode is dropped, also for D != 0
        else:
            want = src / eig(m, l, n1, n2, d1, d2, D)
        assert np.abs(got - want).max() < 1e-12 * max(1.0, np.abs(want).max())


@pytest.mark.parametrize("D", [0.0, 10.0])
@pytest.mark.parametrize("shape", [(32, 32), (30, 50), (64, 96), (100, 100)])
def test_solve_residual_5point_stencil(oracle, D, shape):
    """(Delta_h - D) u = src - mean(src) with an independent periodic 5-point stencil."""
    n1, n2 = shape
    d1, d2 = 20.0 / n1, 24.0 / n2
    rng = np.random.default_rng(n1 + 7 * n2)
    src = rng.standard_normal(shape)
    u = oracle.solve_poisson(src, d1, d2, D) / (n1 * n2)
    lap = ((np.roll(u, -1, 0) - 2 * u + np.roll(u, 1, 0)) / d1 ** 2 +
           (np.roll(u, -1, 1) - 2 * u + np.roll(u, 1, 1)) / d2 ** 2)
    res = lap - D * u - (src - src.mean())
    if D != 0:
        # the dropped zero mode: (Delta - D) u has zero mean, so only src's mean is removed
        pass
    assert np.abs(res).max() < 1e-10 * np.abs(src).max() * max(1.0, 1.0 / min(d1, d2) ** 2)
    assert abs(u.mean()) < 1e-12 * np.abs(u).max()


def test_solve_matches_numpy_spectral(oracle):
    n1, n2, d1, d2 = 48, 40, 0.5, 0.4
    rng = np.random.default_rng(11)
    src = rng.standard_normal((n1, n2))
    for D in (0.0, 10.0):
        s = np.fft.rfft2(src)
        m = np.arange(n1)[:, None]
        l = np.arange(n2 // 2 + 1)[None, :]
        lam = eig(m, l, n1, n2, d1, d2, D)
        lam[0, 0] = 1.0
        s = s / lam
        s[0, 0] = 0
        want = np.fft.irfft2(s, s=(n1, n2)) * n1 * n2
        got = oracle.solve_poisson(src, d1, d2, D)
        assert np.abs(got - want).max() < 1e-11 * np.abs(want).max()
