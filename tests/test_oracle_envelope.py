"""Pins of the envelope oracle (oracle/envelope.py), CPU only.

The envelope advance has no counterpart in the reference, so the oracle is
pinned by (1) dense real-space solves of the discrete equations on a tiny
grid, assembled from the stencil and the slice coupling without any FFT: with
chi a at the old time (m = 0) and fully Crank-Nicolson (m -> infinity, the
limit of the chi corrections), (2) the vacuum Gaussian beam: spot
w0 sqrt(1 + (tau/Z_R)^2), on-axis amplitude a0 w0/w and Gouy phase
-atan(tau/Z_R), Z_R = k0 w0^2/2, (3) a parabolic channel chi = 4 r^2/w0^4 that
holds the spot of a Gaussian of radius w0, and (4) a plateau in uniform chi,
whose phase turns by -2 atan(dt chi/(4 k0)) per Crank-Nicolson step (tending to
-chi tau/(2 k0)) at constant amplitude.  A wrong sign of i, of d_x or of chi, a
factor 2 in the chi or the diffraction term, transposed axes or the sweep run
the wrong way fail one of them.
"""
import numpy as np
import pytest

from oracle import envelope as env


def lap_matrix(ny, nz, dy, dz):
    """Dense periodic 5-point Laplacian on a [j][k] plane, from the stencil."""
    nn = ny * nz
    L = np.zeros((nn, nn))
    for j in range(ny):
        for k in range(nz):
            r = j * nz + k
            L[r, r] += -2 / dy ** 2 - 2 / dz ** 2
            L[r, ((j + 1) % ny) * nz + k] += 1 / dy ** 2
            L[r, ((j - 1) % ny) * nz + k] += 1 / dy ** 2
            L[r, j * nz + (k + 1) % nz] += 1 / dz ** 2
            L[r, j * nz + (k - 1) % nz] += 1 / dz ** 2
    return L


def dense_step(a, chi, dt, k0, dx, dy, dz, implicit_chi):
    """One step as one linear system in the unknowns a+ of every slice:
    (2/dx)(d_{i+1} - d_i) + 2 i k0 d_i = (dt/2)(chi_i - L)(a+_i + a_i), d_i = a+_i - a_i,
    or, with c_i = (min chi_i + max chi_i)/2, with dt (chi_i - c_i) a_i
    + (dt/2)(c_i - L)(a+_i + a_i) on the right."""
    nx, ny, nz = a.shape
    nn = ny * nz
    L = lap_matrix(ny, nz, dy, dz)
    I = np.eye(nn)
    A = np.zeros((nx * nn, nx * nn), complex)
    b = np.zeros(nx * nn, complex)
    av = a.reshape(nx, nn)
    alpha = -2 / dx + 2j * k0
    for i in range(nx):
        c = 0.5 * (chi[i].min() + chi[i].max())
        X = np.diag(chi[i].ravel())
        Ximp = X if implicit_chi else c * I
        rows = slice(i * nn, (i + 1) * nn)
        # alpha a+_i + (2/dx) a+_{i+1} + (dt/2)(L - Ximp) a+_i = known terms
        A[rows, rows] = alpha * I + 0.5 * dt * (L - Ximp)
        rhs = alpha * av[i] - 0.5 * dt * (L - Ximp) @ av[i]
        if not implicit_chi:
            rhs = rhs + dt * (X - c * I) @ av[i]
        if i + 1 < nx:
            A[rows, (i + 1) * nn:(i + 2) * nn] = (2 / dx) * I
            rhs = rhs + (2 / dx) * av[i + 1]
        b[rows] = rhs
    return np.linalg.solve(A, b).reshape(nx, ny, nz)


@pytest.mark.parametrize("m", [0, 60])
def test_step_equals_dense_solve(m):
    rng = np.random.default_rng(5)
    nx, ny, nz = 4, 5, 3
    a = rng.normal(size=(nx, ny, nz)) + 1j * rng.normal(size=(nx, ny, nz))
    chi = rng.uniform(0.0, 2.0, size=(nx, ny, nz))
    args = (0.7, 3.0, 0.3, 0.8, 1.1)
    want = dense_step(a, chi, *args, implicit_chi=m > 0)
    got = env.advance(a, chi, *args, m=m)
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()
    want2 = dense_step(want, chi, *args, implicit_chi=m > 0)
    got2 = env.advance(a, chi, *args, steps=2, m=m)
    assert np.abs(got2 - want2).max() <= 1e-12 * np.abs(want2).max()


# Gaussian beam: a0 exp(-4 y^2/ys^2) has field radius w0 = ys/2.
K0, W0, L, N, NX, DXI = 20.0, 4.0, 48.0, 128, 40, 0.25
ZR = K0 * W0 ** 2 / 2


def beam(nx=NX):
    d = L / N
    a = env.gaussian(nx, N, N, DXI, d, d, 1.0, 6.0, 2 * W0, 2 * W0, nx * DXI / 2, L / 2, L / 2)
    return a, d


def spot(plane, d):
    """1/e field radius from the second moment of |a|^2 about its centroid."""
    p = np.abs(plane) ** 2
    y = np.arange(plane.shape[0]) * d
    z = np.arange(plane.shape[1]) * d
    s = p.sum()
    yc = (p.sum(1) * y).sum() / s
    zc = (p.sum(0) * z).sum() / s
    r2 = (y[:, None] - yc) ** 2 + (z[None, :] - zc) ** 2
    return np.sqrt(2 * (p * r2).sum() / s)


@pytest.mark.parametrize("frac", [0.5, 1.0])
def test_vacuum_gaussian_beam(frac):
    a, d = beam()
    c = NX // 2
    n_steps = int(40 * frac)
    b = env.advance(a, np.zeros(a.shape), ZR / 40, K0, DXI, d, d, steps=n_steps)
    tau = ZR / 40 * n_steps
    w = W0 * np.sqrt(1 + (tau / ZR) ** 2)
    mid = N // 2
    assert abs(spot(b[c], d) / spot(a[c], d) - w / W0) < 0.005
    assert abs(abs(b[c, mid, mid]) / abs(a[c, mid, mid]) - W0 / w) < 0.005
    assert abs(np.angle(b[c, mid, mid]) + np.arctan(tau / ZR)) < 0.01


def test_matched_channel_holds_the_spot():
    """chi = 4 r^2/w0^4 (capped at r = 3 w0, where |a|^2 is e^-18 of its peak)."""
    a, d = beam()
    c = NX // 2
    y = np.arange(N) * d - L / 2
    r2 = y[:, None] ** 2 + y[None, :] ** 2
    chi = np.broadcast_to(np.minimum(4 * r2, 36 * W0 ** 2) / W0 ** 4, a.shape).copy()
    b = env.advance(a, chi, ZR / 40, K0, DXI, d, d, steps=40, m=2)
    assert abs(spot(b[c], d) / spot(a[c], d) - 1) < 0.005
    assert abs(np.abs(b[c]).max() / np.abs(a[c]).max() - 1) < 0.005
    # the chi corrections have converged to the Crank-Nicolson solution
    b30 = env.advance(a, chi, ZR / 40, K0, DXI, d, d, steps=40, m=6)
    assert np.abs(b - b30).max() < 1e-3 * np.abs(b30).max()
    # a channel twice as deep focuses the beam: the spot shrinks
    b2 = env.advance(a, 2 * chi, ZR / 40, K0, DXI, d, d, steps=20, m=2)
    assert spot(b2[c], d) < 0.97 * spot(a[c], d)


def test_plateau_in_uniform_chi_turns_its_phase():
    """Uniform chi equals its midrange: one transform per slice (m = 0) is the
    Crank-Nicolson step."""
    nx, dxi, k0, chi0 = 120, 0.2, 10.0, 0.5
    x = np.arange(nx) * dxi
    front = x[-1] - 4.0
    prof = 0.5 * (1 - np.tanh((x - front) / 1.0))   # plateau behind a smooth front
    a = np.broadcast_to(prof[:, None, None], (nx, 4, 4)).astype(complex)
    chi = np.full(a.shape, chi0)
    dt, n = 0.5, 40
    b = env.advance(a, chi, dt, k0, dxi, 1.0, 1.0, steps=n, m=0)
    plateau = slice(0, nx // 2)
    ratio = b[plateau, 0, 0] / a[plateau, 0, 0]
    cn = -2 * n * np.arctan(dt * chi0 / (4 * k0))
    assert abs(cn + chi0 * dt * n / (2 * k0)) < 1e-5
    assert np.abs(np.angle(ratio) - cn).max() < 1e-8
    assert np.abs(np.abs(ratio) - 1).max() < 1e-7
