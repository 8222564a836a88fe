"""CPU restatement of the laser-envelope advance (numpy, FP64).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the device
envelope (qp_envelope_*).  The product path never imports this module.

No counterpart in the reference: there <a^2> is fixed for the whole run
(src/System_3d.cpp:805-816 is its only writer) and chi is deposited
(src/System_3d.cpp:396,526) but never read.  The envelope a(x, y, z) of the
vector potential Re[a exp(i k0 (x - t))] (x toward the front, slice i at
x = i dx, arrays [i][j][k]) obeys
    lap_perp a + 2 d_tau (d_x + i k0) a = chi a.
One step tau -> tau + dt, slices swept from the front (i = nx - 1) down
(Crank-Nicolson in tau, d_x as the difference toward the front):
    (2/dx)(d_{i+1} - d_i) + 2 i k0 d_i = (dt/2)(chi_i - lap_perp)(a+_i + a_i),
    d_i = a+_i - a_i,  d_{nx} = 0,
with lap_perp the periodic 5-point Laplacian.  With c_i the midrange
(min + max)/2 of chi_i, (chi_i - c_i) d_i couples the transverse modes and is
iterated: the first solve takes it as 0, each of m further solves the previous
d_i.  Each solve is done per transverse mode in the form a+_i = M_i S_i + C_i a_i
(complex 2D FFTs), S_i = -(2/dx) d_{i+1} + dt (chi_i - c_i) a_i
+ (dt/2)(chi_i - c_i) d_i, M_i = 1/(alpha + (lam - c_i) dt/2),
C_i = M_i (alpha - (lam - c_i) dt/2), alpha = -2/dx + 2 i k0.

Pinned in tests/test_oracle_envelope.py against a dense real-space solve of the
same discrete equations (tiny grid), the vacuum Gaussian-beam closed form
(spot, amplitude, Gouy phase), a matched parabolic channel (constant spot)
and the phase rotation -chi tau / (2 k0) of a plateau in uniform chi.
"""
import numpy as np


def lam5(ny, nz, dy, dz):
    """Eigenvalues of the periodic 5-point Laplacian on the ny x nz FFT modes."""
    j = np.arange(ny)
    k = np.arange(nz)
    l1 = (2 * np.cos(2 * np.pi * j / ny) - 2) / dy ** 2
    l2 = (2 * np.cos(2 * np.pi * k / nz) - 2) / dz ** 2
    return l1[:, None] + l2[None, :]


def gaussian(nx, ny, nz, dx, dy, dz, a0, xsigma, ysigma, zsigma, x0, y0, z0):
    """a = a0 exp(-4 sum (x-x0)^2/sigma^2) on the nodes (x = i dx, y = j dy, z = k dz):
    |a|^2/2 is the Gaussian <a^2> of the wake's laser with the same arguments."""
    x = np.arange(nx) * dx - x0
    y = np.arange(ny) * dy - y0
    z = np.arange(nz) * dz - z0
    e = (4 * x[:, None, None] ** 2 / xsigma ** 2 + 4 * y[None, :, None] ** 2 / ysigma ** 2 +
         4 * z[None, None, :] ** 2 / zsigma ** 2)
    return (a0 * np.exp(-e)).astype(np.complex128)


def advance(a, chi, dt, k0, dx, dy, dz, steps=1, m=1):
    """steps steps of dt with m chi corrections per slice; a complex [nx, ny, nz],
    chi real [nx, ny, nz]; returns the new envelope."""
    a = np.array(a, dtype=np.complex128)
    nx, ny, nz = a.shape
    lam = lam5(ny, nz, dy, dz)
    alpha = -2.0 / dx + 2j * k0
    chi = np.asarray(chi, dtype=np.float64)
    cm = 0.5 * (chi.min(axis=(1, 2)) + chi.max(axis=(1, 2)))
    for _ in range(steps):
        new = np.empty_like(a)
        dprev = np.zeros((ny, nz), np.complex128)
        for i in range(nx - 1, -1, -1):
            h = 0.5 * dt * (lam - cm[i])
            M = 1.0 / (alpha + h)
            C = M * (alpha - h)
            dchi = chi[i] - cm[i]
            ca = np.fft.fft2(a[i])
            base = -(2.0 / dx) * dprev + dt * dchi * a[i]
            d = np.zeros((ny, nz), np.complex128)
            for it in range(m + 1):
                S = base + 0.5 * dt * dchi * d if it else base
                d = np.fft.ifft2(M * np.fft.fft2(S) + C * ca) - a[i]
            new[i] = a[i] + d
            dprev = d
        a = new
    return a


def a_sqr(a):
    """<a^2> = |a|^2 / 2."""
    return 0.5 * np.abs(a) ** 2
