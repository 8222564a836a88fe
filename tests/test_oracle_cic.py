"""Pins of the oracle's CIC deposit / gather kernels (src/array_utils.cpp:50-288).

Closed forms and invariants that a plausible indexing mistake breaks:
charge conservation, node and cell-centre special cases, periodic wrap,
adjointness of deposit and gather, exact reproduction of bilinear fields, and
exact derivatives of linear / quadratic fields by the staggered-derivative
gathers.
"""
import numpy as np
import pytest

N1, N2, D1, D2 = 8, 6, 0.5, 0.25


def dep(oracle, y, z, v, o1=0.0, o2=0.0, raw=False):
    a = np.zeros((N1, N2))
    lib = oracle.load()
    p = oracle._ptr(a)
    if raw:
        lib.qpo_deposit_raw(y, z, v, p, N1, N2, D1, D2)
    else:
        lib.qpo_deposit(y, z, v, p, N1, N2, D1, D2, o1, o2)
    return a


def gat(oracle, y, z, a, o1=0.0, o2=0.0):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return oracle.load().qpo_gather(y, z, oracle._ptr(a), N1, N2, D1, D2, o1, o2)


def rand_points(n, seed):
    rng = np.random.default_rng(seed)
    return rng.uniform(0, N1 * D1, n), rng.uniform(0, N2 * D2, n)


@pytest.mark.parametrize("origin", [(0.0, 0.0), (0.5 * D1, 0.0), (0.0, 0.5 * D2)])
def test_deposit_conserves_charge(oracle, origin):
    ys, zs = rand_points(50, 1)
    for y, z in zip(ys, zs):
        a = dep(oracle, y, z, 1.7, *origin)
        assert abs(a.sum() - 1.7) < 1e-14
        assert (a >= 0).all()
        assert np.count_nonzero(a) <= 4


def test_deposit_node_and_centre(oracle):
    a = dep(oracle, 3 * D1, 2 * D2, 1.0)
    assert a[3, 2] == 1.0 and a.sum() == 1.0
    a = dep(oracle, 3.5 * D1, 2.5 * D2, 1.0)
    assert np.allclose(a[3:5, 2:4], 0.25, atol=0, rtol=0) and a.sum() == 1.0


def test_deposit_periodic_wrap(oracle):
    # last cell spreads onto row/column 0
    a = dep(oracle, (N1 - 0.5) * D1, (N2 - 0.5) * D2, 1.0)
    assert np.isclose(a[N1 - 1, N2 - 1], 0.25) and np.isclose(a[0, 0], 0.25)
    assert np.isclose(a[0, N2 - 1], 0.25) and np.isclose(a[N1 - 1, 0], 0.25)
    # half-cell origin below the first node wraps j1 = -1 to n-1
    a = dep(oracle, 0.1 * D1, 1.0 * D2, 1.0, 0.5 * D1, 0.0)
    assert np.isclose(a[N1 - 1, 1], 0.4) and np.isclose(a[0, 1], 0.6) and np.isclose(a.sum(), 1.0)


def test_raw_deposit_equals_node_deposit_for_nonnegative_coordinates(oracle):
    ys, zs = rand_points(30, 2)
    for y, z in zip(ys, zs):
        assert np.array_equal(dep(oracle, y, z, 0.3, raw=True), dep(oracle, y, z, 0.3))


@pytest.mark.parametrize("origin", [(0.0, 0.0), (0.5 * D1, 0.0), (0.0, 0.5 * D2)])
def test_gather_is_adjoint_of_deposit(oracle, origin):
    rng = np.random.default_rng(5)
    f = rng.standard_normal((N1, N2))
    ys, zs = rand_points(40, 3)
    for y, z in zip(ys, zs):
        lhs = (f * dep(oracle, y, z, 1.0, *origin)).sum()
        rhs = gat(oracle, y, z, f, *origin)
        assert abs(lhs - rhs) < 1e-13 * np.abs(f).max()


def test_gather_reproduces_bilinear_field(oracle):
    # f = c0 + c1 y + c2 z + c3 y z on the nodes is reproduced exactly inside a cell
    c = (0.3, -1.2, 0.7, 2.0)
    jj, kk = np.meshgrid(np.arange(N1), np.arange(N2), indexing="ij")
    f = c[0] + c[1] * jj * D1 + c[2] * kk * D2 + c[3] * jj * D1 * kk * D2
    rng = np.random.default_rng(9)
    for _ in range(40):
        y = rng.uniform(0, (N1 - 1) * D1 - 1e-9)
        z = rng.uniform(0, (N2 - 1) * D2 - 1e-9)
        want = c[0] + c[1] * y + c[2] * z + c[3] * y * z
        assert abs(gat(oracle, y, z, f) - want) < 1e-12


def test_derivative_gathers_exact_for_linear_and_quadratic(oracle):
    lib = oracle.load()
    jj, kk = np.meshgrid(np.arange(N1), np.arange(N2), indexing="ij")
    y_n, z_n = jj * D1, kk * D2
    rng = np.random.default_rng(4)
    # linear: exact slope; quadratic in y: centred half-node differences are
    # exact at the half nodes and interpolate linearly, so d/dy y^2 = 2y exactly.
    for f, dfy, dfz in ((0.5 + 1.5 * y_n - 0.25 * z_n, lambda y, z: 1.5, lambda y, z: -0.25),
                        (y_n ** 2 + 3 * z_n, lambda y, z: 2 * y, lambda y, z: 3.0),
                        (y_n - z_n ** 2, lambda y, z: 1.0, lambda y, z: -2 * z)):
        f = np.ascontiguousarray(f)
        p = oracle._ptr(f)
        for _ in range(30):
            # stay away from the periodic seam, where the polynomial is not periodic
            y = rng.uniform(1.0 * D1, (N1 - 2) * D1)
            z = rng.uniform(1.0 * D2, (N2 - 2) * D2)
            gy = lib.qpo_gather_yder(y, z, p, N1, N2, D1, D2)
            gz = lib.qpo_gather_zder(y, z, p, N1, N2, D1, D2)
            assert abs(gy - dfy(y, z)) < 1e-12, (gy, dfy(y, z))
            assert abs(gz - dfz(y, z)) < 1e-12, (gz, dfz(y, z))


def test_derivative_gather_periodic(oracle):
    """A periodic sine: derivative gathers follow the centred difference across the seam."""
    lib = oracle.load()
    jj, kk = np.meshgrid(np.arange(N1), np.arange(N2), indexing="ij")
    f = np.ascontiguousarray(np.sin(2 * np.pi * jj / N1) + np.cos(2 * np.pi * kk / N2))
    p = oracle._ptr(f)
    y, z = 0.1 * D1, 0.2 * D2  # j1 = floor(y/d - 1/2) = -1 wraps to n1-1
    g = lib.qpo_gather_yder(y, z, p, N1, N2, D1, D2)
    t = y / D1 - 0.5 - (-1)
    d_lo = (f[0, 0] - f[N1 - 1, 0]) / D1
    d_hi = (f[1, 0] - f[0, 0]) / D1
    d_lo2 = (f[0, 1] - f[N1 - 1, 1]) / D1
    d_hi2 = (f[1, 1] - f[0, 1]) / D1
    zf = z / D2
    want = (d_lo * (1 - t) + d_hi * t) * (1 - zf) + (d_lo2 * (1 - t) + d_hi2 * t) * zf
    assert abs(g - want) < 1e-13


def test_normalize_and_clamp(oracle):
    lib = oracle.load()
    a = np.arange(12, dtype=np.float64).reshape(3, 4) - 3.0
    b = a.copy()
    lib.qpo_normalize_potential(oracle._ptr(b), 3, 4)
    border = (a[:, 0].sum() + a[0, 1:].sum()) / (3 + 4 - 1)
    assert np.allclose(b, a - border, rtol=0, atol=1e-15)
    lib.qpo_increase_minimum(oracle._ptr(b), 12, -0.9999)
    assert np.allclose(b, np.maximum(a - border, -0.9999), rtol=0, atol=1e-15)


def test_normalize_coordinates(oracle):
    import ctypes
    lib = oracle.load()
    for y, z, wy, wz in ((-0.5, 10.5, 9.5, 0.5), (20.0, -20.0, 0.0, 0.0), (3.25, 4.5, 3.25, 4.5)):
        yy = ctypes.c_double(y)
        zz = ctypes.c_double(z)
        lib.qpo_normalize_coordinates(ctypes.byref(yy), ctypes.byref(zz), 10.0, 10.0)
        assert abs(yy.value - wy) < 1e-14 and abs(zz.value - wz) < 1e-14
