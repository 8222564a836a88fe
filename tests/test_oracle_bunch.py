"""Pins of the oracle's driver-bunch functions (CPU): the 3D CIC deposit and
trilinear gather (src/array_utils.cpp:290-374), the bunch fields and QED
parameter (src/System_3d.cpp:843-860) and the explicit push
(src/System_3d.cpp:862-890).

Closed forms: charge conservation and the x-cut of the deposit, adjointness of
gather and deposit, exact reproduction of a trilinear field on node and
staggered arrays (src/System_3d.cpp:179-183) and through the bunch gather, node and
periodic-seam cases, free streaming and constant-force momentum gain, the
photon branch, and chi for pure transverse E and pure B.  Grids have
ny != nz and dy != dz so that a transposed axis fails.
"""
import numpy as np
import pytest

G_ARGS = (4.0, 6.0, 5.0, 0.25, 0.5, 0.3125)   # lx, ly, lz, dx, dy, dz -> 16 x 12 x 16


@pytest.fixture
def g(oracle):
    return oracle.grid_from_box(*G_ARGS)


def test_charge_conservation_and_x_cut(oracle, g):
    rng = np.random.default_rng(1)
    n = 300
    x = rng.uniform(0, (g.nx - 1) * g.dx, n)
    y = rng.uniform(-g.dy, g.ly, n)          # includes y in [-dy, 0): j1 = -1 wraps
    z = rng.uniform(0, g.lz, n)
    q = rng.normal(size=n)
    a = np.zeros((g.nx, g.ny, g.nz))
    for p in range(n):
        oracle.deposit3d(g, x[p], y[p], z[p], q[p], a)
    assert abs(a.sum() - q.sum()) <= 1e-12 * np.abs(q).sum()
    # x-cells outside [0, nx-1) deposit nothing (src/array_utils.cpp:299-301)
    b = np.zeros_like(a)
    for xx in (-0.01, (g.nx - 1) * g.dx + 1e-9, g.nx * g.dx + 0.3):
        oracle.deposit3d(g, xx, 1.0, 1.0, 1.0, b)
    assert not b.any()


def test_gather_is_the_adjoint_of_deposit(oracle, g):
    rng = np.random.default_rng(2)
    f = rng.normal(size=(g.nx, g.ny, g.nz))
    for _ in range(40):
        x = rng.uniform(0, (g.nx - 1) * g.dx)
        y, z, v = rng.uniform(0, g.ly), rng.uniform(0, g.lz), rng.normal()
        a = np.zeros_like(f)
        oracle.deposit3d(g, x, y, z, v, a)
        assert abs((f * a).sum() - v * oracle.gather3d(g, x, y, z, f)) <= 1e-12 * abs(v) * np.abs(f).max()


@pytest.mark.parametrize("origin", [(0.0, 0.0, 0.0), "ex", "ey", "ez"])
def test_gather_reproduces_trilinear_fields(oracle, g, origin):
    """Node (i, j, k) of an array with origin o sits at o + (i dx, j dy, k dz)
    (the staggered field arrays, src/System_3d.cpp:179-183)."""
    o = oracle.field_origins(g)[origin] if isinstance(origin, str) else origin
    i, j, k = np.meshgrid(np.arange(g.nx), np.arange(g.ny), np.arange(g.nz), indexing="ij")
    X, Y, Z = o[0] + i * g.dx, o[1] + j * g.dy, o[2] + k * g.dz

    def fun(x, y, z):
        return 0.3 + 1.7 * x - 0.6 * y + 2.2 * z + 0.5 * x * y - 0.8 * y * z + 0.25 * x * z + 0.1 * x * y * z

    f = fun(X, Y, Z)
    rng = np.random.default_rng(3)
    for _ in range(40):
        x = o[0] + rng.uniform(0, (g.nx - 1) * g.dx)
        y = o[1] + rng.uniform(0, (g.ny - 1) * g.dy)   # away from the periodic seam
        z = o[2] + rng.uniform(0, (g.nz - 1) * g.dz)
        assert abs(oracle.gather3d(g, x, y, z, f, o) - fun(x, y, z)) <= 1e-12 * np.abs(f).max()
    assert oracle.gather3d(g, o[0] - 0.01, 1.0, 1.0, f, o) == 0.0
    assert oracle.gather3d(g, o[0] + (g.nx - 1) * g.dx + 0.01, 1.0, 1.0, f, o) == 0.0


def test_bunch_fields_use_the_staggered_origins(oracle, g):
    """Each field of the bunch gather is read with its own array origin: a field
    linear in the staggered coordinate comes back exactly at the particle."""
    rng = np.random.default_rng(9)
    b = bunch(40, rng)
    b["x"] = rng.uniform(0.2, (g.nx - 2) * g.dx, 40)
    b["y"] = rng.uniform(0.5 * g.dy, (g.ny - 2) * g.dy, 40)
    b["z"] = rng.uniform(0.5 * g.dz, (g.nz - 2) * g.dz, 40)
    i, j, k = np.meshgrid(np.arange(g.nx), np.arange(g.ny), np.arange(g.nz), indexing="ij")
    f3 = {}
    for name, o in oracle.field_origins(g).items():
        f3[name] = 0.2 + 0.9 * (o[0] + i * g.dx) - 0.3 * (o[1] + j * g.dy) + 0.7 * (o[2] + k * g.dz)
    out = oracle.bunch_fields(g, b, f3, 1.0)
    want = 0.2 + 0.9 * b["x"] - 0.3 * b["y"] + 0.7 * b["z"]
    for name in f3:
        assert np.abs(out[name] - want).max() < 1e-12, name


def test_node_and_seam_cases(oracle, g):
    a = np.zeros((g.nx, g.ny, g.nz))
    oracle.deposit3d(g, 3 * g.dx, 5 * g.dy, 7 * g.dz, 2.0, a)
    assert a[3, 5, 7] == 2.0 and a.sum() == 2.0
    # y in the last cell: shares between j = ny-1 and j = 0; z likewise with k
    b = np.zeros_like(a)
    oracle.deposit3d(g, 2.5 * g.dx, (g.ny - 0.5) * g.dy, (g.nz - 0.25) * g.dz, 1.0, b)
    nz_idx = set(zip(*np.nonzero(b)))
    assert nz_idx == {(i, j, k) for i in (2, 3) for j in (g.ny - 1, 0) for k in (g.nz - 1, 0)}
    assert abs(b[2, g.ny - 1, g.nz - 1] - 0.5 * 0.5 * 0.25) < 1e-15


def bunch(n, rng, photon_frac=0.0):
    b = {"x": rng.uniform(0.5, 3.0, n), "y": rng.uniform(1, 5, n), "z": rng.uniform(1, 4, n),
         "px": rng.uniform(10, 100, n), "py": rng.normal(0, 0.5, n), "pz": rng.normal(0, 0.5, n)}
    b["gamma"] = np.sqrt(1 + b["px"] ** 2 + b["py"] ** 2 + b["pz"] ** 2)
    b["q"] = rng.uniform(-1, 0, n)
    b["cmr"] = np.full(n, -1.0)
    b["photon"] = (rng.uniform(size=n) < photon_frac).astype(np.int32)
    b["gamma"][b["photon"] == 1] = np.sqrt(b["px"] ** 2 + b["py"] ** 2 + b["pz"] ** 2)[b["photon"] == 1]
    return b


def const_fields(g, **vals):
    shape = (g.nx, g.ny, g.nz)
    return {k: np.full(shape, vals.get(k, 0.0)) for k in ("ex", "ey", "ez", "by", "bz")}


def test_push_closed_forms(oracle, g):
    rng = np.random.default_rng(4)
    b = bunch(50, rng, photon_frac=0.3)
    dt = 0.7
    # no fields: straight lines at v - (1, 0, 0); momenta unchanged
    f0 = {k: np.zeros(50) for k in ("ex", "ey", "ez", "by", "bz")}
    nb = oracle.bunch_push(b, f0, dt)
    assert np.allclose(nb["x"], b["x"] + dt * (b["px"] / b["gamma"] - 1), rtol=0, atol=1e-14)
    assert np.allclose(nb["y"], b["y"] + dt * b["py"] / b["gamma"], rtol=0, atol=1e-14)
    assert np.array_equal(nb["px"], b["px"])
    # uniform E_x on charges: p_x grows by dt * cmr * E_x per step; photons feel nothing
    ex = 0.37
    fx = {k: np.full(50, ex if k == "ex" else 0.0) for k in ("ex", "ey", "ez", "by", "bz")}
    s = b
    for _ in range(3):
        s = oracle.bunch_push(s, fx, dt)
    ch = b["photon"] == 0
    assert np.allclose(s["px"][ch], b["px"][ch] + 3 * dt * (-1.0) * ex, rtol=0, atol=1e-12)
    assert np.array_equal(s["px"][~ch], b["px"][~ch])
    assert np.allclose(s["gamma"][ch], np.sqrt(1 + s["px"] ** 2 + s["py"] ** 2 + s["pz"] ** 2)[ch], rtol=1e-15)
    assert np.allclose(s["gamma"][~ch], np.sqrt(s["px"] ** 2 + s["py"] ** 2 + s["pz"] ** 2)[~ch], rtol=1e-15)


def test_fields_and_chi_closed_forms(oracle, g):
    rng = np.random.default_rng(5)
    b = bunch(30, rng)
    b["py"][:] = 0
    b["pz"][:] = 0
    b["gamma"] = np.sqrt(1 + b["px"] ** 2)
    fs = 2.5
    out = oracle.bunch_fields(g, b, const_fields(g, ey=0.4), fs)
    assert np.allclose(out["ey"], 0.4, rtol=0, atol=1e-15) and not out["ex"].any()
    assert np.allclose(out["chi"], b["gamma"] * 0.4 / fs, rtol=1e-13)
    out = oracle.bunch_fields(g, b, const_fields(g, by=0.3), fs)
    assert np.allclose(out["chi"], b["px"] * 0.3 / fs, rtol=1e-13)
    # a particle outside the x range of the grid sees no field
    b2 = {k: np.array([v[0]]) for k, v in b.items()}
    b2["x"][:] = -1.0
    assert oracle.bunch_fields(g, b2, const_fields(g, ey=0.4), fs)["ey"][0] == 0.0


def test_bunch_deposit_sums(oracle, g):
    rng = np.random.default_rng(6)
    b = bunch(200, rng)
    rho, jx = oracle.bunch_deposit(g, b)
    assert abs(rho.sum() - b["q"].sum()) <= 1e-12 * np.abs(b["q"]).sum()
    assert abs(jx.sum() - (b["q"] * b["px"] / b["gamma"]).sum()) <= 1e-12 * np.abs(b["q"]).sum()
