"""The oracle's windowed entry points (CPU only): the first S slices of the
loop with 3D arrays holding only those slices equal the same slices of the
full-array loop exactly, and so do the windowed laser fill and bunch deposit.
They let the large configurations (1024^2 x 4000 slices: 33 GB per 3D FP64
array) be checked over their leading slices."""
import numpy as np


def test_window_loop_equals_full(oracle):
    L, n, lx, dx = 16.0, 24, 3.0, 0.1
    g = oracle.grid_from_box(lx, L, L, dx, L / n, L / n, K=3)
    y, z, w = oracle.init_lattice(g, 2, 2)
    laser = (1.5, 2.0, 5.0, 5.0, lx - 0.8, L / 2, L / 2)
    a = oracle.gaussian_a_sqr(g, *laser)
    rng = np.random.default_rng(5)
    rho_b = np.zeros((g.nx, g.ny, g.nz))
    rho_b[g.nx - 12:g.nx - 4, 10:14, 10:14] = rng.uniform(0.1, 0.5, (8, 4, 4))
    jx_b = 0.9 * rho_b
    S = 17
    full, st, _ = oracle.wake_run(g, y, z, w, a, rho_b=rho_b, jx_b=jx_b, n_slices=S)
    aw = oracle.gaussian_a_sqr_window(g, S, *laser)
    assert np.array_equal(aw, a[g.nx - S:])
    win, st2, _ = oracle.wake_run_window(g, y, z, w, aw, S, rho_b=rho_b[g.nx - S:], jx_b=jx_b[g.nx - S:])
    assert st == st2 == 0
    for f in full:
        assert np.array_equal(win[f], full[f][g.nx - S:]), f
        assert np.abs(win[f]).max() > 0 or f in ("by", "bz") and np.abs(full[f]).max() == 0


def test_window_bunch_deposit_equals_full(oracle):
    g = oracle.grid_from_box(2.0, 8.0, 8.0, 0.1, 0.5, 0.5, K=1)
    rng = np.random.default_rng(9)
    nb = 300
    b = {"x": rng.uniform(-0.2, 2.2, nb), "y": rng.uniform(0, 8, nb), "z": rng.uniform(0, 8, nb),
         "px": rng.uniform(10, 20, nb), "q": rng.uniform(-1, 1, nb)}
    b["gamma"] = np.sqrt(1 + b["px"] ** 2)
    rho, jx = oracle.bunch_deposit(g, b)
    for base, nheld in ((0, g.nx), (5, 7), (g.nx - 3, 3)):
        rw, jw = oracle.bunch_deposit_window(g, b, base, nheld)
        assert np.array_equal(rw, rho[base:base + nheld])
        assert np.array_equal(jw, jx[base:base + nheld])
