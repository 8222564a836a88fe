"""Whole-loop pins of the oracle's slice loop (CPU only).

* no driver -> no wake (all fields vanish to rounding);
* transverse symmetries of a centred driver;
* the linear-wake known answer of notebooks/Test3D.ipynb (cell 1), mirrored to
  today's slice direction (the loop runs from x = lx down to 0,
  src/System_3d.cpp:328): psi(x) = int_x^lx a^2(x')/2 sin(x' - x) dx',
  e_x(x) = int_x^lx a^2(x')/2 cos(x' - x) dx' on the axis;
* the soft stop rule fires on a bunch source that blows B up;
* committed golden fixtures reproduce (regression pin of the oracle itself).
"""
import os

import numpy as np
import pytest

from conftest import rel_err

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def small(oracle, n=32, L=20.0, lx=3.0, dx=0.1, K=5, a0=1.0):
    g = oracle.grid_from_box(lx, L, L, dx, L / n, L / n, K=K)
    y, z, w = oracle.init_lattice(g, 2, 2)
    a = oracle.gaussian_a_sqr(g, a0, 4.0, 6.0, 6.0, lx - 1.5, L / 2, L / 2)
    return g, y, z, w, a


def test_no_driver_no_wake(oracle):
    """Cold uniform plasma, no driver: gamma = 1, v_x = 0, so rho = n = -1 and the
    susceptibility -n/((1-v_x) gamma) = +1 on every node; every field vanishes."""
    g, y, z, w, a = small(oracle)
    out, st, _ = oracle.wake_run(g, y, z, w, np.zeros_like(a))
    assert st == 0
    for f, v in out.items():
        if f == "rho":
            assert np.abs(v + 1.0).max() < 1e-12
        elif f == "chi":
            assert np.abs(v - 1.0).max() < 1e-12
        else:
            assert np.abs(v).max() < 1e-12, f


def test_symmetries_of_centred_driver(oracle):
    """The driver sits on node 16 of 32: reflection k -> 32 - k is exact in z.  In y it
    holds exactly without the B iteration (K = 0); with K > 0 it is broken only by the
    reference's B_z source, whose last row reads bz(i, 0, k) (src/System_3d.cpp:563)."""
    g, y, z, w, a = small(oracle, K=0)
    out, _, _ = oracle.wake_run(g, y, z, w, a, fields=("psi", "ex", "rho"))
    for f in ("psi", "ex", "rho"):
        v = out[f]
        zr = np.roll(v[:, :, ::-1], 1, axis=2)
        yr = np.roll(v[:, ::-1, :], 1, axis=1)
        assert np.abs(v - zr).max() < 1e-12 * max(1.0, np.abs(v).max()), f
        assert np.abs(v - yr).max() < 1e-12 * max(1.0, np.abs(v).max()), f
        assert np.abs(v - np.swapaxes(v, 1, 2)).max() < 1e-12 * max(1.0, np.abs(v).max()), f
    g, y, z, w, a = small(oracle, K=5)
    out, _, _ = oracle.wake_run(g, y, z, w, a, fields=("psi", "by", "bz"))
    psi = out["psi"]
    assert np.abs(psi - np.roll(psi[:, :, ::-1], 1, axis=2)).max() < 1e-12
    yasym = np.abs(psi - np.roll(psi[:, ::-1, :], 1, axis=1)).max()
    assert 0 < yasym < 1e-3 * np.abs(psi).max()


def linear_wake_case(oracle, a0=0.05):
    L, n, lx, dx = 30.0, 96, 16.0, 0.05
    g = oracle.grid_from_box(lx, L, L, dx, L / n, L / n, K=5)
    y, z, w = oracle.init_lattice(g, 2, 2)
    xs, rs, x0 = 4.0, 10.0, lx - 4.0   # full 1/e^2 widths: w_x = 4/sqrt(8), w_r = 10/sqrt(8)
    a = oracle.gaussian_a_sqr(g, a0, xs, rs, rs, x0, L / 2, L / 2)
    return g, y, z, w, a, (a0, xs, x0)


def test_linear_wake_known_answer(oracle):
    g, y, z, w, a, (a0, xs, x0) = linear_wake_case(oracle)
    out, st, _ = oracle.wake_run(g, y, z, w, a, fields=("psi", "ex"))
    assert st == 0
    c = g.ny // 2
    x = np.arange(g.nx) * g.dx
    half_a2 = 0.5 * a[:, c, c]
    # trapezoid integrals on a fine grid of the analytic driver
    wx = xs / np.sqrt(8.0)
    xf = np.linspace(0, g.nx * g.dx, 20001)
    af = 0.5 * (0.5 * a0 * a0) * np.exp(-(xf - x0) ** 2 / wx ** 2)
    psi_t = np.array([np.trapezoid(af[xf >= xi] * np.sin(xf[xf >= xi] - xi), xf[xf >= xi]) for xi in x])
    # e_x(i) = (psi_i - psi_{i+1})/dx is centred at x_i + dx/2 (the reference's
    # origin attribute says -dx/2, src/System_3d.cpp:182; the difference quotient
    # itself sits half a step ahead)
    xe = x + 0.5 * g.dx
    ex_t = np.array([np.trapezoid(af[xf >= xi] * np.cos(xf[xf >= xi] - xi), xf[xf >= xi]) for xi in xe])
    psi_n = out["psi"][:, c, c]
    ex_n = out["ex"][:, c, c]
    sel = slice(0, g.nx - 1)
    amp = np.abs(psi_t).max()
    # O(dx, dy^2) discretisation + finite driver width: measured 2.0 % (psi), 2.4 % (e_x)
    assert np.abs(psi_n[sel] - psi_t[sel]).max() < 0.04 * amp
    assert np.abs(ex_n[sel] - ex_t[sel]).max() < 0.04 * np.abs(ex_t).max()
    # the wake oscillates at k_p = 1 behind the driver: psi and e_x in quadrature
    behind = x < x0 - 3.0
    corr = np.corrcoef(psi_n[behind], psi_t[behind])[0, 1]
    assert corr > 0.995
    assert np.abs(half_a2).max() > 0


def test_linear_wake_scales_as_a0_squared(oracle):
    g, y, z, w, a1, _ = linear_wake_case(oracle, a0=0.02)
    o1, _, _ = oracle.wake_run(g, y, z, w, a1, fields=("psi",))
    g, y, z, w, a2, _ = linear_wake_case(oracle, a0=0.04)
    o2, _, _ = oracle.wake_run(g, y, z, w, a2, fields=("psi",))
    ratio = np.abs(o2["psi"]).max() / np.abs(o1["psi"]).max()
    assert abs(ratio - 4.0) < 0.01


def test_stop_rule_fires(oracle):
    """A huge bunch source drives |B| above THRESHOLD_B = 100 (src/System_3d.cpp:326, 541-591)."""
    g, y, z, w, a = small(oracle, lx=2.0)
    rho_b = np.zeros((g.nx, g.ny, g.nz))
    jx_b = np.zeros_like(rho_b)
    i0 = g.nx - 5
    jx_b[i0, 10:14, 10:14] = 5e4
    rho_b[i0, 10:14, 10:14] = 5e4
    out, st, ss = oracle.wake_run(g, y, z, w, a, rho_b=rho_b, jx_b=jx_b, fields=("psi", "by", "bz", "ex"))
    assert st == 1 and ss == i0
    assert np.abs(out["bz"][i0]).max() > 100 or np.abs(out["by"][i0]).max() > 100
    # slices below the stop are never computed except the output_step(0) overwrite
    assert np.abs(out["by"][1:i0]).max() == 0


@pytest.mark.parametrize("name", ["small_k5", "small_k1", "channel_thermal"])
def test_golden_fixtures_reproduce(oracle, name):
    path = os.path.join(GOLDEN, f"{name}.npz")
    if not os.path.exists(path):
        pytest.skip("fixture not generated")
    d = np.load(path, allow_pickle=False)
    from golden.make_golden import case_inputs
    g, y, z, w, py, pz, a = case_inputs(oracle, name)
    out, st, _ = oracle.wake_run(g, y, z, w, a, py=py, pz=pz)
    for f in out:
        assert rel_err(out[f], d[f])[0] < 1e-13, f
