"""Why the transverse solves stay FP64: a measured lower bound.

The oracle's precision-study switch rounds every solve's input, spectrum
(before and after the eigenvalue division) and output to single precision.
Any transform that stores or accumulates in FP32 -- an FP32 FFT, or a dense
DFT as split-BF16 products accumulated in FP32 on the matrix cores -- rounds
at least that much.  The whole slice loop is then compared with the FP64
oracle on the same inputs.

psi itself stays at FP32 rounding (~1e-7), but e_x = (psi_i - psi_(i+1))/dx
(src/System_3d.cpp:633-640) divides a slice-to-slice difference by dx, so the
rounding of psi is amplified by ~|psi| / (dx |e_x|).  At C1's dx = 0.05 that
is already ~2.5e-6; the bar is 1e-5 and C2 (dx = 0.02) crosses it within its
first 300 slices (QP_SLOW=1 runs that case: 1.02e-5 max, 4.8e-6 RMS).  C3
(dx = 0.01) and C5 (dx = 0.005) amplify 2x and 4x more.
"""
import os

import numpy as np
import pytest


def _run(oracle, name, S):
    from quetzalpic_amd import workloads
    wl = workloads.get(name)
    lib = oracle.load("omp")
    g = oracle.grid_from_box(wl.lx, wl.L, wl.L, wl.dx0, wl.L / wl.n0, wl.L / wl.n0, K=wl.K)
    y, z, w, py, pz = wl.particles()
    a = oracle.gaussian_a_sqr_window(g, S, *wl.laser)
    res = {}
    try:
        for f32 in (0, 1):
            lib.qpo_set_solve_f32(f32)
            out, st, _ = oracle.wake_run_window(g, y, z, w, a, S, py=py, pz=pz, fields=("psi", "ex"),
                                                variant="omp")
            assert st == 0
            res[f32] = out
    finally:
        lib.qpo_set_solve_f32(0)
    err = {}
    for f in ("psi", "ex"):
        d = res[1][f] - res[0][f]
        err[f] = (np.abs(d).max() / np.abs(res[0][f]).max(), np.linalg.norm(d) / np.linalg.norm(res[0][f]))
    return err


def test_fp32_transforms_amplified_in_ex(oracle):
    err = _run(oracle, "C1", 200)
    print("C1 FP32-rounded solves:", err)
    assert 1e-8 < err["psi"][0] < 3e-7          # psi: single-precision rounding
    assert err["ex"][0] > 1e-6                   # e_x: amplified by ~1/dx
    assert err["ex"][0] > 20 * err["psi"][0]


@pytest.mark.skipif(os.environ.get("QP_SLOW") != "1", reason="70 s of CPU: set QP_SLOW=1")
def test_fp32_transforms_miss_the_bar_at_c2(oracle):
    err = _run(oracle, "C2", 300)
    print("C2 (300 slices) FP32-rounded solves:", err)
    assert err["ex"][0] > 1e-5
