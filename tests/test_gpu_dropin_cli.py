"""End-to-end drop-in check on the GPU: the C++14 CLI (Config_reader ->
System_3d -> C-ABI slice loop -> Output_writer) on a TOML deck, its HDF5
output read back and compared with the FP64 oracle on the same configuration."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, rel_err

pytestmark = pytest.mark.gpu

DECKS = os.path.join(ROOT, "tests", "golden", "decks_small")
EXE = os.path.join(ROOT, "quetzalpic_amd", "bin", "quasistatic_pic")


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-10), ("mixed", 1e-5)])
def test_cli_matches_oracle(oracle, tmp_path, prec, tol):
    from quetzalpic_amd import h5read
    deck = os.path.join(DECKS, f"laser_small_{prec}.toml")
    r = subprocess.run([EXE, deck], cwd=str(tmp_path), stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout[-3000:]
    assert "Finished in" in r.stdout
    path = str(tmp_path / "Fields_000.h5")
    g = oracle.grid_from_box(3.0, 20.0, 20.0, 0.1, 0.625, 0.625, K=5)
    y, z, w = oracle.init_lattice(g, 2, 2)
    a = oracle.gaussian_a_sqr(g, 1.0, 4.0, 6.0, 6.0, 1.5, 10.0, 10.0)
    want, st, _ = oracle.wake_run(g, y, z, w, a)
    for f in ("psi", "ex", "ey", "ez", "by", "bz", "rho", "jx", "jy", "jz"):
        got = h5read.read_field(path, f)
        assert got.shape == (g.nx, g.ny, g.nz)
        mx, rms = rel_err(got, want[f])
        lim = tol if f in ("psi", "ex") else max(tol, 1e-4)
        assert mx < lim, (f, mx, rms)
    chi = h5read.read_field(path, "susceptibility")
    assert rel_err(chi, want["chi"])[0] < max(tol, 1e-4)
    asq = h5read.read_field(path, "aSqr")
    assert rel_err(asq, a)[0] < 1e-13
    # attributes of the reference layout
    assert np.allclose(h5read.read_attr(path, "ex", "origin", 3), [-0.05, 0.0, 0.0])
    assert np.allclose(h5read.read_attr(path, "jy", "origin", 3), [-0.05, 0.3125, 0.0])
    xy = h5read.read_field(str(tmp_path / "Fields_xy_000.h5"), "psi")
    assert xy.shape == (g.nx, g.ny)
    assert os.path.exists(tmp_path / "Photons_000.h5")  # output_photons defaults to true
