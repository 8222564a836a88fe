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


def test_cli_device_bunch_equals_host_bunch(tmp_path):
    """run() with the driver bunch on the device (deposit and gather streamed
    through the slice loop, push on the device) against the host bunch of the
    same CLI, three beam steps: fields and bunch particles of the last step."""
    from quetzalpic_amd import h5read
    runs = {}
    for mode in ("device", "host"):
        d = tmp_path / mode
        d.mkdir()
        deck = os.path.join(DECKS, f"bunch_small_{mode}.toml")
        r = subprocess.run([EXE, deck], cwd=str(d), stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                           timeout=300)
        assert r.returncode == 0, r.stdout[-3000:]
        assert ("Driver bunch: on the device" in r.stdout) == (mode == "device")
        runs[mode] = d
    for f in ("psi", "ex", "ey", "ez", "by", "bz", "rho_bunch", "jx_bunch"):
        a = h5read.read_field(str(runs["device"] / "Fields_002.h5"), f)
        b = h5read.read_field(str(runs["host"] / "Fields_002.h5"), f)
        assert np.abs(b).max() > 0, f
        assert rel_err(a, b)[0] < 1e-9, (f, rel_err(a, b))
    for k in ("x", "y", "z", "px", "py", "pz", "gamma", "ex", "ey", "ez", "by", "bz", "chi", "weight"):
        a = h5read.read_field(str(runs["device"] / "Bunch_002.h5"), k)
        b = h5read.read_field(str(runs["host"] / "Bunch_002.h5"), k)
        assert a.shape == b.shape and a.size > 1000
        assert rel_err(a, b)[0] < 1e-9, (k, rel_err(a, b))
