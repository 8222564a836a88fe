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


def run_cli(tmp_path, deck_name):
    deck = os.path.join(DECKS, deck_name)
    r = subprocess.run([EXE, deck], cwd=str(tmp_path), stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stdout[-3000:]
    return r.stdout


def test_cli_evolving_laser(oracle, tmp_path):
    """[laser] evolve: the envelope starts from the deck's <a^2>, the wake stores
    chi into it, and one envelope step of dt separates the two beam steps."""
    from oracle import envelope as envo
    from quetzalpic_amd import h5read
    out = run_cli(tmp_path, "laser_evolve_fp64.toml")
    assert "Laser envelope: evolving" in out
    g = oracle.grid_from_box(3.0, 20.0, 20.0, 0.1, 0.625, 0.625, K=5)
    a0 = oracle.gaussian_a_sqr(g, 1.5, 4.0, 6.0, 6.0, 1.5, 10.0, 10.0)
    f0, f1 = str(tmp_path / "Fields_000.h5"), str(tmp_path / "Fields_001.h5")
    asq0 = h5read.read_field(f0, "aSqr")
    assert rel_err(asq0, a0)[0] < 1e-13
    chi0 = h5read.read_field(f0, "susceptibility")
    y, z, w = oracle.init_lattice(g, 2, 2)
    want0, _, _ = oracle.wake_run(g, y, z, w, a0, fields=("chi", "psi"))
    assert rel_err(chi0, want0["chi"])[0] < 1e-9
    a1 = envo.advance(np.sqrt(2 * asq0), chi0, 4.0, 10.0, g.dx, g.dy, g.dz, steps=1, m=1)
    asq1 = h5read.read_field(f1, "aSqr")
    assert rel_err(asq1, envo.a_sqr(a1))[0] < 1e-10
    assert rel_err(asq1, asq0)[0] > 1e-3          # the laser did change
    want1, _, _ = oracle.wake_run(g, y, z, w, asq1, fields=("psi", "ex"))
    for f in ("psi", "ex"):
        assert rel_err(h5read.read_field(f1, f), want1[f])[0] < 1e-9, f


def test_cli_ramp_profile(oracle, tmp_path):
    """[plasma] profile = "ramp": n = max(0, 1 + slope (y - ly/2)) on the lattice of
    init_wake_particles (src/System_3d.cpp:724-748)."""
    from quetzalpic_amd import h5read
    run_cli(tmp_path, "ramp_fp64.toml")
    g = oracle.grid_from_box(3.0, 20.0, 20.0, 0.1, 0.625, 0.625, K=5)
    ppc = 2
    yy = (np.arange(ppc * g.ny) + 0.5) * g.dy / ppc
    zz = (np.arange(ppc * g.nz) + 0.5) * g.dz / ppc
    y = np.repeat(yy, len(zz))
    z = np.tile(zz, len(yy))
    v = -np.maximum(0.0, 1.0 + 0.08 * (y - 10.0))
    keep = v != 0
    a = oracle.gaussian_a_sqr(g, 1.0, 4.0, 6.0, 6.0, 1.5, 10.0, 10.0)
    want, _, _ = oracle.wake_run(g, y[keep], z[keep], v[keep] / ppc / ppc, a, fields=("psi", "ex", "rho"))
    for f in ("psi", "ex", "rho"):
        assert rel_err(h5read.read_field(str(tmp_path / "Fields_000.h5"), f), want[f])[0] < 1e-9, f


def lattice(g, ppcy, ppcz, profile):
    """init_wake_particles (src/System_3d.cpp:724-748): i outer, zero cells skipped."""
    yy = (np.arange(ppcy * g.ny) + 0.5) * g.dy / ppcy
    zz = (np.arange(ppcz * g.nz) + 0.5) * g.dz / ppcz
    y = np.repeat(yy, len(zz))
    z = np.tile(zz, len(yy))
    v = profile(y, z)
    keep = v != 0
    return y[keep], z[keep], v[keep] / ppcy / ppcz


def test_cli_wake_batch(qp_ctx, oracle, tmp_path):
    """[[wake]] members run as one batch; each member's files equal that member
    run alone in its own mixed wake, and the first member equals the oracle."""
    from quetzalpic_amd import capi, h5read
    out = run_cli(tmp_path, "wake_batch_mixed.toml")
    assert "Batch of 3 wakes" in out
    g = oracle.grid_from_box(3.0, 20.0, 20.0, 0.1, 0.625, 0.625, K=5)
    members = {
        "base": ((1.0, 4.0, 6.0, 6.0, 1.5, 10.0, 10.0), 2, lambda y, z: -np.ones_like(y)),
        "channel_a2": ((2.0, 4.0, 6.0, 6.0, 1.5, 9.5, 10.0), 2,
                       lambda y, z: -(1.0 + 0.02 * ((y - 10.0) ** 2 + (z - 10.0) ** 2))),
        "ramp": ((1.0, 4.0, 6.0, 6.0, 1.5, 10.0, 10.0), 1, lambda y, z: -np.maximum(0.0, 1.0 + 0.08 * (y - 10.0))),
    }
    fields = ("psi", "rho", "jx", "jy", "jz", "ex", "ey", "ez", "by", "bz", "chi")
    for name, (laser, ppcy, prof) in members.items():
        path = str(tmp_path / f"{name}_Fields_000.h5")
        assert os.path.exists(tmp_path / f"{name}_Fields_xy_000.h5")
        y, z, w = lattice(g, ppcy, 2, prof)
        wk = capi.Wake(qp_ctx, g.nx, g.ny, g.nz, g.dx, g.dy, g.dz, g.ly, g.lz, K=5, precision=capi.QP_PREC_MIXED)
        wk.set_particles(y, z, w)
        wk.set_laser_gaussian(*laser)
        single, st, _ = wk.run(fields=fields)
        wk.close()
        assert st == 0
        for f in fields:
            got = h5read.read_field(path, "susceptibility" if f == "chi" else f)
            assert rel_err(got, single[f])[0] < 1e-9, (name, f, rel_err(got, single[f]))
        a = oracle.gaussian_a_sqr(g, *laser)
        assert rel_err(h5read.read_field(path, "aSqr"), a)[0] < 1e-13, name
        if name == "base":
            want, _, _ = oracle.wake_run(g, y, z, w, a, fields=("psi", "ex"))
            for f in ("psi", "ex"):
                assert rel_err(h5read.read_field(path, f), want[f])[0] < 1e-5, f
    assert not os.path.exists(tmp_path / "Fields_000.h5")
