"""CPU checks of the C++14 host layer (TOML reader, Config_reader on the
reference's input decks, Species_register, HDF5 layout) via its self-test."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "quetzalpic_amd", "bin")


def test_host_selftest(tmp_path):
    exe = os.path.join(BIN, "qp_host_selftest")
    assert os.path.exists(exe), "host layer not built (run __graft_entry__.build())"
    r = subprocess.run([exe, os.path.join(ROOT, "tests", "golden", "decks"), str(tmp_path)],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120)
    fails = [l for l in r.stdout.splitlines() if l.startswith("FAIL")]
    assert r.returncode == 0 and not fails, r.stdout
    assert sum(1 for l in r.stdout.splitlines() if l.startswith("PASS")) >= 40


def test_cli_needs_an_input():
    exe = os.path.join(BIN, "quasistatic_pic")
    assert os.path.exists(exe)
    r = subprocess.run([exe], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=60)
    assert "Provide a path to the input .toml file" in r.stdout


def test_c4_members_are_seeded_and_cover_every_grid_size():
    """The C4 workload: 64 members, grids drawn from {64, 96, 128, 192, 256},
    the same members on every call (seeded per instance), ramp densities >= 0."""
    from quetzalpic_amd import workloads
    a = workloads.c4_members()
    b = workloads.c4_members()
    assert len(a) == 64
    assert {m.ny for m in a} == set(workloads.C4_SIZES)
    for x, y in zip(a, b):
        assert (x.ny, x.a0, x.w0, x.profile) == (y.ny, y.a0, y.w0, y.profile)
        assert 0.5 <= x.a0 <= 3.0 and 2.0 <= x.w0 <= 5.0 and 0.0 <= x.profile[1] <= 0.1
        assert x.nx == 800 and abs(x.dx - 0.025) < 1e-12
    yy, zz, w, _, _ = a[3].particles()
    assert (w <= 0).all() and len(w) == a[3].n_particles
