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
