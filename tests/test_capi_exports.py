"""The HIP C-ABI library loads and exports every entry point that
include/quetzal_hip.h declares (no compute calls: runs without a GPU)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = []
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names += re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(qp_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_declares_entry_points():
    names = declared_symbols()
    for must in ("qp_solve2d", "qp_wake_run", "qp_fourier2d_forward", "qp_wake_set_particles"):
        assert must in names
    assert len(names) >= 20


def test_library_exports_every_declared_symbol():
    from quetzalpic_amd import capi
    lib = ctypes.CDLL(capi.LIB_PATH)
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, missing


def test_no_cpu_fallback_when_library_missing(tmp_path, monkeypatch):
    from quetzalpic_amd import capi
    monkeypatch.setattr(capi, "LIB_PATH", str(tmp_path / "missing.so"))
    monkeypatch.setattr(capi, "_lib", None)
    try:
        capi.load()
    except capi.QpError:
        return
    raise AssertionError("load() must fail loudly without the HIP library")


def test_version_and_error_strings():
    from quetzalpic_amd import capi
    lib = capi.load()
    assert b"gfx950" in lib.qp_version()
    assert lib.qp_last_error() is not None


def test_workload_lattice_matches_oracle(oracle):
    """The bench's numpy lattice (init_wake_particles, src/System_3d.cpp:724-748) equals the oracle's."""
    import numpy as np
    from quetzalpic_amd import workloads
    wl = workloads.get("C1")
    g = oracle.grid_from_box(wl.lx, wl.L, wl.L, wl.dx0, wl.L / wl.n0, wl.L / wl.n0)
    y, z, w, _, _ = wl.particles()
    y2, z2, w2 = oracle.init_lattice(g, 2, 2)
    assert np.array_equal(y, y2) and np.array_equal(z, z2) and np.array_equal(w, w2)
    assert (wl.nx, wl.ny) == (g.nx, g.ny)
