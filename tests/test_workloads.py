"""Synthetic inputs (CPU only): the counter-based SplitMix64 stream of the
workloads equals the oracle's (same key layout), the thermal momenta have the
stated N(0, sigma^2) statistics, and the C4 draws are reproducible and cover
every grid class."""
import numpy as np

from quetzalpic_amd import workloads


def test_splitmix64_matches_oracle(oracle):
    lib = oracle.load()
    rng = np.random.default_rng(0)
    keys = [(1, 2, 3), (0, 0, 0), (2024, 63, 3), (2 ** 64 - 1, 2 ** 40, 2 ** 33 + 7)]
    keys += [tuple(int(v) for v in rng.integers(0, 2 ** 63, 3)) for _ in range(20)]
    for s, t, i in keys:
        assert int(workloads.splitmix64(s, t, i)) == lib.qpo_splitmix64(s, t, i)
    # vectorised over the index equals the scalar calls
    idx = np.arange(1000, 1100)
    v = workloads.splitmix64(7, 3, idx)
    assert all(int(v[k]) == lib.qpo_splitmix64(7, 3, int(idx[k])) for k in range(len(idx)))


def test_thermal_momenta_statistics():
    wl = workloads.get("C3")
    _, _, _, py, pz = wl.particles()
    n = len(py)
    assert n == 1048576
    for p in (py, pz):
        assert abs(p.mean()) < 5 * 1e-3 / np.sqrt(n)
        assert abs(p.std() / 1e-3 - 1) < 5 / np.sqrt(2 * n)
    assert abs(np.corrcoef(py, pz)[0, 1]) < 5 / np.sqrt(n)
    # reproducible (counter-based, no generator state)
    assert np.array_equal(wl.particles()[3], py)


def test_c4_members_reproducible_and_cover_classes():
    a = workloads.c4_members()
    b = workloads.c4_members()
    assert [(m.ny, m.a0, m.w0, m.profile) for m in a] == [(m.ny, m.a0, m.w0, m.profile) for m in b]
    assert {m.ny for m in a} == set(workloads.C4_SIZES)
    assert all(0.5 <= m.a0 <= 3 and 2 <= m.w0 <= 5 and 0 <= m.profile[1] <= 0.1 for m in a)
