"""Write the golden fixtures of the slice loop with the FP64 oracle.

The reference ships no expected outputs (its tests are input decks only), so
these fixtures are produced by the oracle after it passed the closed-form and
known-answer pins in tests/test_oracle_*.py.  They pin the oracle against
regressions and give the GPU parity tests committed vectors.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

CASES = ("small_k5", "small_k1", "channel_thermal")


def case_inputs(oc, name):
    """Grid, particles and <a^2> of a fixture (small enough to commit)."""
    py = pz = None
    if name in ("small_k5", "small_k1"):
        K = 5 if name == "small_k5" else 1
        g = oc.grid_from_box(2.0, 20.0, 20.0, 0.1, 20.0 / 24, 20.0 / 24, K=K)
        y, z, w = oc.init_lattice(g, 2, 2)
        a = oc.gaussian_a_sqr(g, 1.5, 3.0, 6.0, 6.0, 1.2, 10.0, 10.0)
    elif name == "channel_thermal":
        g = oc.grid_from_box(1.6, 16.0, 12.0, 0.08, 16.0 / 20, 12.0 / 16, K=5)
        prof = oc.QpoProfile(kind=oc.PROFILE_CHANNEL, y0=8.0, z0=6.0, base=1.0, curv=0.05)
        y, z, w = oc.init_lattice(g, 2, 2, prof)
        rng = np.random.default_rng(2024)
        py = rng.normal(0, 1e-3, len(y))
        pz = rng.normal(0, 1e-3, len(y))
        a = oc.gaussian_a_sqr(g, 2.0, 2.5, 5.0, 5.0, 1.0, 8.0, 6.0)
    else:
        raise KeyError(name)
    return g, y, z, w, py, pz, a


def main():
    from oracle import oracle_ctypes as oc
    for name in CASES:
        g, y, z, w, py, pz, a = case_inputs(oc, name)
        out, st, ss = oc.wake_run(g, y, z, w, a, py=py, pz=pz)
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), status=st, stop_slice=ss,
                            inputs_y=y, inputs_z=z, inputs_w=w, inputs_a_sqr=a,
                            inputs_py=py if py is not None else np.zeros(0),
                            inputs_pz=pz if pz is not None else np.zeros(0),
                            grid=np.array([g.nx, g.ny, g.nz, g.dx, g.dy, g.dz, g.ly, g.lz, g.K]), **out)
        print(name, {k: float(np.abs(v).max()) for k, v in out.items()})


if __name__ == "__main__":
    main()
