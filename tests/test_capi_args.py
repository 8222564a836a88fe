"""Argument checks of the ctypes wrappers (CPU only: they run before any C call).
The C side reads exactly n1*n2, len(y) or nx*ny*nz values through each pointer,
so a mis-shaped array must be refused in Python."""
import numpy as np
import pytest


def test_f64_accepts_and_converts():
    from quetzalpic_amd import capi
    a = capi._f64([[1, 2], [3, 4]], (2, 2), "x")
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    assert capi._f64(None, (2, 2), "x") is None
    c = capi._f64(np.zeros((3, 2)), (3, 2), "c", np.complex128)
    assert c.dtype == np.complex128


@pytest.mark.parametrize("shape", [(2, 3), (3, 2, 1), (6,)])
def test_f64_rejects_wrong_shape(shape):
    from quetzalpic_amd import capi
    with pytest.raises(ValueError, match="expected"):
        capi._f64(np.zeros(shape), (3, 2), "grid")


def test_particles_lengths():
    from quetzalpic_amd import capi
    y = np.zeros(5)
    out = capi._particles(y, np.ones(5), None, None, np.ones(5), "p")
    assert out[2] is None and len(out[4]) == 5
    with pytest.raises(ValueError, match="weight"):
        capi._particles(y, np.ones(5), None, None, np.ones(4), "p")
    with pytest.raises(ValueError, match="pz"):
        capi._particles(y, np.ones(5), np.ones(5), np.ones(6), np.ones(5), "p")
    with pytest.raises(ValueError, match="one-dimensional"):
        capi._particles(np.zeros((2, 2)), np.ones(4), None, None, np.ones(4), "p")
