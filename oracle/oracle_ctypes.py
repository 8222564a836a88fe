"""ctypes bindings to the FP64 CPU oracle (oracle/qp_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by
bench.py's cpu_baseline leg, always as the checker.  The product path
(quetzalpic_amd) never imports this module.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_BUILD = os.path.join(_HERE, "_build")

_dp = ctypes.POINTER(ctypes.c_double)


class QpoGrid(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("dx", ctypes.c_double), ("dy", ctypes.c_double), ("dz", ctypes.c_double),
        ("ly", ctypes.c_double), ("lz", ctypes.c_double),
        ("K", ctypes.c_int), ("D", ctypes.c_double),
        ("psi_threshold", ctypes.c_double), ("b_threshold", ctypes.c_double),
    ]


class QpoFields(ctypes.Structure):
    _fields_ = [(n, _dp) for n in ("psi", "rho", "jx", "jy", "jz", "ex", "ey", "ez", "by", "bz", "chi")]


class QpoProfile(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int), ("rho0", ctypes.c_double),
        ("power", ctypes.c_double), ("radius", ctypes.c_double), ("min_factor", ctypes.c_double),
        ("max_factor", ctypes.c_double), ("y0", ctypes.c_double), ("z0", ctypes.c_double),
        ("base", ctypes.c_double), ("curv", ctypes.c_double),
    ]


FIELD_NAMES = ("psi", "rho", "jx", "jy", "jz", "ex", "ey", "ez", "by", "bz", "chi")
PROFILE_CONSTANT, PROFILE_POWERLAW, PROFILE_CHANNEL = 0, 1, 2


def _ptr(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


_libs = {}


def load(variant="det"):
    """variant 'det' = deterministic serial oracle, 'omp' = OpenMP -Ofast timing build,
    or a key registered with load_path."""
    if variant in _libs:
        return _libs[variant]
    name = {"det": "liboracle.so", "omp": "liboracle_omp.so"}[variant]
    path = os.path.join(_BUILD, name)
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} missing: run `make -C oracle` (or __graft_entry__.build())")
    return load_path(variant, path)


def load_path(variant, path):
    """Bind an oracle build at `path` under the key `variant` (the mutation tests
    load deliberately broken copies of the oracle this way)."""
    lib = ctypes.CDLL(path)
    lib.qpo_fft.argtypes = [ctypes.c_int, _dp, ctypes.c_int]
    lib.qpo_set_solve_f32.argtypes = [ctypes.c_int]
    lib.qpo_set_pstar_f32.argtypes = [ctypes.c_int]
    lib.qpo_rfft2.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp]
    lib.qpo_irfft2.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp]
    lib.qpo_solve_poisson.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                      ctypes.c_double, _dp]
    lib.qpo_deposit.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp, ctypes.c_int,
                                ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                ctypes.c_double]
    lib.qpo_deposit_raw.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp, ctypes.c_int,
                                    ctypes.c_int, ctypes.c_double, ctypes.c_double]
    lib.qpo_gather.argtypes = [ctypes.c_double, ctypes.c_double, _dp, ctypes.c_int, ctypes.c_int,
                               ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double]
    lib.qpo_gather.restype = ctypes.c_double
    for f in (lib.qpo_gather_yder, lib.qpo_gather_zder):
        f.argtypes = [ctypes.c_double, ctypes.c_double, _dp, ctypes.c_int, ctypes.c_int,
                      ctypes.c_double, ctypes.c_double]
        f.restype = ctypes.c_double
    lib.qpo_normalize_potential.argtypes = [_dp, ctypes.c_int, ctypes.c_int]
    lib.qpo_increase_minimum.argtypes = [_dp, ctypes.c_int, ctypes.c_double]
    lib.qpo_normalize_coordinates.argtypes = [_dp, _dp, ctypes.c_double, ctypes.c_double]
    lib.qpo_profile_eval.argtypes = [ctypes.POINTER(QpoProfile), ctypes.c_double, ctypes.c_double]
    lib.qpo_profile_eval.restype = ctypes.c_double
    lib.qpo_init_lattice.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                     ctypes.c_int, ctypes.c_int, ctypes.POINTER(QpoProfile), _dp, _dp, _dp]
    lib.qpo_init_lattice.restype = ctypes.c_size_t
    lib.qpo_fill_a_sqr_gaussian.argtypes = [ctypes.POINTER(QpoGrid)] + [ctypes.c_double] * 7 + [_dp]
    lib.qpo_fill_a_sqr_gaussian_window.argtypes = ([ctypes.POINTER(QpoGrid), ctypes.c_int, ctypes.c_int] +
                                                   [ctypes.c_double] * 7 + [_dp])
    run_args = [ctypes.POINTER(QpoGrid), ctypes.c_size_t, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp,
                ctypes.POINTER(QpoFields), ctypes.POINTER(ctypes.c_int)]
    lib.qpo_wake_run.argtypes = run_args
    lib.qpo_wake_run.restype = ctypes.c_int
    lib.qpo_wake_run_partial.argtypes = run_args + [ctypes.c_int]
    lib.qpo_wake_run_partial.restype = ctypes.c_int
    lib.qpo_wake_run_window.argtypes = run_args + [ctypes.c_int]
    lib.qpo_wake_run_window.restype = ctypes.c_int
    lib.qpo_bunch_deposit_window.argtypes = ([ctypes.POINTER(QpoGrid), ctypes.c_int, ctypes.c_int, ctypes.c_size_t] +
                                             [_dp] * 8)
    lib.qpo_deposit3d.argtypes = [ctypes.POINTER(QpoGrid)] + [ctypes.c_double] * 4 + [_dp]
    lib.qpo_gather3d.argtypes = [ctypes.POINTER(QpoGrid)] + [ctypes.c_double] * 6 + [_dp]
    lib.qpo_gather3d.restype = ctypes.c_double
    lib.qpo_bunch_deposit.argtypes = [ctypes.POINTER(QpoGrid), ctypes.c_size_t] + [_dp] * 8
    lib.qpo_bunch_fields.argtypes = ([ctypes.POINTER(QpoGrid), ctypes.c_size_t] + [_dp] * 12 +
                                     [ctypes.c_double] + [_dp] * 6)
    _ip = ctypes.POINTER(ctypes.c_int)
    lib.qpo_bunch_push.argtypes = [ctypes.c_size_t] + [_dp] * 8 + [_ip] + [_dp] * 5 + [ctypes.c_double]
    lib.qpo_num_threads.restype = ctypes.c_int
    lib.qpo_splitmix64.argtypes = [ctypes.c_ulonglong] * 3
    lib.qpo_splitmix64.restype = ctypes.c_ulonglong
    _libs[variant] = lib
    return lib


# ---------------------------------------------------------------- wrappers
def fft(x, sign=-1, variant="det"):
    x = np.ascontiguousarray(x, dtype=np.complex128).copy()
    buf = x.view(np.float64)
    load(variant).qpo_fft(len(x), _ptr(buf), sign)
    return x


def rfft2(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    n1, n2 = a.shape
    out = np.zeros((n1, n2 // 2 + 1), dtype=np.complex128)
    load().qpo_rfft2(n1, n2, _ptr(a), _ptr(out.view(np.float64)))
    return out


def irfft2(spec, n2):
    spec = np.ascontiguousarray(spec, dtype=np.complex128)
    n1 = spec.shape[0]
    out = np.zeros((n1, n2))
    load().qpo_irfft2(n1, n2, _ptr(spec.view(np.float64)), _ptr(out))
    return out


def solve_poisson(src, d1, d2, D=0.0):
    """Unnormalised periodic solve (returns n1*n2 times the solution)."""
    a = np.ascontiguousarray(src, dtype=np.float64).copy()
    n1, n2 = a.shape
    load().qpo_solve_poisson(n1, n2, d1, d2, D, _ptr(a))
    return a


def make_grid(nx, ny, nz, lx, ly, lz, dx, dy, dz, K=5, D=10.0, psi_threshold=1e-4, b_threshold=100.0):
    """Grid-size rule n = (int)(l/d), d = l/n (src/System_3d.cpp:87-95)."""
    nx_, ny_, nz_ = int(lx / dx), int(ly / dy), int(lz / dz)
    if nx is not None:
        assert nx == nx_ and ny == ny_ and nz == nz_, (nx_, ny_, nz_)
    return QpoGrid(nx_, ny_, nz_, lx / nx_, ly / ny_, lz / nz_, ly, lz, K, D, psi_threshold, b_threshold)


def grid_from_box(lx, ly, lz, dx, dy, dz, K=5, **kw):
    return make_grid(None, None, None, lx, ly, lz, dx, dy, dz, K=K, **kw)


def init_lattice(g, ppcy, ppcz, profile=None):
    prof = profile or QpoProfile(kind=PROFILE_CONSTANT, rho0=-1.0)
    lib = load()
    n = lib.qpo_init_lattice(g.ny, g.nz, g.dy, g.dz, ppcy, ppcz, ctypes.byref(prof), None, None, None)
    y, z, w = np.zeros(n), np.zeros(n), np.zeros(n)
    lib.qpo_init_lattice(g.ny, g.nz, g.dy, g.dz, ppcy, ppcz, ctypes.byref(prof), _ptr(y), _ptr(z), _ptr(w))
    return y, z, w


def gaussian_a_sqr(g, a0, xsigma, ysigma, zsigma, x0, y0, z0):
    a = np.zeros((g.nx, g.ny, g.nz))
    load().qpo_fill_a_sqr_gaussian(ctypes.byref(g), a0, xsigma, ysigma, zsigma, x0, y0, z0, _ptr(a))
    return a


def wake_run(g, y, z, w, a_sqr, py=None, pz=None, rho_b=None, jx_b=None, n_slices=0,
             fields=FIELD_NAMES, variant="det"):
    """Run the oracle slice loop; returns (dict of [nx,ny,nz] arrays, status, stop_slice)."""
    lib = load(variant)
    shape = (g.nx, g.ny, g.nz)
    out = {k: np.zeros(shape) for k in fields}
    f = QpoFields(*[_ptr(out[k]) if k in out else None for k in FIELD_NAMES])
    stop = ctypes.c_int(-1)
    y = np.ascontiguousarray(y, dtype=np.float64)
    z = np.ascontiguousarray(z, dtype=np.float64)
    w = np.ascontiguousarray(w, dtype=np.float64)
    py = None if py is None else np.ascontiguousarray(py, dtype=np.float64)
    pz = None if pz is None else np.ascontiguousarray(pz, dtype=np.float64)
    a_sqr = np.ascontiguousarray(a_sqr, dtype=np.float64)
    rho_b = None if rho_b is None else np.ascontiguousarray(rho_b, dtype=np.float64)
    jx_b = None if jx_b is None else np.ascontiguousarray(jx_b, dtype=np.float64)
    st = lib.qpo_wake_run_partial(ctypes.byref(g), len(y), _ptr(y), _ptr(z), _ptr(py), _ptr(pz), _ptr(w),
                                  _ptr(a_sqr), _ptr(rho_b), _ptr(jx_b), ctypes.byref(f), ctypes.byref(stop),
                                  int(n_slices))
    return out, st, stop.value


def wake_run_window(g, y, z, w, a_sqr, n_slices, py=None, pz=None, rho_b=None, jx_b=None,
                    fields=FIELD_NAMES, variant="det"):
    """The first n_slices slices of the loop with windowed 3D arrays: a_sqr,
    rho_b, jx_b are [n_slices, ny, nz] (slice i at i - (nx - n_slices)); the
    outputs are returned in the same layout."""
    lib = load(variant)
    shape = (n_slices, g.ny, g.nz)
    assert a_sqr.shape == shape
    for b in (rho_b, jx_b):
        assert b is None or b.shape == shape
    out = {k: np.zeros(shape) for k in fields}
    f = QpoFields(*[_ptr(out[k]) if k in out else None for k in FIELD_NAMES])
    stop = ctypes.c_int(-1)
    c = [None if v is None else np.ascontiguousarray(v, dtype=np.float64) for v in (y, z, py, pz, w)]
    a_sqr = np.ascontiguousarray(a_sqr, dtype=np.float64)
    rho_b = None if rho_b is None else np.ascontiguousarray(rho_b, dtype=np.float64)
    jx_b = None if jx_b is None else np.ascontiguousarray(jx_b, dtype=np.float64)
    st = lib.qpo_wake_run_window(ctypes.byref(g), len(c[0]), *[_ptr(v) for v in c], _ptr(a_sqr), _ptr(rho_b),
                                 _ptr(jx_b), ctypes.byref(f), ctypes.byref(stop), int(n_slices))
    return out, st, stop.value


def gaussian_a_sqr_window(g, n_slices, a0, xsigma, ysigma, zsigma, x0, y0, z0):
    """The top n_slices slices of gaussian_a_sqr (i from nx - n_slices to nx - 1)."""
    a = np.zeros((n_slices, g.ny, g.nz))
    load().qpo_fill_a_sqr_gaussian_window(ctypes.byref(g), g.nx - n_slices, n_slices, a0, xsigma, ysigma, zsigma,
                                          x0, y0, z0, _ptr(a))
    return a


# ------------------------------------------------------------ driver bunch
BUNCH_STATE = ("x", "y", "z", "px", "py", "pz", "gamma")
BUNCH_FIELDS = ("ex", "ey", "ez", "by", "bz", "chi")


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def deposit3d(g, x, y, z, v, a3):
    load().qpo_deposit3d(ctypes.byref(g), x, y, z, v, _ptr(a3))


def gather3d(g, x, y, z, a3, origin=(0.0, 0.0, 0.0)):
    return load().qpo_gather3d(ctypes.byref(g), *origin, x, y, z, _ptr(_f64(a3)))


def field_origins(g):
    """Origins of the 3D field arrays (src/System_3d.cpp:179-183)."""
    return {"ex": (-0.5 * g.dx, 0.0, 0.0), "ey": (0.0, 0.5 * g.dy, 0.0), "ez": (0.0, 0.0, 0.5 * g.dz),
            "by": (0.0, 0.0, 0.5 * g.dz), "bz": (0.0, 0.5 * g.dy, 0.0)}


def bunch_deposit(g, b):
    """deposit_bunch_densities (src/System_3d.cpp:926-946) of the bunch dict b
    (x, y, z, px, gamma, q arrays); returns (rho_b, jx_b) [nx, ny, nz]."""
    rho = np.zeros((g.nx, g.ny, g.nz))
    jx = np.zeros((g.nx, g.ny, g.nz))
    a = [_f64(b[k]) for k in ("x", "y", "z", "px", "gamma", "q")]
    load().qpo_bunch_deposit(ctypes.byref(g), len(a[0]), *[_ptr(v) for v in a], _ptr(rho), _ptr(jx))
    return rho, jx


def bunch_deposit_window(g, b, base, nheld):
    """bunch_deposit into arrays holding slices base .. base + nheld - 1 only."""
    rho = np.zeros((nheld, g.ny, g.nz))
    jx = np.zeros((nheld, g.ny, g.nz))
    a = [_f64(b[k]) for k in ("x", "y", "z", "px", "gamma", "q")]
    load().qpo_bunch_deposit_window(ctypes.byref(g), int(base), int(nheld), len(a[0]), *[_ptr(v) for v in a],
                                    _ptr(rho), _ptr(jx))
    return rho, jx


def bunch_fields(g, b, f3, field_schwinger):
    """deposit_wakefield_to_bunch (src/System_3d.cpp:843-860): dict of the five
    gathered fields and chi, from the 3D arrays f3['ex'..'bz']."""
    a = [_f64(b[k]) for k in BUNCH_STATE]
    fl = [_f64(f3[k]) for k in ("ex", "ey", "ez", "by", "bz")]
    n = len(a[0])
    out = {k: np.zeros(n) for k in BUNCH_FIELDS}
    load().qpo_bunch_fields(ctypes.byref(g), n, *[_ptr(v) for v in a], *[_ptr(v) for v in fl],
                            field_schwinger, *[_ptr(out[k]) for k in BUNCH_FIELDS])
    return out


def bunch_push(b, fields, dt):
    """advance_bunch_particles without QED (src/System_3d.cpp:862-890); returns
    a new dict with the advanced state (cmr, photon, q carried over)."""
    nb = {k: _f64(b[k]).copy() for k in BUNCH_STATE}
    ph = np.ascontiguousarray(b["photon"], dtype=np.int32)
    fl = [_f64(fields[k]) for k in ("ex", "ey", "ez", "by", "bz")]
    load().qpo_bunch_push(len(nb["x"]), *[_ptr(nb[k]) for k in BUNCH_STATE], _ptr(_f64(b["cmr"])),
                          ph.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), *[_ptr(v) for v in fl], dt)
    for k in ("q", "cmr", "photon"):
        nb[k] = b[k]
    return nb
