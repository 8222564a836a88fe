"""ctypes binding of the C-ABI in include/quetzal_hip.h (libquetzal_hip.so).

This is the Python-side mirror of the drop-in boundary used by the tests and
bench.py.  It has no fallback: if the HIP library is missing, loading raises.
"""
import ctypes
import os
import weakref

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libquetzal_hip.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "quetzal_hip.h")

QP_OK, QP_STOPPED = 0, 1
QP_ERR_ARG, QP_ERR_HIP, QP_ERR_STATE, QP_ERR_ALLOC = -1, -2, -3, -4
QP_PREC_FP64, QP_PREC_MIXED = 0, 1
QP_SLICE_NORMAL, QP_SLICE_STOPPED, QP_SLICE_FINAL = 0, 1, 2
FIELD_NAMES = ("psi", "rho", "jx", "jy", "jz", "ex", "ey", "ez", "by", "bz", "chi")

_dp = ctypes.POINTER(ctypes.c_double)


class QpError(RuntimeError):
    pass


class WakeParams(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("dx", ctypes.c_double), ("dy", ctypes.c_double), ("dz", ctypes.c_double),
        ("ly", ctypes.c_double), ("lz", ctypes.c_double),
        ("K", ctypes.c_int), ("D", ctypes.c_double),
        ("psi_threshold", ctypes.c_double), ("b_threshold", ctypes.c_double),
        ("precision", ctypes.c_int), ("use_graph", ctypes.c_int), ("resort_every", ctypes.c_int),
        ("longitudinal_b", ctypes.c_int), ("async_output", ctypes.c_int), ("slice_limit", ctypes.c_int),
    ]


class SliceView(ctypes.Structure):
    _fields_ = [("slice", ctypes.c_int), ("kind", ctypes.c_int)] + [(n, _dp) for n in FIELD_NAMES] + [("bx", _dp)]


SLICE_CB = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.POINTER(SliceView), ctypes.c_void_p)
BATCH_CB = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_int, ctypes.POINTER(SliceView), ctypes.c_void_p)
GROUP_CB = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(SliceView), ctypes.c_void_p)

_lib = None


def load():
    """Load libquetzal_hip.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise QpError(f"{LIB_PATH} not built: run __graft_entry__.build()")
    lib = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    sig = {
        "qp_last_error": ([], ctypes.c_char_p),
        "qp_version": ([], ctypes.c_char_p),
        "qp_ctx_create": ([ctypes.c_int, ctypes.POINTER(vp)], ctypes.c_int),
        "qp_ctx_destroy": ([vp], ctypes.c_int),
        "qp_ctx_synchronize": ([vp], ctypes.c_int),
        "qp_fourier2d_create": ([vp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(vp)], ctypes.c_int),
        "qp_fourier2d_destroy": ([vp], ctypes.c_int),
        "qp_fourier2d_forward": ([vp, _dp, _dp], ctypes.c_int),
        "qp_fourier2d_backward": ([vp, _dp, _dp], ctypes.c_int),
        "qp_solve2d": ([vp, ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp, _dp], ctypes.c_int),
        "qp_solve2d_device": ([vp, ctypes.c_double, ctypes.c_double, ctypes.c_double, vp, vp], ctypes.c_int),
        "qp_solve2d_device_noalloc": ([vp, ctypes.c_int, ctypes.c_int] + [ctypes.c_double] * 3 + [vp, vp],
                                      ctypes.c_int),
        "qp_wake_create": ([vp, ctypes.POINTER(WakeParams), ctypes.POINTER(vp)], ctypes.c_int),
        "qp_wake_destroy": ([vp], ctypes.c_int),
        "qp_wake_set_particles": ([vp, ctypes.c_size_t, _dp, _dp, _dp, _dp, _dp], ctypes.c_int),
        "qp_wake_set_a_sqr": ([vp, _dp], ctypes.c_int),
        "qp_wake_set_laser_gaussian": ([vp] + [ctypes.c_double] * 7, ctypes.c_int),
        "qp_wake_set_bunch_sources": ([vp, _dp, _dp], ctypes.c_int),
        "qp_wake_run": ([vp, SLICE_CB, vp, ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
        "qp_wake_last_time_ms": ([vp, _dp], ctypes.c_int),
        "qp_wake_profile_kernel": ([vp, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int), _dp], ctypes.c_int),
        "qp_wake_get_particles": ([vp, ctypes.c_size_t, _dp, _dp, _dp, _dp], ctypes.c_int),
        "qp_wake_describe": ([vp, ctypes.c_char_p, ctypes.c_size_t], ctypes.c_int),
        "qp_wake_window_stats": ([vp, ctypes.c_int, ctypes.POINTER(ctypes.c_longlong),
                                  ctypes.POINTER(ctypes.c_longlong)], ctypes.c_int),
        "qp_wake_create_batch": ([vp, ctypes.POINTER(WakeParams), ctypes.c_int, ctypes.POINTER(vp)], ctypes.c_int),
        "qp_wake_batch_size": ([vp, ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
        "qp_wake_set_particles_at": ([vp, ctypes.c_int, ctypes.c_size_t, _dp, _dp, _dp, _dp, _dp], ctypes.c_int),
        "qp_wake_set_laser_gaussian_at": ([vp, ctypes.c_int] + [ctypes.c_double] * 7, ctypes.c_int),
        "qp_wake_run_batch": ([vp, BATCH_CB, vp, ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
        "qp_wake_get_particles_at": ([vp, ctypes.c_int, ctypes.c_size_t, _dp, _dp, _dp, _dp], ctypes.c_int),
        "qp_group_create": ([vp, ctypes.c_int, ctypes.POINTER(vp), ctypes.POINTER(vp)], ctypes.c_int),
        "qp_group_destroy": ([vp], ctypes.c_int),
        "qp_group_run": ([vp, GROUP_CB, vp, ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
        "qp_group_last_time_ms": ([vp, _dp], ctypes.c_int),
        "qp_group_set_tile": ([vp, ctypes.c_int, ctypes.c_int], ctypes.c_int),
        "qp_group_set_fork": ([vp, ctypes.c_int], ctypes.c_int),
        "qp_wake_set_tile": ([vp, ctypes.c_int, ctypes.c_int], ctypes.c_int),
        "qp_envelope_create": ([vp] + [ctypes.c_int] * 3 + [ctypes.c_double] * 4 + [ctypes.POINTER(vp)],
                               ctypes.c_int),
        "qp_envelope_destroy": ([vp], ctypes.c_int),
        "qp_envelope_set": ([vp, _dp, _dp], ctypes.c_int),
        "qp_envelope_set_gaussian": ([vp] + [ctypes.c_double] * 7, ctypes.c_int),
        "qp_envelope_get": ([vp, _dp, _dp], ctypes.c_int),
        "qp_envelope_set_chi": ([vp, _dp], ctypes.c_int),
        "qp_envelope_get_chi": ([vp, _dp], ctypes.c_int),
        "qp_envelope_get_a_sqr": ([vp, _dp], ctypes.c_int),
        "qp_envelope_advance": ([vp, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp], ctypes.c_int),
        "qp_wake_attach_envelope": ([vp, vp], ctypes.c_int),
        "qp_wake_set_envelope_step": ([vp, ctypes.c_double, ctypes.c_int], ctypes.c_int),
        "qp_bunch_create": ([vp] + [ctypes.c_int] * 3 + [ctypes.c_double] * 4 + [ctypes.POINTER(vp)], ctypes.c_int),
        "qp_bunch_destroy": ([vp], ctypes.c_int),
        "qp_bunch_set_particles": ([vp, ctypes.c_size_t] + [_dp] * 9 + [ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
        "qp_bunch_size": ([vp, ctypes.POINTER(ctypes.c_size_t)], ctypes.c_int),
        "qp_bunch_get_particles": ([vp, ctypes.c_size_t] + [_dp] * 7, ctypes.c_int),
        "qp_bunch_get_fields": ([vp, ctypes.c_size_t] + [_dp] * 6, ctypes.c_int),
        "qp_bunch_push": ([vp, ctypes.c_double], ctypes.c_int),
        "qp_bunch_deposit": ([vp, _dp, _dp], ctypes.c_int),
        "qp_wake_attach_bunch": ([vp, vp], ctypes.c_int),
    }
    for name, (args, res) in sig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = res
    _lib = lib
    return lib


def _check(st, what):
    if st < 0:
        raise QpError(f"{what} failed ({st}): {load().qp_last_error().decode()}")
    return st


def _f64(a, shape, what, dtype=np.float64):
    """a as a C-contiguous array of `dtype` with exactly `shape` (an int: that
    many elements in one dimension, any shape for a tuple match), else
    ValueError: the C side reads exactly that many values through the pointer."""
    if a is None:
        return None
    a = np.ascontiguousarray(a, dtype=dtype)
    want = shape if isinstance(shape, tuple) else (int(shape),)
    if a.shape != want:
        raise ValueError(f"{what}: shape {a.shape}, expected {want}")
    return a


def _particles(y, z, py, pz, w, what):
    """Particle SoA arrays, all of len(y)."""
    y = np.ascontiguousarray(y, dtype=np.float64)
    if y.ndim != 1:
        raise ValueError(f"{what}: y must be one-dimensional, got shape {y.shape}")
    n = len(y)
    z = _f64(z, n, f"{what} z")
    w = _f64(w, n, f"{what} weight")
    py = _f64(py, n, f"{what} py")
    pz = _f64(pz, n, f"{what} pz")
    return y, z, py, pz, w


def _ptr(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


class Context:
    """qp_ctx: one HIP device and one stream."""

    def __init__(self, device=0):
        self.h = ctypes.c_void_p()
        self._children = weakref.WeakSet()
        _check(load().qp_ctx_create(device, ctypes.byref(self.h)), "qp_ctx_create")

    def _adopt(self, child):
        self._children.add(child)

    def synchronize(self):
        _check(load().qp_ctx_synchronize(self.h), "qp_ctx_synchronize")

    def close(self):
        """Destroy the context; objects created on it are closed first."""
        for c in list(self._children):
            c.close()
        if self.h:
            load().qp_ctx_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Fourier2d:
    """Mirror of the reference's Fourier2d (src/Fourier2d.h:4-19): r2c / c2r,
    unnormalised, n1 x n2 real <-> n1 x (n2/2+1) complex."""

    def __init__(self, ctx, n1, n2):
        self.ctx, self.n1, self.n2 = ctx, n1, n2
        self.h = ctypes.c_void_p()
        _check(load().qp_fourier2d_create(ctx.h, n1, n2, ctypes.byref(self.h)), "qp_fourier2d_create")
        ctx._adopt(self)

    def forward_transform(self, real):
        real = _f64(real, (self.n1, self.n2), "forward_transform")
        out = np.zeros((self.n1, self.n2 // 2 + 1), dtype=np.complex128)
        _check(load().qp_fourier2d_forward(self.h, _ptr(real), _ptr(out.view(np.float64))), "forward")
        return out

    def backward_transform(self, spec):
        spec = _f64(spec, (self.n1, self.n2 // 2 + 1), "backward_transform", np.complex128)
        out = np.zeros((self.n1, self.n2))
        _check(load().qp_fourier2d_backward(self.h, _ptr(spec.view(np.float64)), _ptr(out)), "backward")
        return out

    def solve_poisson_equation(self, src, d1, d2, D=0.0):
        """System_3d::solve_poisson_equation (src/System_3d.cpp:686-706); unnormalised."""
        src = _f64(src, (self.n1, self.n2), "solve_poisson_equation")
        out = np.zeros_like(src)
        _check(load().qp_solve2d(self.h, d1, d2, D, _ptr(src), _ptr(out)), "qp_solve2d")
        return out

    def close(self):
        if self.h:
            load().qp_fourier2d_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Wake:
    """One wake on the device: System_3d::solve_wakefield behind the C-ABI."""

    def __init__(self, ctx, nx, ny, nz, dx, dy, dz, ly, lz, K=5, D=10.0, psi_threshold=1e-4,
                 b_threshold=100.0, precision=QP_PREC_FP64, use_graph=1, resort_every=0, longitudinal_b=0,
                 async_output=0, slice_limit=0):
        self.ctx = ctx
        self.shape = (nx, ny, nz)
        self.params = WakeParams(nx, ny, nz, dx, dy, dz, ly, lz, K, D, psi_threshold, b_threshold,
                                 precision, use_graph, resort_every, longitudinal_b, async_output, slice_limit)
        self.h = ctypes.c_void_p()
        self._env = None
        self._bunch = None
        _check(load().qp_wake_create(ctx.h, ctypes.byref(self.params), ctypes.byref(self.h)), "qp_wake_create")
        self.np = 0
        ctx._adopt(self)

    def set_particles(self, y, z, w, py=None, pz=None):
        y, z, py, pz, w = _particles(y, z, py, pz, w, "Wake.set_particles")
        self.np = len(y)
        _check(load().qp_wake_set_particles(self.h, len(y), _ptr(y), _ptr(z), _ptr(py), _ptr(pz), _ptr(w)),
               "qp_wake_set_particles")

    def set_a_sqr(self, a):
        a = _f64(a, self.shape, "set_a_sqr")
        _check(load().qp_wake_set_a_sqr(self.h, _ptr(a)), "qp_wake_set_a_sqr")

    def set_laser_gaussian(self, a0, xsigma, ysigma, zsigma, x0, y0, z0):
        _check(load().qp_wake_set_laser_gaussian(self.h, a0, xsigma, ysigma, zsigma, x0, y0, z0),
               "qp_wake_set_laser_gaussian")

    def set_bunch_sources(self, rho_b=None, jx_b=None):
        rho_b = _f64(rho_b, self.shape, "set_bunch_sources rho_b")
        jx_b = _f64(jx_b, self.shape, "set_bunch_sources jx_b")
        _check(load().qp_wake_set_bunch_sources(self.h, _ptr(rho_b), _ptr(jx_b)), "qp_wake_set_bunch_sources")

    def run(self, fields=None, callback=None, slices=None):
        """Run one pass.  fields: iterable of names to collect into [nx,ny,nz]
        arrays with the reference's semantics (2D per-slice writes for psi..jz,
        3D arrays for ex..chi).  slices=(lo, hi): collect only slices lo..hi-1,
        into [hi-lo,ny,nz] arrays (slice i at i - lo).  Returns (dict, status,
        stop_slice)."""
        nx, ny, nz = self.shape
        lo, hi = slices if slices is not None else (0, nx)
        assert 0 <= lo < hi <= nx
        out = {}
        if fields:
            out = {k: np.zeros((hi - lo, ny, nz)) for k in fields}
        stop = ctypes.c_int(-1)
        if fields or callback:
            nn = ny * nz

            def _cb(vp, _user):
                v = vp.contents
                i = v.slice - lo
                for k in (out if 0 <= i < hi - lo else ()):
                    if v.kind == QP_SLICE_STOPPED and k not in ("by", "bz", "chi"):
                        continue
                    ptr = getattr(v, k)
                    if not ptr:  # bx without longitudinal_b
                        continue
                    out[k][i] = np.ctypeslib.as_array(ptr, shape=(nn,)).reshape(ny, nz)
                if callback:
                    return int(callback(v) or 0)
                return 0

            cb = SLICE_CB(_cb)
        else:
            cb = SLICE_CB()
        st = _check(load().qp_wake_run(self.h, cb, None, ctypes.byref(stop)), "qp_wake_run")
        return out, st, stop.value

    def last_time_ms(self):
        ms = ctypes.c_double()
        _check(load().qp_wake_last_time_ms(self.h, ctypes.byref(ms)), "qp_wake_last_time_ms")
        return ms.value

    def profile_kernel(self, cls):
        n = ctypes.c_int()
        us = ctypes.c_double()
        _check(load().qp_wake_profile_kernel(self.h, cls.encode(), ctypes.byref(n), ctypes.byref(us)),
               "qp_wake_profile_kernel")
        return n.value, us.value

    def describe(self):
        buf = ctypes.create_string_buffer(512)
        _check(load().qp_wake_describe(self.h, buf, 512), "qp_wake_describe")
        return buf.value.decode()

    def set_tile(self, threads_per_block, window_nodes):
        """Force the particle tile of the mixed launches ((0, 0): let the autotuner pick)."""
        _check(load().qp_wake_set_tile(self.h, threads_per_block, window_nodes), "qp_wake_set_tile")

    def window_stats(self, member=0):
        """(overflow_blocks, deposit_blocks) of the last mixed pass."""
        o, t = ctypes.c_longlong(), ctypes.c_longlong()
        _check(load().qp_wake_window_stats(self.h, member, ctypes.byref(o), ctypes.byref(t)), "qp_wake_window_stats")
        return o.value, t.value

    def get_particles(self):
        y, z, py, pz = (np.zeros(self.np) for _ in range(4))
        _check(load().qp_wake_get_particles(self.h, self.np, _ptr(y), _ptr(z), _ptr(py), _ptr(pz)),
               "qp_wake_get_particles")
        return y, z, py, pz

    def attach_envelope(self, env):
        """Read <a^2> = |a|^2/2 from `env` and store each slice's chi into it (None detaches)."""
        _check(load().qp_wake_attach_envelope(self.h, env.h if env is not None else None),
               "qp_wake_attach_envelope")
        if self._env is not None:
            self._env._wakes.discard(self)
        self._env = env
        if env is not None:
            env._wakes.add(self)

    def set_envelope_step(self, dtau, chi_iterations=1):
        """dtau > 0: every run also advances the attached envelope by one step of dtau
        with that pass's chi, enqueued behind the pass; 0 turns it off."""
        _check(load().qp_wake_set_envelope_step(self.h, dtau, chi_iterations), "qp_wake_set_envelope_step")

    def attach_bunch(self, bunch):
        """Stream `bunch` through the slice loop: per-slice sources and field gathers (None detaches)."""
        _check(load().qp_wake_attach_bunch(self.h, bunch.h if bunch is not None else None), "qp_wake_attach_bunch")
        if self._bunch is not None:
            self._bunch._wakes.discard(self)
        self._bunch = bunch
        if bunch is not None:
            bunch._wakes.add(self)

    def close(self):
        if self.h:
            load().qp_wake_destroy(self.h)
            self.h = ctypes.c_void_p()
        if getattr(self, "_env", None) is not None:
            self._env._wakes.discard(self)
            self._env = None
        if getattr(self, "_bunch", None) is not None:
            self._bunch._wakes.discard(self)
            self._bunch = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Envelope:
    """qp_envelope: complex laser envelope a[i][j][k] on a wake's grid, advanced in tau
    (no counterpart in the reference; the scheme is stated in include/quetzal_hip.h)."""

    def __init__(self, ctx, nx, ny, nz, dx, dy, dz, k0):
        self.h = ctypes.c_void_p()
        self.shape = (nx, ny, nz)
        self._wakes = weakref.WeakSet()
        _check(load().qp_envelope_create(ctx.h, nx, ny, nz, dx, dy, dz, k0, ctypes.byref(self.h)),
               "qp_envelope_create")
        ctx._adopt(self)

    def set(self, a):
        a = np.asarray(a)
        if a.shape != self.shape:
            raise ValueError(f"Envelope.set: shape {a.shape}, expected {self.shape}")
        re = np.ascontiguousarray(a.real, dtype=np.float64)
        im = np.ascontiguousarray(a.imag, dtype=np.float64) if np.iscomplexobj(a) else None
        _check(load().qp_envelope_set(self.h, _ptr(re), _ptr(im)), "qp_envelope_set")

    def set_gaussian(self, a0, xsigma, ysigma, zsigma, x0, y0, z0):
        _check(load().qp_envelope_set_gaussian(self.h, a0, xsigma, ysigma, zsigma, x0, y0, z0),
               "qp_envelope_set_gaussian")

    def get(self):
        re, im = np.zeros(self.shape), np.zeros(self.shape)
        _check(load().qp_envelope_get(self.h, _ptr(re), _ptr(im)), "qp_envelope_get")
        return re + 1j * im

    def set_chi(self, chi):
        c = _f64(chi, self.shape, "Envelope.set_chi")
        _check(load().qp_envelope_set_chi(self.h, _ptr(c)), "qp_envelope_set_chi")

    def get_chi(self):
        c = np.zeros(self.shape)
        _check(load().qp_envelope_get_chi(self.h, _ptr(c)), "qp_envelope_get_chi")
        return c

    def get_a_sqr(self):
        c = np.zeros(self.shape)
        _check(load().qp_envelope_get_a_sqr(self.h, _ptr(c)), "qp_envelope_get_a_sqr")
        return c

    def advance(self, dtau, steps=1, chi_iterations=2, use_graph=1):
        """Returns the device time of the call in ms."""
        ms = ctypes.c_double()
        _check(load().qp_envelope_advance(self.h, dtau, steps, chi_iterations, use_graph, ctypes.byref(ms)),
               "qp_envelope_advance")
        return ms.value

    def close(self):
        for w in list(self._wakes):
            w.attach_envelope(None)
        if self.h:
            _check(load().qp_envelope_destroy(self.h), "qp_envelope_destroy")
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Bunch:
    """qp_bunch: driver-bunch particles on the device (deposit, field gather and
    push of System_3d's bunch, src/System_3d.cpp:843-946, without QED emission)."""

    STATE = ("x", "y", "z", "px", "py", "pz", "gamma")
    FIELDS = ("ex", "ey", "ez", "by", "bz", "chi")

    def __init__(self, ctx, nx, ny, nz, dx, dy, dz, field_schwinger=1.0):
        self.h = ctypes.c_void_p()
        self.shape = (nx, ny, nz)
        self.np = 0
        self._wakes = weakref.WeakSet()
        _check(load().qp_bunch_create(ctx.h, nx, ny, nz, dx, dy, dz, field_schwinger, ctypes.byref(self.h)),
               "qp_bunch_create")
        ctx._adopt(self)

    def set_particles(self, b):
        """b: dict with x, y, z, px, py, pz, gamma, q, cmr (float) and photon (int) arrays."""
        n = len(np.asarray(b["x"]))
        a = [_f64(b[k], n, f"Bunch.set_particles {k}") for k in self.STATE + ("q", "cmr")]
        ph = _f64(b["photon"], n, "Bunch.set_particles photon", np.int32)
        _check(load().qp_bunch_set_particles(self.h, n, *[_ptr(v) for v in a],
                                             ph.ctypes.data_as(ctypes.POINTER(ctypes.c_int))),
               "qp_bunch_set_particles")
        self.np = n

    def get_particles(self):
        out = {k: np.zeros(self.np) for k in self.STATE}
        _check(load().qp_bunch_get_particles(self.h, self.np, *[_ptr(out[k]) for k in self.STATE]),
               "qp_bunch_get_particles")
        return out

    def get_fields(self):
        out = {k: np.zeros(self.np) for k in self.FIELDS}
        _check(load().qp_bunch_get_fields(self.h, self.np, *[_ptr(out[k]) for k in self.FIELDS]),
               "qp_bunch_get_fields")
        return out

    def push(self, dt):
        _check(load().qp_bunch_push(self.h, dt), "qp_bunch_push")

    def deposit(self):
        rho, jx = np.zeros(self.shape), np.zeros(self.shape)
        _check(load().qp_bunch_deposit(self.h, _ptr(rho), _ptr(jx)), "qp_bunch_deposit")
        return rho, jx

    def close(self):
        for w in list(self._wakes):
            w.attach_bunch(None)
        if self.h:
            _check(load().qp_bunch_destroy(self.h), "qp_bunch_destroy")
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class WakeBatch:
    """nwakes independent wakes with one geometry, advanced together so every
    launch carries all of them (qp_wake_create_batch; mixed precision).  Each
    member has its own particles, Gaussian laser, fields and stop rule."""

    def __init__(self, ctx, nwakes, nx, ny, nz, dx, dy, dz, ly, lz, K=5, D=10.0, psi_threshold=1e-4,
                 b_threshold=100.0, use_graph=1, longitudinal_b=0, async_output=0, slice_limit=0):
        self.ctx = ctx
        self.nwakes = nwakes
        self.shape = (nx, ny, nz)
        self.params = WakeParams(nx, ny, nz, dx, dy, dz, ly, lz, K, D, psi_threshold, b_threshold,
                                 QP_PREC_MIXED, use_graph, 0, longitudinal_b, async_output, slice_limit)
        self.h = ctypes.c_void_p()
        _check(load().qp_wake_create_batch(ctx.h, ctypes.byref(self.params), nwakes, ctypes.byref(self.h)),
               "qp_wake_create_batch")
        self.nps = [0] * nwakes
        ctx._adopt(self)

    def set_particles(self, member, y, z, w, py=None, pz=None):
        if not 0 <= member < self.nwakes:
            raise ValueError(f"WakeBatch.set_particles: member {member} outside [0, {self.nwakes})")
        y, z, py, pz, w = _particles(y, z, py, pz, w, "WakeBatch.set_particles")
        self.nps[member] = len(y)
        _check(load().qp_wake_set_particles_at(self.h, member, len(y), _ptr(y), _ptr(z), _ptr(py), _ptr(pz),
                                               _ptr(w)), "qp_wake_set_particles_at")

    def set_laser_gaussian(self, member, a0, xsigma, ysigma, zsigma, x0, y0, z0):
        _check(load().qp_wake_set_laser_gaussian_at(self.h, member, a0, xsigma, ysigma, zsigma, x0, y0, z0),
               "qp_wake_set_laser_gaussian_at")

    def run(self, fields=None, callback=None, members=None):
        """One pass of every member.  Returns (list of per-member field dicts as
        Wake.run, status, list of stop slices); members: indices whose fields are
        collected (default all; the others get empty dicts)."""
        nx, ny, nz = self.shape
        keep = set(range(self.nwakes)) if members is None else set(members)
        outs = [{k: np.zeros(self.shape) for k in fields} if fields and m in keep else {}
                for m in range(self.nwakes)]
        stops = (ctypes.c_int * self.nwakes)()
        if fields or callback:
            nn = ny * nz

            def _cb(member, vp, _user):
                v = vp.contents
                out = outs[member]
                for k in out:
                    if v.kind == QP_SLICE_STOPPED and k not in ("by", "bz", "chi"):
                        continue
                    ptr = getattr(v, k)
                    if not ptr:
                        continue
                    out[k][v.slice] = np.ctypeslib.as_array(ptr, shape=(nn,)).reshape(ny, nz)
                if callback:
                    return int(callback(member, v) or 0)
                return 0

            cb = BATCH_CB(_cb)
        else:
            cb = BATCH_CB()
        st = _check(load().qp_wake_run_batch(self.h, cb, None, stops), "qp_wake_run_batch")
        return outs, st, list(stops)

    def last_time_ms(self):
        ms = ctypes.c_double()
        _check(load().qp_wake_last_time_ms(self.h, ctypes.byref(ms)), "qp_wake_last_time_ms")
        return ms.value

    def profile_kernel(self, cls):
        n = ctypes.c_int()
        us = ctypes.c_double()
        _check(load().qp_wake_profile_kernel(self.h, cls.encode(), ctypes.byref(n), ctypes.byref(us)),
               "qp_wake_profile_kernel")
        return n.value, us.value

    def describe(self):
        buf = ctypes.create_string_buffer(512)
        _check(load().qp_wake_describe(self.h, buf, 512), "qp_wake_describe")
        return buf.value.decode()

    def set_tile(self, threads_per_block, window_nodes):
        """Force the particle tile of the mixed launches ((0, 0): let the autotuner pick)."""
        _check(load().qp_wake_set_tile(self.h, threads_per_block, window_nodes), "qp_wake_set_tile")

    def window_stats(self, member=0):
        """(overflow_blocks, deposit_blocks) of the last pass for one member."""
        o, t = ctypes.c_longlong(), ctypes.c_longlong()
        _check(load().qp_wake_window_stats(self.h, member, ctypes.byref(o), ctypes.byref(t)), "qp_wake_window_stats")
        return o.value, t.value

    def get_particles(self, member):
        n = self.nps[member]
        y, z, py, pz = (np.zeros(n) for _ in range(4))
        _check(load().qp_wake_get_particles_at(self.h, member, n, _ptr(y), _ptr(z), _ptr(py), _ptr(pz)),
               "qp_wake_get_particles_at")
        return y, z, py, pz

    def close(self):
        if self.h:
            load().qp_wake_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class WakeGroup:
    """Several WakeBatch objects (each its own transverse grid) advanced in
    lockstep, one launch per phase of the slice step over all of them
    (qp_group_create).  The group borrows the batches: close it first."""

    def __init__(self, ctx, batches):
        self.ctx = ctx
        self.batches = list(batches)
        arr = (ctypes.c_void_p * len(self.batches))(*[b.h for b in self.batches])
        self.h = ctypes.c_void_p()
        _check(load().qp_group_create(ctx.h, len(self.batches), arr, ctypes.byref(self.h)), "qp_group_create")
        ctx._adopt(self)

    def run(self, fields=None, callback=None, members=None):
        """One pass of every member of every batch.  Returns (outs, status, stops):
        outs[b][m] is a field dict as WakeBatch.run (empty unless (b, m) is in
        `members`, default all), stops[b][m] the stop slice or -1."""
        keep = None if members is None else set(members)
        outs = []
        for bi, b in enumerate(self.batches):
            outs.append([{k: np.zeros(b.shape) for k in fields} if fields and (keep is None or (bi, m) in keep)
                         else {} for m in range(b.nwakes)])
        ntot = sum(b.nwakes for b in self.batches)
        stops = (ctypes.c_int * ntot)()
        if fields or callback:
            def _cb(bi, member, vp, _user):
                v = vp.contents
                out = outs[bi][member]
                ny, nz = self.batches[bi].shape[1:]
                for k in out:
                    if v.kind == QP_SLICE_STOPPED and k not in ("by", "bz", "chi"):
                        continue
                    ptr = getattr(v, k)
                    if not ptr:
                        continue
                    out[k][v.slice] = np.ctypeslib.as_array(ptr, shape=(ny * nz,)).reshape(ny, nz)
                if callback:
                    return int(callback(bi, member, v) or 0)
                return 0

            cb = GROUP_CB(_cb)
        else:
            cb = GROUP_CB()
        st = _check(load().qp_group_run(self.h, cb, None, stops), "qp_group_run")
        flat = list(stops)
        res, off = [], 0
        for b in self.batches:
            res.append(flat[off:off + b.nwakes])
            off += b.nwakes
        return outs, st, res

    def last_time_ms(self):
        ms = ctypes.c_double()
        _check(load().qp_group_last_time_ms(self.h, ctypes.byref(ms)), "qp_group_last_time_ms")
        return ms.value

    def set_tile(self, threads_per_block, window_nodes):
        """Force the particle tile of every group launch ((0, 0): tune again)."""
        _check(load().qp_group_set_tile(self.h, threads_per_block, window_nodes), "qp_group_set_tile")

    def set_fork(self, fork_families):
        """Run the two transform families' solver launches on two streams (True, the
        default) or in sequence on the context's stream (False)."""
        _check(load().qp_group_set_fork(self.h, 1 if fork_families else 0), "qp_group_set_fork")

    def close(self):
        if self.h:
            load().qp_group_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
