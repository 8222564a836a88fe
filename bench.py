#!/usr/bin/env python3
"""Benchmark of the quasistatic slice loop on one MI355X.

Default workload: C5, the largest single-GPU configuration (1024^2 grid,
4 194 304 plasma particles, 4000 slices, a 7.7 M-particle electron bunch
streamed through the slice loop plus a trailing laser).  A step is one beam
time step of System_3d::run (src/System_3d.cpp:199-236): one mixed-precision
wake pass with the bunch's per-slice sources and field gathers on the device,
then the bunch push.  value = plasma particle-slice pushes of the timed steps
/ wall time = N_p (n_x - 1) K / t.  C1-C3 (one pass per step) and C4 (one
pass of 64 wakes per step) are selected with --workload.  Prints one JSON line
on stdout (progress on stderr).
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# C4: groups (streams) the five grid-size batches run in (measured, see bench_c4).
C4_CHAINS = 3
FIELD_NAMES_ALL = ("psi", "rho", "jx", "jy", "jz", "ex", "ey", "ez", "by", "bz", "chi")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def bytes_per_particle(cls, K):
    """Algorithmic HBM bytes per particle and launch of a particle phase in the
    build's own layout (persistent state S = 36 B: u32 cell, 2 x u32 in-cell
    offsets, 2 x f64 momenta, f64 weight; scratch T = 16 B: 2 x f64 predicted
    momenta p*):
      A: read cell, offsets, weight, p* (36), write cell, offsets, momenta (28) = 64
      B: read S (+ T after the first iteration), write T                      = 52 | 68
      C: read S + T                                                           = 52
      S: read S                                                               = 36
    648 B per particle-slice at K = 5."""
    if cls == "push_A":
        return 64.0
    if cls == "push_B":
        return (52.0 + 68.0 * (K - 1)) / K if K > 0 else 52.0
    if cls == "push_C":
        return 52.0
    if cls == "push_S":
        return 36.0
    raise KeyError(cls)


def bytes_per_particle_spec(cls, K):
    """The same for the reference layout of the target statement (persistent
    S = 24 B: int cell 4, 2 x f32 offset, 2 x f32 momenta, f32 weight; scratch
    T = 8 B: 2 x f32 p*): A reads S + T and writes position and momentum (20 B)
    = 52; B_1 reads S and writes T = 32; B_k reads S + T and writes T = 40;
    C reads S + T = 32.  116 + 72 (K - 1) = 404 B per particle-slice at K = 5.
    The roofline's `achieved` uses these bytes."""
    if cls == "push_A":
        return 52.0
    if cls == "push_B":
        return (32.0 + 40.0 * (K - 1)) / K if K > 0 else 32.0
    if cls == "push_C":
        return 32.0
    if cls == "push_S":
        return 24.0
    raise KeyError(cls)


def spec_bytes_per_particle_slice(K):
    return 116.0 + 72.0 * (K - 1)


def solve_bytes(n1, n2):
    """Algorithmic bytes of one three-launch solve: real source read + half
    spectrum write, spectrum read + write, spectrum read + real write."""
    nh = n2 // 2 + 1
    real = 8.0 * n1 * n2
    spec = 16.0 * n1 * nh
    return (real + spec) + 2 * spec + (spec + real)


def solve_bytes_slice(n1, n2, K, bunch):
    """Algorithmic bytes of one slice's solves with the sources and sinks fused
    into their first and last launches: 1 + K psi solves (source: the deposit
    accumulator and the ion background, plus the bunch's rho and j_x planes when
    a bunch drives the wake; src/System_3d.cpp:342-365, 437-469) and K B_y / B_z
    pair solves (sources from B_y, B_z, j_x, j_y, j_z and the previous j_y, j_z;
    :529-564; results B_y, B_z).  Every plane read once, each half spectrum
    written, read and written by the column pass and read once, each result
    written once."""
    nh = n2 // 2 + 1
    real = 8.0 * n1 * n2
    spec = 16.0 * n1 * nh
    psi = (4 if bunch else 2) * real + spec + 2 * spec + spec + real
    bpair = 7 * real + 2 * spec + 4 * spec + 2 * spec + 2 * real
    return (1 + K) * psi + K * bpair


def read_traffic(kernel_cls, workload):
    """Per-launch HBM bytes of a kernel from the committed rocprofv3 --pmc summary
    (profiles/traffic_<class>_<workload>.txt: lines 'FETCH_KIB <v>', 'WRITE_KIB <v>',
    'FETCH_FACTOR <f>', 'WRITE_FACTOR <f>', the factors from a device copy of
    known size); None when there is no summary for this workload."""
    p = os.path.join(ROOT, "profiles", f"traffic_{kernel_cls}_{workload}.txt")
    if not os.path.exists(p):
        return None
    vals = {}
    with open(p) as f:
        for line in f:
            parts = line.split()
            if len(parts) == 2:
                try:
                    vals[parts[0]] = float(parts[1])
                except ValueError:
                    pass
    if "FETCH_KIB" not in vals or "WRITE_KIB" not in vals:
        return None
    return 1024.0 * (vals["FETCH_KIB"] * vals.get("FETCH_FACTOR", 1.0) +
                     vals["WRITE_KIB"] * vals.get("WRITE_FACTOR", 1.0))


# Laser wavenumber of the envelope leg in units of k_p (not a reference input:
# the reference has no envelope advance), its step and chi corrections.
ENV_K0, ENV_DTAU, ENV_CHI_IT = 20.0, 2.0, 1


def envelope_bytes_per_point(m):
    """Algorithmic HBM bytes per transverse node and step of the envelope
    advance with m chi corrections per slice.  One transform: row pass reads
    a (re, im) 16 + chi 8 + d_{i+1} 16 (+ previous d_i 16 after the first) and
    writes two half spectra 16; column pass reads and writes them 32; last row
    pass reads them 16 and writes d_i 16.  The last transform of a slice also
    reads and writes a (32); the <a^2> pass reads a 16 and writes 8."""
    first = 16 + 8 + 16 + 16 + 32 + 16 + 16
    return first + m * (first + 16) + 32 + 24


def bench_envelope(capi, ctx, wl, steps, use_graph=1):
    """Envelope advance on the workload's grid: chi from one mixed wake pass with
    the attached Gaussian envelope, then `steps` steps (graph-replayed sweeps
    unless use_graph = 0)."""
    env = capi.Envelope(ctx, wl.nx, wl.ny, wl.nz, wl.dx, wl.dy, wl.dz, ENV_K0)
    env.set_gaussian(*wl.laser)
    wk = make_wake(capi, ctx, wl, capi.QP_PREC_MIXED, use_graph=use_graph)
    wk.attach_envelope(env)
    wk.run()
    wk.close()
    env.advance(ENV_DTAU, steps=1, chi_iterations=ENV_CHI_IT, use_graph=use_graph)   # graph capture
    ms = env.advance(ENV_DTAU, steps=steps, chi_iterations=ENV_CHI_IT, use_graph=use_graph) / steps
    # one coupled time step (wake pass + envelope step): two host calls, and the
    # step enqueued by the pass itself (qp_wake_set_envelope_step)
    wk = make_wake(capi, ctx, wl, capi.QP_PREC_MIXED, use_graph=use_graph)
    wk.attach_envelope(env)
    coupled = {}
    for mode in ("two_calls", "in_pass"):
        wk.set_envelope_step(ENV_DTAU if mode == "in_pass" else 0.0, ENV_CHI_IT)
        for rep in range(2):   # first: capture
            ctx.synchronize()
            t0 = time.perf_counter()
            wk.run()
            if mode == "two_calls":
                env.advance(ENV_DTAU, steps=1, chi_iterations=ENV_CHI_IT, use_graph=use_graph)
            ctx.synchronize()
            coupled[mode] = 1e3 * (time.perf_counter() - t0)
    wk.close()
    env.close()
    log(f"[bench] coupled wake pass + envelope step: two calls {coupled['two_calls']:.1f} ms, "
        f"in the pass {coupled['in_pass']:.1f} ms")
    pts = wl.nx * wl.ny * wl.nz
    byts = envelope_bytes_per_point(ENV_CHI_IT) * pts
    launches = wl.nx * 3 * (ENV_CHI_IT + 1) + 1
    achieved = byts / (ms * 1e-3) / 1e9
    res = {"k0": ENV_K0, "dtau": ENV_DTAU, "chi_iterations": ENV_CHI_IT, "graph": use_graph, "ms_per_step": ms,
           "slices_per_s": wl.nx / (ms * 1e-3), "launches_per_step": launches,
           "us_per_launch": 1e3 * ms / launches, "algorithmic_bytes_per_step": byts,
           "achieved_GBs": achieved, "frac_hbm": achieved / HBM_PEAK_GBS, "coupled_step_ms": coupled}
    log(f"[bench] envelope: {ms:.2f} ms/step ({launches} launches, {1e3 * ms / launches:.2f} us each), "
        f"{achieved:.0f} GB/s algorithmic")
    return res


# Schwinger field in plasma units at n0 = 1e18 cm^-3 (only scales the bunch's chi).
FIELD_SCHWINGER = 1.4e7


def cpu_info():
    """Host CPU model and logical core count (the CPU baseline's hardware)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model, os.cpu_count()


def class_profile(wk, classes):
    """Mean device duration per launch of each kernel class (hipEvents on the
    wake's stream around every launch of the class, one instrumented pass each)."""
    out = {}
    for cls in classes:
        n, us = wk.profile_kernel(cls)
        out[cls] = {"launches": n, "mean_us": us, "total_ms": n * us * 1e-3}
        log(f"[bench]   {cls}: {n} launches x {us:.2f} us = {n * us * 1e-3:.1f} ms")
    return out


def roofline_of(classes, K, n_particles, workload):
    """Roofline object of the push/deposit class with the largest total time:
    achieved = the reference-layout algorithmic bytes of one launch (the
    target's 404-B-per-particle-slice accounting) / its mean duration; the
    build's own-layout bytes are reported beside it, and `traffic` is the
    counter-measured HBM bytes per launch when a summary is committed."""
    dom = max(("push_A", "push_B", "push_C"), key=lambda c: classes[c]["total_ms"] if c in classes else -1)
    t = classes[dom]["mean_us"] * 1e-6
    spec = bytes_per_particle_spec(dom, K) * n_particles
    own = bytes_per_particle(dom, K) * n_particles
    ach = spec / t / 1e9
    return {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": read_traffic(dom, workload),
            "kernel": f"k_mx_{dom[-1]}", "algorithmic_bytes_per_launch": spec,
            "own_layout_bytes_per_launch": own, "achieved_own_layout": round(own / t / 1e9, 1),
            "frac_own_layout": round(own / t / 1e9 / HBM_PEAK_GBS, 4), "mean_us": classes[dom]["mean_us"]}


def solve_roofline(classes, n1, n2, K, bunch=False):
    """Bytes per transform launch (a third of a three-launch solve) / mean
    duration; beside it the whole solver class of a slice (3 + 6K launches) with
    its fused sources and sinks (solve_bytes_slice) / the class's time per slice."""
    if "solve" not in classes:
        return None
    t = classes["solve"]["mean_us"] * 1e-6
    sb = solve_bytes(n1, n2) / 3.0
    ach = sb / t / 1e9
    slice_b = solve_bytes_slice(n1, n2, K, bunch)
    slice_t = (3 + 6 * K) * t
    ach_s = slice_b / slice_t / 1e9
    return {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 4), "algorithmic_bytes_per_launch": sb,
            "mean_us": classes["solve"]["mean_us"], "launches": classes["solve"]["launches"],
            "with_sources": {"bytes_per_slice": slice_b, "us_per_slice": round(slice_t * 1e6, 1),
                             "achieved": round(ach_s, 1), "frac": round(ach_s / HBM_PEAK_GBS, 4)}}


def err_pair(got, ref):
    d = got - ref
    return [float(np.abs(d).max() / max(np.abs(ref).max(), 1e-300)),
            float(np.linalg.norm(d) / max(np.linalg.norm(ref), 1e-300))]


def leading_slices(capi, ctx, wl, precision, S, step, bunch=None, fields=("psi", "ex")):
    """The first S slices of one pass (slice_limit); keeps every step-th slice."""
    wk = make_wake(capi, ctx, wl, precision, slice_limit=S)
    if bunch is not None:
        wk.attach_bunch(bunch)
    keep = {f: {} for f in fields}
    nn = wl.ny * wl.nz

    def cb(v):
        if v.kind == 0 and (wl.nx - 1 - v.slice) % step == 0:
            for f in fields:
                keep[f][v.slice] = np.ctypeslib.as_array(getattr(v, f), shape=(nn,)).reshape(wl.ny, wl.nz).copy()
        return 0

    t0 = time.perf_counter()
    wk.run(callback=cb)
    t = time.perf_counter() - t0
    wk.close()
    order = sorted(keep[fields[0]], reverse=True)
    return {f: np.stack([keep[f][i] for i in order]) for f in fields}, order, t


def cpu_baseline(wl, S, bunch=None):
    """The oracle's OpenMP -Ofast build (the reference's loop structure and
    Release flags, own FFT: FFTW is absent) over the first S slices of the same
    workload, with windowed (slice-streamed) 3D arrays; timed on the host's
    cores.  Returns the baseline object and the oracle's psi / e_x slices."""
    os.environ.setdefault("OMP_PROC_BIND", "close")
    from oracle import oracle_ctypes as oc
    lib = oc.load("omp")
    g = oc.grid_from_box(wl.lx, wl.L, wl.L, wl.dx0, wl.L / wl.n0, wl.L / wl.n0, K=wl.K)
    y, z, w, py, pz = wl.particles()
    a = oc.gaussian_a_sqr_window(g, S, *wl.laser)
    rb = jb = None
    if bunch is not None:
        rb, jb = oc.bunch_deposit_window(g, bunch, g.nx - S, S)
    t0 = time.perf_counter()
    out, st, _ = oc.wake_run_window(g, y, z, w, a, S, py=py, pz=pz, rho_b=rb, jx_b=jb, fields=("psi", "ex"),
                                    variant="omp")
    t = time.perf_counter() - t0
    model, ncpu = cpu_info()
    np_ = len(y)
    res = {"value": np_ * S / t, "unit": "particle-slice pushes/s", "cores": int(lib.qpo_num_threads()),
           "kind": "port", "cpu_model": model, "host_logical_cpus": ncpu,
           "omp_proc_bind": os.environ.get("OMP_PROC_BIND"),
           "sample": f"first {S} of {g.nx} slices of the same workload ({np_} particles), slice-streamed "
                     f"3D arrays, FP64, OpenMP atomics, own mixed-radix FFT (FFTW absent), -Ofast; {t:.1f} s"}
    return res, {f: out[f][::-1] for f in ("psi", "ex")}


def timed_steps(ctx, steps, one_step):
    """W untimed steps are the caller's; here: synchronize, K steps, synchronize."""
    ctx.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        one_step(k)
    ctx.synchronize()
    return time.perf_counter() - t0


def bench_c5(args, capi, workloads):
    """C5: the beam-driven job.  A step is one beam time step of run()
    (src/System_3d.cpp:199-236) without QED: the mixed wake pass with the bunch
    streamed through it (per-slice sources and field gathers on the device),
    then the bunch push.  value = plasma particle-slice pushes / wall time."""
    wl = workloads.get("C5")
    bp = wl.bunch
    b0 = wl.bunch_particles()
    nb = len(b0["x"])
    log(f"[bench] workload {wl.name}: nx={wl.nx} n={wl.ny} particles={wl.n_particles} bunch={nb}")
    ctx = capi.Context(0)
    wk = make_wake(capi, ctx, wl, capi.QP_PREC_MIXED, use_graph=0 if args.no_graph else 1,
                   slice_limit=args.slice_limit)
    db = capi.Bunch(ctx, wl.nx, wl.ny, wl.nz, wl.dx, wl.dy, wl.dz, FIELD_SCHWINGER)
    wk.attach_bunch(db)
    db.set_particles(b0)
    dev_ms = []

    def step(k):
        wk.run()
        dev_ms.append(wk.last_time_ms())
        db.push(bp["dt"])
        log(f"[bench]   beam step {k}: pass {dev_ms[-1]:.1f} ms")

    for k in range(args.warmup):
        step(k)
    db.set_particles(b0)
    dev_ms.clear()
    t = timed_steps(ctx, args.steps, step)
    slices = args.slice_limit or wl.nx
    pushes = wl.n_particles * (slices - 1) * args.steps
    value = pushes / t
    final = db.get_particles()
    gamma_mean = float(np.mean(final["gamma"]))
    log(f"[bench] C5: {args.steps} beam steps in {t:.2f} s -> {value:.3e} pushes/s; bunch <gamma> {gamma_mean:.2f}")
    classes, roofline, sroof = {}, None, None
    if not args.no_profile:
        db.set_particles(b0)
        classes = class_profile(wk, ("push_A", "push_B", "push_C", "solve"))
        roofline = roofline_of(classes, wl.K, wl.n_particles, "C5")
        sroof = solve_roofline(classes, wl.ny, wl.nz, wl.K, bunch=True)
        log(f"[bench] roofline {roofline}")
    wk.close()
    err64 = fp64_rate = None
    mixed_lead = None
    if not args.no_fp64:
        S, step_k = 300, 3
        db.set_particles(b0)
        mixed_lead, order, _ = leading_slices(capi, ctx, wl, capi.QP_PREC_MIXED, S, step_k, bunch=db)
        db.set_particles(b0)
        f64, order64, t64 = leading_slices(capi, ctx, wl, capi.QP_PREC_FP64, S, step_k, bunch=db)
        assert order == order64
        err64 = {f: err_pair(mixed_lead[f], f64[f]) for f in ("psi", "ex")}
        fp64_rate = wl.n_particles * S / t64
        log(f"[bench] mixed vs fp64 GPU over the first {S} slices (max, rms rel): {err64}")
    cpu = err_cpu = None
    if not args.no_cpu:
        S = args.cpu_slices
        cpu, oref = cpu_baseline(wl, S, bunch=b0)
        if mixed_lead is not None:
            k = (S + 2) // 3   # the kept slices among the first S (every 3rd)
            ref = {f: oref[f][::3][:k] for f in ("psi", "ex")}
            if all(np.abs(ref[f]).max() == 0 and np.abs(mixed_lead[f][:k]).max() == 0 for f in ref):
                # the sample's slices lie ahead of the bunch's head (and of any
                # laser): the fields are zero on both sides and a relative error
                # says nothing; the oracle comparison through the bunch is the
                # C5 parity test (tests/test_gpu_configs.py)
                err_cpu = {"fields_zero_on_both_sides": True}
            else:
                err_cpu = {f: err_pair(mixed_lead[f][:k], ref[f]) for f in ref}
        log(f"[bench] cpu baseline: {cpu['value']:.3e} pushes/s on {cpu['cores']} threads ({cpu['cpu_model']}); "
            f"err vs oracle {err_cpu}")
    windows = None if args.no_profile else window_report(capi, ctx, wl, bunch=db, bunch_particles=b0)
    db.close()
    ctx.close()
    # the extras open their own contexts: HIP spreads streams over a few hardware
    # queues, and an idle C5 stream left open may share one with a C4 chain
    extra = {} if args.no_extra else bench_extras(args, capi, workloads)
    line = {
        "metric": "particle-slice pushes/s (mixed precision, 1 MI355X, <=1e-5 rel err in psi and Ez vs FP64)",
        "value": value, "unit": "particle-slice pushes/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * t / args.steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "storage": "plasma particles: u32 cell + u32x2 offsets, f64x2 momenta, f64 weight, f64x2 predicted momenta; "
                   "grids and solves f64; bunch f64 SoA",
        "config": {**wl.describe(), "bunch_particles": nb,
                   "step": "one beam time step: mixed wake pass with the bunch streamed through it + bunch push"},
        "solves_per_s": ((1 + 3 * wl.K) * (slices - 1) + 1) * args.steps / t,
        "device_ms_per_pass": float(np.median(dev_ms)) if dev_ms else None,
        "bunch_mean_gamma_after": gamma_mean,
        "kernel_classes": classes, "roofline": roofline, "solve_roofline": sroof,
        "fp64_gpu_pushes_per_s_leading": fp64_rate,
        "err_vs_fp64_gpu_leading": err64, "err_vs_oracle_sample": err_cpu,
        "cpu_baseline": cpu, "speedup_vs_cpu": (value / cpu["value"]) if cpu else None,
        "deposit_windows": windows,
        "other_workloads": extra,
        "note": "3D arrays of the reference's layout (281 GiB at this size) are never formed: the bunch "
                "sources and field gathers are streamed slice by slice",
    }
    print(json.dumps(line), flush=True)


def pass_rate(capi, ctx, wl, steps):
    """Whole-pass rate of one wake (mixed, hipGraph slice step): (ms, pushes/s)."""
    wk = make_wake(capi, ctx, wl, capi.QP_PREC_MIXED)
    wk.run()
    t = timed_steps(ctx, steps, lambda k: wk.run())
    wk.close()
    ms = 1e3 * t / steps
    return ms, wl.n_particles * (wl.nx - 1) / (ms * 1e-3)


def bench_extras(args, capi, workloads):
    """The smaller configurations beside the headline (no checks: their parity
    is in tests/test_gpu_configs.py)."""
    out = {}
    ctx = capi.Context(0)
    for name, steps in (("C2", 3), ("C3", 2)):
        wl = workloads.get(name)
        ms, rate = pass_rate(capi, ctx, wl, steps)
        out[name] = {"ms_per_pass": ms, "pushes_per_s": rate, "particles": wl.n_particles, "slices": wl.nx}
        log(f"[bench] {name}: {ms:.1f} ms per pass -> {rate:.3e} pushes/s")
        if name == "C3" and not args.no_profile:
            out[name]["deposit_windows"] = window_report(capi, ctx, wl)
    ctx.close()
    for chains in (C4_CHAINS, 1):
        ms, rate, desc = c4_pass(capi, workloads, 2, chains=chains)
        key = "C4" if chains == C4_CHAINS else f"C4_{chains}_chain"
        out[key] = {"ms_per_pass": ms, "pushes_per_s": rate, **desc}
        log(f"[bench] C4 ({chains} chains): {ms:.1f} ms per pass -> {rate:.3e} pushes/s")
    return out


def calib_copy(mib):
    """Device-to-device copy of a known size through the HIP runtime (its copy
    kernel), so FETCH_SIZE / WRITE_SIZE can be scaled to bytes."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    n = mib << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(n)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(n)) == 0
    assert hip.hipMemset(a, 1, ctypes.c_size_t(n)) == 0
    for _ in range(3):
        assert hip.hipMemcpy(b, a, ctypes.c_size_t(n), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)


def make_wake(capi, ctx, wl, precision, use_graph=1, async_output=0, slice_limit=0, resort_every=0):
    w = capi.Wake(ctx, wl.nx, wl.ny, wl.nz, wl.dx, wl.dy, wl.dz, wl.L, wl.L, K=wl.K, precision=precision,
                  use_graph=use_graph, resort_every=resort_every, async_output=async_output, slice_limit=slice_limit)
    y, z, wt, py, pz = wl.particles()
    w.set_particles(y, z, wt, py, pz)
    w.set_laser_gaussian(*wl.laser)
    return w


def window_report(capi, ctx, wl, resort_values=(0, 50), bunch=None, bunch_particles=None):
    """Deposit-window overflow (blocks that fall back to global atomics) over one
    whole pass, with the particles left in their initial cell order (0) and
    re-binned by tile every N slices; with the pass time."""
    out = []
    for r in resort_values:
        wk = make_wake(capi, ctx, wl, capi.QP_PREC_MIXED, resort_every=r)
        if bunch is not None:
            bunch.set_particles(bunch_particles)
            wk.attach_bunch(bunch)
        wk.run()
        ovf, tot = wk.window_stats()
        ms = wk.last_time_ms()
        wk.close()
        out.append({"resort_every": r, "overflow_blocks": ovf, "deposit_blocks": tot,
                    "overflow_fraction": ovf / tot if tot else 0.0, "pass_ms": ms})
        log(f"[bench] {wl.name[:2]} resort_every={r}: {ovf} of {tot} deposit blocks overflow the LDS window "
            f"({100.0 * ovf / max(tot, 1):.3f} %), pass {ms:.1f} ms")
    return out


class C4Job:
    """C4: 64 independent wakes.  The members of one grid size form a batch
    (qp_wake_create_batch: every launch of a batch carries all its members).
    The five batches are split into `chains` groups (qp_group_create: one
    launch per phase over all batches of the group), balanced by particle
    count; each group runs on its own context (stream) from its own host thread,
    so the groups' dependency chains overlap.  chains = 1: all 64 wakes in one
    launch per phase; chains = 5: one batch per stream."""

    def __init__(self, capi, workloads, use_graph=1, only_class=0, chains=1, slice_limit=0):
        members = workloads.c4_members()
        if only_class:  # one grid-size class only (profiling runs)
            members = [m for m in members if m.ny == only_class]
        self.members = members
        self.classes = {}
        for m in members:
            self.classes.setdefault(m.ny, []).append(m)
        self.sizes = sorted(self.classes)
        load = {n: sum(m.n_particles for m in self.classes[n]) for n in self.sizes}
        chains = max(1, min(chains, len(self.sizes)))
        bins = [[] for _ in range(chains)]
        for n in sorted(self.sizes, key=lambda n: -load[n]):   # largest first, into the lightest chain
            min(bins, key=lambda b: sum(load[k] for k in b)).append(n)
        self.chains = [sorted(b) for b in bins if b]
        self.batches = {}
        self.groups = []
        for chain in self.chains:
            ctx = capi.Context(0)
            for n in chain:
                ms = self.classes[n]
                m0 = ms[0]
                b = capi.WakeBatch(ctx, len(ms), m0.nx, m0.ny, m0.nz, m0.dx, m0.dy, m0.dz, m0.L, m0.L, K=m0.K,
                                   use_graph=use_graph, slice_limit=slice_limit)
                for i, m in enumerate(ms):
                    y, z, w, py, pz = m.particles()
                    b.set_particles(i, y, z, w, py, pz)
                    b.set_laser_gaussian(i, *m.laser)
                self.batches[n] = (ctx, b, sum(m.n_particles for m in ms))
                log(f"[bench] C4 class {n}x{n}: {len(ms)} members, {self.batches[n][2]} particles")
            if len(chain) > 1:
                g = capi.WakeGroup(ctx, [self.batches[n][1] for n in chain])
                g.set_fork(chains == 1)   # a second stream per group only pays when the group runs alone
                self.groups.append(g)
            else:
                self.groups.append(self.batches[chain[0]][1])
        slices = slice_limit if 0 < slice_limit < members[0].nx else members[0].nx
        self.pushes = sum(m.n_particles * (slices - 1) for m in members)

    def describe(self):
        return {"members": {f"{n}x{n}": len(self.classes[n]) for n in self.sizes},
                "particles": sum(m.n_particles for m in self.members), "slices": self.members[0].nx,
                "chains": [[f"{n}x{n}" for n in c] for c in self.chains]}

    def run(self):
        if len(self.groups) == 1:
            self.groups[0].run()
            return
        import threading
        errs = []

        def one(g):
            try:
                g.run()
            except Exception as e:  # surfaced below
                errs.append(e)

        ths = [threading.Thread(target=one, args=(g,)) for g in self.groups]
        for t in ths:
            t.start()
        for t in ths:
            t.join()
        if errs:
            raise errs[0]

    def synchronize(self):
        for (ctx, _, _) in self.batches.values():
            ctx.synchronize()

    def close(self):
        for g in self.groups:
            g.close()
        ctxs = []
        for (ctx, b, _) in self.batches.values():
            b.close()
            if ctx not in ctxs:
                ctxs.append(ctx)
        for ctx in ctxs:
            ctx.close()


def c4_pass(capi, workloads, steps, warmup=1, chains=C4_CHAINS):
    job = C4Job(capi, workloads, chains=chains)
    for _ in range(warmup):
        job.run()
    job.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        job.run()
    job.synchronize()
    ms = 1e3 * (time.perf_counter() - t0) / steps
    desc = job.describe()
    rate = job.pushes / (ms * 1e-3)
    job.close()
    return ms, rate, desc


def bench_c4(args, capi, workloads):
    """value = all members' particle-slice pushes of a pass / wall time."""
    job = C4Job(capi, workloads, use_graph=0 if args.no_graph else 1, only_class=args.c4_class, chains=args.c4_chains,
                slice_limit=args.slice_limit)
    for _ in range(args.warmup):
        job.run()
    job.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        job.run()
    job.synchronize()
    ms_per_step = 1e3 * (time.perf_counter() - t0) / args.steps
    value = job.pushes / (ms_per_step * 1e-3)
    log(f"[bench] C4: {ms_per_step:.1f} ms per pass of {len(job.members)} wakes = {value:.3e} pushes/s")
    alone = {}
    for n in (job.sizes if not args.no_profile else []):
        ctx, b, npart = job.batches[n]
        b.run()
        alone[f"{n}x{n}"] = {"members": len(job.classes[n]), "particles": npart, "ms": b.last_time_ms(),
                             "pushes_per_s": npart * (job.classes[n][0].nx - 1) / (b.last_time_ms() * 1e-3)}
    # roofline of the dominant push kernel in the largest class
    ctx, b, npart = job.batches[job.sizes[-1]]
    kc = {}
    roofline = None
    if not args.no_profile:
        kc = class_profile(b, ("push_A", "push_B", "push_C", "solve"))
        roofline = roofline_of(kc, job.members[0].K, npart, f"C4_{job.sizes[-1]}")
        roofline["class"] = f"{job.sizes[-1]}x{job.sizes[-1]}"
    err = cpu = None
    if not args.no_check:
        from oracle import oracle_ctypes as oc
        m = job.classes[job.sizes[0]][0]
        ctx, b, _ = job.batches[job.sizes[0]]
        outs, _, _ = b.run(fields=("psi", "ex"), members=[0])
        g = oc.grid_from_box(m.lx, m.L, m.L, m.dx0, m.L / m.n0, m.L / m.n0, K=m.K)
        y, z, w, py, pz = m.particles()
        a = oc.gaussian_a_sqr(g, *m.laser)
        os.environ.setdefault("OMP_PROC_BIND", "close")
        t0 = time.perf_counter()
        want, _, _ = oc.wake_run(g, y, z, w, a, py=py, pz=pz, fields=("psi", "ex"), variant="omp")
        tc = time.perf_counter() - t0
        err = {f: err_pair(outs[0][f], want[f]) for f in ("psi", "ex")}
        lib = oc.load("omp")
        model, ncpu = cpu_info()
        cpu = {"value": len(y) * (g.nx - 1) / tc, "unit": "particle-slice pushes/s",
               "cores": int(lib.qpo_num_threads()), "kind": "port", "cpu_model": model, "host_logical_cpus": ncpu,
               "sample": f"one whole C4 member ({m.ny}x{m.nz}, {len(y)} particles, {g.nx} slices), "
                         f"FP64, OpenMP atomics, -Ofast; {tc:.1f} s"}
        log(f"[bench] C4 member vs oracle: {err}")
    desc = job.describe()
    job.close()
    line = {
        "metric": "particle-slice pushes/s (mixed precision, 1 MI355X, <=1e-5 rel err in psi and Ez vs FP64)",
        "value": value, "unit": "particle-slice pushes/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "C4: 64 wakes over [-12,12]^2, grids from {64,96,128,192,256} (one batch per size, "
                               "the batches dealt into the groups of `chains`, each group on its own stream), "
                               "4 ppc, 800 slices dxi=0.025, a0~U[0.5,3], "
                               "w0~U[2,5], sigma_xi=2, transverse density up-ramp slope~U[0,0.1]", **desc},
        "classes_alone": alone, "kernel_classes_largest": kc, "roofline": roofline,
        "err_vs_oracle_member": err, "cpu_baseline": cpu,
        "speedup_vs_cpu": (value / cpu["value"]) if cpu else None,
    }
    print(json.dumps(line), flush=True)


def bench_single(args, capi, workloads):
    """C1-C3: one wake pass per step (hipGraph slice step)."""
    wl = workloads.get(args.workload)
    log(f"[bench] workload {wl.name}: nx={wl.nx} n={wl.ny} particles={wl.n_particles}")
    ctx = capi.Context(0)
    if args.envelope_only:
        res = bench_envelope(capi, ctx, wl, args.steps, use_graph=0 if args.no_graph else 1)
        print(json.dumps({"envelope": res, "config": wl.describe()}), flush=True)
        ctx.close()
        return
    wk = make_wake(capi, ctx, wl, capi.QP_PREC_MIXED, use_graph=0 if args.no_graph else 1,
                   slice_limit=args.slice_limit)
    for _ in range(args.warmup):
        wk.run()
    dev_ms = []

    def step(k):
        wk.run()
        dev_ms.append(wk.last_time_ms())

    t = timed_steps(ctx, args.steps, step)
    ms_per_step = 1e3 * t / args.steps
    slices = args.slice_limit or wl.nx
    value = wl.n_particles * (slices - 1) / (ms_per_step * 1e-3)
    log(f"[bench] mixed: {ms_per_step:.2f} ms/pass wall, device {np.median(dev_ms):.2f} ms -> {value:.3e} pushes/s")
    tuning = json.loads(wk.describe())
    classes, roofline, sroof = {}, None, None
    if not args.no_profile:
        classes = class_profile(wk, ("push_A", "push_B", "push_C", "solve"))
        roofline = roofline_of(classes, wl.K, wl.n_particles, args.workload)
        sroof = solve_roofline(classes, wl.ny, wl.nz, wl.K)
    wk.close()
    err64 = fp64_rate = None
    if not args.no_fp64:
        S = min(wl.nx, 400)
        m, order, _ = leading_slices(capi, ctx, wl, capi.QP_PREC_MIXED, S, 1)
        f64, _, t64 = leading_slices(capi, ctx, wl, capi.QP_PREC_FP64, S, 1)
        err64 = {f: err_pair(m[f], f64[f]) for f in ("psi", "ex")}
        fp64_rate = wl.n_particles * S / t64
        log(f"[bench] mixed vs fp64 GPU over the first {S} slices: {err64}")
    cpu = None
    if not args.no_cpu:
        cpu, _ = cpu_baseline(wl, min(args.cpu_slices, wl.nx))
        log(f"[bench] cpu baseline: {cpu['value']:.3e} pushes/s on {cpu['cores']} threads")
    ctx.close()
    line = {
        "metric": "particle-slice pushes/s (mixed precision, 1 MI355X, <=1e-5 rel err in psi and Ez vs FP64)",
        "value": value, "unit": "particle-slice pushes/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic", "config": wl.describe(),
        "solves_per_s": ((1 + 3 * wl.K) * (slices - 1) + 1) / (ms_per_step * 1e-3),
        "device_ms_per_step": float(np.median(dev_ms)), "kernel_classes": classes, "launch_shapes": tuning,
        "roofline": roofline, "solve_roofline": sroof, "fp64_gpu_pushes_per_s_leading": fp64_rate,
        "err_vs_fp64_gpu_leading": err64, "cpu_baseline": cpu,
        "speedup_vs_cpu": (value / cpu["value"]) if cpu else None,
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5, help="timed steps (C5: beam time steps; else passes)")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="C5", choices=("C1", "C2", "C3", "C4", "C5"))
    ap.add_argument("--cpu-slices", type=int, default=20, help="slices of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="C5: skip the C2 / C3 / C4 lines beside the headline")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch the slice step directly instead of replaying hipGraphs (for tracers)")
    ap.add_argument("--no-profile", action="store_true", help="skip the per-kernel-class event passes")
    ap.add_argument("--slice-limit", type=int, default=0,
                    help="run only the first N slices of every pass (profiling runs; value counts those)")
    ap.add_argument("--c4-chains", type=int, default=C4_CHAINS,
                    help="C4: groups (streams) the five grid-size batches are split into; 1 = one launch per "
                         "phase over all 64 wakes, 5 = one batch per stream")
    ap.add_argument("--c4-class", type=int, default=0,
                    help="C4 only: run the members of this grid size alone (profiling runs)")
    ap.add_argument("--envelope-only", action="store_true",
                    help="time only the laser-envelope advance on the workload's grid (profiling runs)")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="first copy this many MiB device-to-device (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "the slice loop has no multi-GPU path", file=sys.stderr)
        sys.exit(2)

    from quetzalpic_amd import capi, workloads
    if args.calib_copy_mib:
        calib_copy(args.calib_copy_mib)
    if args.workload == "C4":
        bench_c4(args, capi, workloads)
    elif args.workload == "C5":
        bench_c5(args, capi, workloads)
    else:
        bench_single(args, capi, workloads)


if __name__ == "__main__":
    main()
