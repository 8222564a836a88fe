#!/usr/bin/env python3
"""Benchmark of the quasistatic slice loop on one MI355X.

A step is one pass of the slice loop (System_3d::solve_wakefield) over the
workload's n_x slices: value = particle-slice pushes/s = N_p (n_x - 1) / t.
Default workload: C2 (256^2, 262144 particles, 1000 slices).  Prints one JSON
line on stdout (progress on stderr).
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
FIELD_NAMES_ALL = ("psi", "rho", "jx", "jy", "jz", "ex", "ey", "ez", "by", "bz", "chi")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def bytes_per_particle(cls, K):
    """Algorithmic HBM bytes per particle and launch of a particle phase
    (persistent state S = 36 B: u32 cell, 2 x u32 offsets, 2 x f64 momenta,
    f32 weight; scratch T = 8 B: 2 x f32 increments p* - p):
      A: read S + T, write position + momenta (28 B)          = 72
      B: read S (+ T after the first iteration), write T      = 44 | 52
      C: read S + T                                           = 44
      S: read S                                               = 36"""
    if cls == "push_A":
        return 72.0
    if cls == "push_B":
        return (44.0 + 52.0 * (K - 1)) / K if K > 0 else 44.0
    if cls == "push_C":
        return 44.0
    if cls == "push_S":
        return 36.0
    raise KeyError(cls)


def solve_bytes(n1, n2):
    """Algorithmic bytes of one three-launch solve: real source read + half
    spectrum write, spectrum read + write, spectrum read + real write."""
    nh = n2 // 2 + 1
    real = 8.0 * n1 * n2
    spec = 16.0 * n1 * nh
    return (real + spec) + 2 * spec + (spec + real)


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


def bench_bunch_job(args, capi, workloads):
    """C5: the whole beam-driven job, `steps` beam time steps of run()
    (src/System_3d.cpp:199-236) without QED: the mixed wake pass with the
    bunch streamed through it (per-slice sources and field gathers on the
    device), then the bunch push.  value = plasma pushes of all passes / time."""
    wl = workloads.get(args.workload)
    bp = wl.bunch
    steps = args.beam_steps or bp["steps"]
    b0 = wl.bunch_particles()
    nb = len(b0["x"])
    log(f"[bench] workload {wl.name}: nx={wl.nx} n={wl.ny} particles={wl.n_particles} bunch={nb}")
    ctx = capi.Context(0)
    wk = make_wake(capi, ctx, wl, capi.QP_PREC_MIXED, use_graph=0 if args.no_graph else 1)
    db = capi.Bunch(ctx, wl.nx, wl.ny, wl.nz, wl.dx, wl.dy, wl.dz, FIELD_SCHWINGER)
    wk.attach_bunch(db)
    db.set_particles(b0)
    for _ in range(args.warmup):     # one beam step, then back to the initial bunch
        wk.run()
        db.push(bp["dt"])
    db.set_particles(b0)
    ctx.synchronize()
    dev_ms = []
    t0 = time.perf_counter()
    for s_ in range(steps):
        _, st, stop = wk.run()
        dev_ms.append(wk.last_time_ms())
        db.push(bp["dt"])
        log(f"[bench]   beam step {s_}: pass {dev_ms[-1]:.1f} ms (status {st})")
    ctx.synchronize()
    t = time.perf_counter() - t0
    pushes = wl.n_particles * (wl.nx - 1) * steps
    value = pushes / t
    final = db.get_particles()
    log(f"[bench] C5: {steps} beam steps in {t:.2f} s -> {value:.3e} pushes/s; bunch <gamma> "
        f"{float(np.mean(final['gamma'])):.2f}")
    classes, roofline = {}, None
    if not args.no_profile:
        for cls in ("push_B", "push_C"):
            n, us = wk.profile_kernel(cls)
            classes[cls] = {"launches": n, "mean_us": us, "total_ms": n * us * 1e-3}
            log(f"[bench]   {cls}: {n} launches x {us:.2f} us")
        dom = max(classes, key=lambda c: classes[c]["total_ms"])
        byts = bytes_per_particle(dom, wl.K) * wl.n_particles
        ach = byts / (classes[dom]["mean_us"] * 1e-6) / 1e9
        roofline = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None, "kernel": f"k_mx_{dom[-1]}",
                    "algorithmic_bytes_per_launch": byts}
    wk.close()
    db.close()
    ctx.close()
    line = {
        "metric": "particle-slice pushes/s (mixed precision, 1 MI355X, beam-driven job)",
        "value": value, "unit": "particle-slice pushes/s", "n_gpus": 1, "steps": steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * t / steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic", "config": wl.describe(),
        "bunch_particles": nb, "device_ms_per_pass": float(np.median(dev_ms)),
        "kernel_classes": classes, "roofline": roofline, "cpu_baseline": None,
        "note": "3D arrays of the reference's layout (281 GiB at this size) are never formed: the bunch "
                "sources and field gathers are streamed slice by slice",
    }
    print(json.dumps(line), flush=True)


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


def make_wake(capi, ctx, wl, precision, use_graph=1, async_output=0):
    w = capi.Wake(ctx, wl.nx, wl.ny, wl.nz, wl.dx, wl.dy, wl.dz, wl.L, wl.L, K=wl.K, precision=precision,
                  use_graph=use_graph, resort_every=0, async_output=async_output)
    y, z, wt, py, pz = wl.particles()
    w.set_particles(y, z, wt, py, pz)
    w.set_laser_gaussian(*wl.laser)
    return w


def cpu_baseline(wl, budget_s, gpu_fields):
    """The oracle's OpenMP -Ofast build (the reference's loop structure and
    Release flags) on the first S slices of the same workload."""
    from oracle import oracle_ctypes as oc
    lib = oc.load("omp")
    g = oc.grid_from_box(wl.lx, wl.L, wl.L, wl.dx0, wl.L / wl.n0, wl.L / wl.n0, K=wl.K)
    y, z, w, py, pz = wl.particles()
    a = oc.gaussian_a_sqr(g, *wl.laser)
    # calibrate on 3 slices, then size the sample to the time budget
    t0 = time.perf_counter()
    oc.wake_run(g, y, z, w, a, py=py, pz=pz, n_slices=3, fields=(), variant="omp")
    per = (time.perf_counter() - t0) / 3
    s = int(max(3, min(g.nx - 1, budget_s / max(per, 1e-6))))
    t0 = time.perf_counter()
    out, _, _ = oc.wake_run(g, y, z, w, a, py=py, pz=pz, n_slices=s, fields=("psi", "ex"), variant="omp")
    t = time.perf_counter() - t0
    np_ = len(y)
    res = {"value": np_ * s / t, "unit": "particle-slice pushes/s", "cores": int(lib.qpo_num_threads()),
           "kind": "port",
           "sample": f"first {s} of {g.nx} slices of the same workload ({np_} particles), FP64, "
                     f"OpenMP atomics, own mixed-radix FFT, -Ofast; {t:.1f} s"}
    errs = {}
    if gpu_fields is not None:
        lo = g.nx - s
        for f in ("psi", "ex"):
            ref = out[f][lo:g.nx]
            got = gpu_fields[f][lo:g.nx]
            if f == "ex":  # the last slice of ex is zero in both
                ref, got = ref[:-1], got[:-1]
            d = got - ref
            errs[f] = [float(np.abs(d).max() / max(np.abs(ref).max(), 1e-300)),
                       float(np.linalg.norm(d) / max(np.linalg.norm(ref), 1e-300))]
    return res, errs


def bench_c4(args, capi, workloads):
    """C4: 64 independent wakes.  The members of one grid size form a batch
    (qp_wake_create_batch: every launch of a batch carries all its members); the
    batches run concurrently, each on its own context (stream) from its own host
    thread.  value = all members' particle-slice pushes of a pass / wall time."""
    import threading
    members = workloads.c4_members()
    if args.c4_class:  # one grid-size class only (profiling runs)
        members = [m for m in members if m.ny == args.c4_class]
    classes = {}
    for m in members:
        classes.setdefault(m.ny, []).append(m)
    sizes = sorted(classes)
    batches = {}
    for n in sizes:
        ms = classes[n]
        m0 = ms[0]
        ctx = capi.Context(0)
        b = capi.WakeBatch(ctx, len(ms), m0.nx, m0.ny, m0.nz, m0.dx, m0.dy, m0.dz, m0.L, m0.L, K=m0.K,
                           use_graph=0 if args.no_graph else 1)
        for i, m in enumerate(ms):
            y, z, w, py, pz = m.particles()
            b.set_particles(i, y, z, w, py, pz)
            b.set_laser_gaussian(i, *m.laser)
        batches[n] = (ctx, b, sum(m.n_particles for m in ms))
        log(f"[bench] C4 class {n}x{n}: {len(ms)} members, {batches[n][2]} particles")
    pushes = sum(m.n_particles * m.nx for m in members)

    def run_all():
        errs = []

        def one(b):
            try:
                b.run()
            except Exception as e:  # surfaced below
                errs.append(e)

        ths = [threading.Thread(target=one, args=(b,)) for (_, b, _) in batches.values()]
        for t in ths:
            t.start()
        for t in ths:
            t.join()
        if errs:
            raise errs[0]

    for _ in range(args.warmup):
        run_all()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        run_all()
    t1 = time.perf_counter()
    ms_per_step = 1e3 * (t1 - t0) / args.steps
    value = pushes / (ms_per_step * 1e-3)
    log(f"[bench] C4: {ms_per_step:.1f} ms per pass of 64 wakes = {value:.3e} pushes/s")
    # each class alone (device time of one pass)
    alone = {}
    for n in (sizes if not args.no_profile else []):
        ctx, b, npart = batches[n]
        b.run()
        alone[f"{n}x{n}"] = {"members": len(classes[n]), "particles": npart, "ms": b.last_time_ms(),
                             "pushes_per_s": npart * classes[n][0].nx / (b.last_time_ms() * 1e-3)}
    # roofline of the dominant push kernel in the largest class
    ctx, b, npart = batches[sizes[-1]]
    kc = {}
    for cls in (("push_A", "push_B", "push_C", "solve") if not args.no_profile else ()):
        nl, us = b.profile_kernel(cls)
        kc[cls] = {"launches": nl, "mean_us": us, "total_ms": nl * us * 1e-3}
    roofline = None
    if kc:
        dom = max(("push_A", "push_B", "push_C"), key=lambda c: kc[c]["total_ms"])
        dom_bytes = bytes_per_particle(dom, members[0].K) * npart
        achieved = dom_bytes / (kc[dom]["mean_us"] * 1e-6) / 1e9
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "kernel": f"k_mx_{dom[-1]}",
                    "algorithmic_bytes_per_launch": dom_bytes, "class": f"{sizes[-1]}x{sizes[-1]}"}
    # parity: the first member of the smallest class against the FP64 oracle, whole pass
    err = None
    cpu = None
    if not args.no_check:
        from oracle import oracle_ctypes as oc
        m = classes[sizes[0]][0]
        ctx, b, _ = batches[sizes[0]]
        outs, _, _ = b.run(fields=("psi", "ex"))
        g = oc.grid_from_box(m.lx, m.L, m.L, m.dx0, m.L / m.n0, m.L / m.n0, K=m.K)
        y, z, w, py, pz = m.particles()
        a = oc.gaussian_a_sqr(g, *m.laser)
        t0 = time.perf_counter()
        want, _, _ = oc.wake_run(g, y, z, w, a, py=py, pz=pz, fields=("psi", "ex"), variant="omp")
        tc = time.perf_counter() - t0
        err = {}
        for f in ("psi", "ex"):
            d = outs[0][f] - want[f]
            err[f] = [float(np.abs(d).max() / np.abs(want[f]).max()),
                      float(np.linalg.norm(d) / np.linalg.norm(want[f]))]
        lib = oc.load("omp")
        cpu = {"value": len(y) * g.nx / tc, "unit": "particle-slice pushes/s", "cores": int(lib.qpo_num_threads()),
               "kind": "port", "sample": f"one whole C4 member ({m.ny}x{m.nz}, {len(y)} particles, {g.nx} slices), "
                                         f"FP64, OpenMP atomics, -Ofast; {tc:.1f} s"}
        log(f"[bench] C4 member vs oracle: {err}")
    for (ctx, b, _) in batches.values():
        b.close()
        ctx.close()
    line = {
        "metric": "particle-slice pushes/s (mixed precision, 1 MI355X, <=1e-5 rel err in psi and Ez vs FP64)",
        "value": value, "unit": "particle-slice pushes/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "C4: 64 wakes over [-12,12]^2, grids from {64,96,128,192,256} (one batch per size, "
                               "batches on concurrent streams), 4 ppc, 800 slices dxi=0.025, a0~U[0.5,3], "
                               "w0~U[2,5], sigma_xi=2, transverse density up-ramp slope~U[0,0.1]",
                   "members": {f"{n}x{n}": len(classes[n]) for n in sizes}, "particles": sum(
                       m.n_particles for m in members), "slices": members[0].nx},
        "classes_alone": alone, "kernel_classes_largest": kc, "roofline": roofline,
        "err_vs_oracle_member": err, "cpu_baseline": cpu,
        "speedup_vs_cpu": (value / cpu["value"]) if cpu else None,
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="C2")
    ap.add_argument("--cpu-budget", type=float, default=15.0, help="seconds of CPU baseline work")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch the slice step directly instead of replaying hipGraphs (for tracers)")
    ap.add_argument("--no-profile", action="store_true", help="skip the per-kernel-class event pass")
    ap.add_argument("--c4-class", type=int, default=0,
                    help="C4 only: run the members of this grid size alone (profiling runs)")
    ap.add_argument("--beam-steps", type=int, default=0,
                    help="C5: beam time steps to time (default: the configuration's 10; fewer for profiling)")
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
        return
    if args.workload == "C5":
        bench_bunch_job(args, capi, workloads)
        return
    wl = workloads.get(args.workload)
    log(f"[bench] workload {wl.name}: nx={wl.nx} n={wl.ny} particles={wl.n_particles}")
    ctx = capi.Context(0)
    if args.envelope_only:
        res = bench_envelope(capi, ctx, wl, args.steps, use_graph=0 if args.no_graph else 1)
        print(json.dumps({"envelope": res, "config": wl.describe()}), flush=True)
        ctx.close()
        return

    # ---- timed region: the mixed fused path, hipGraph slice step
    wk = make_wake(capi, ctx, wl, capi.QP_PREC_MIXED, use_graph=0 if args.no_graph else 1)
    for _ in range(args.warmup):
        wk.run()
    ctx.synchronize()
    dev_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        wk.run()
        dev_ms.append(wk.last_time_ms())
    ctx.synchronize()
    t1 = time.perf_counter()
    ms_per_step = 1e3 * (t1 - t0) / args.steps
    pushes = wl.n_particles * (wl.nx - 1)
    solves = (1 + 3 * wl.K) * (wl.nx - 1) + 1
    value = pushes / (ms_per_step * 1e-3)
    log(f"[bench] mixed: {ms_per_step:.2f} ms/pass wall, device {np.median(dev_ms):.2f} ms -> {value:.3e} pushes/s")
    tuning = json.loads(wk.describe())
    log(f"[bench] launch shapes: {tuning}")

    # ---- per-kernel-class device times (separate instrumented pass, same work):
    # hipEvents on the wake's stream around every launch of the class
    classes = {}
    roofline = None
    if not args.no_profile:
        for cls in ("push_A", "push_B", "push_C", "solve"):
            n, us = wk.profile_kernel(cls)
            classes[cls] = {"launches": n, "mean_us": us, "total_ms": n * us * 1e-3}
            log(f"[bench]   {cls}: {n} launches x {us:.2f} us = {n * us * 1e-3:.2f} ms")
        # push/deposit kernel with the largest total time (the particle streaming the metric counts)
        dom = max(("push_A", "push_B", "push_C"), key=lambda c: classes[c]["total_ms"])
        dom_bytes = bytes_per_particle(dom, wl.K) * wl.n_particles
        achieved = dom_bytes / (classes[dom]["mean_us"] * 1e-6) / 1e9
        traffic = read_traffic(dom, args.workload)
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": f"k_mx_{dom[-1]}",
                    "algorithmic_bytes_per_launch": dom_bytes}
        sb = solve_bytes(wl.ny, wl.nz) / 3.0
        classes["solve"]["algorithmic_bytes_per_launch"] = sb
        classes["solve"]["achieved_GBs"] = sb / (classes["solve"]["mean_us"] * 1e-6) / 1e9

    # ---- plain FP64 GPU path (one kernel per reference loop) on the same workload
    fp64 = None
    fields64 = None
    if not args.no_fp64:
        wf = make_wake(capi, ctx, wl, capi.QP_PREC_FP64)
        wf.run()
        t0 = time.perf_counter()
        wf.run()
        t64 = time.perf_counter() - t0
        fp64 = pushes / t64
        log(f"[bench] fp64 plain: {t64 * 1e3:.1f} ms/pass -> {fp64:.3e} pushes/s")
        if not args.no_check:
            fields64, _, _ = wf.run(fields=("psi", "ex"))
        wf.close()

    # ---- error of the mixed path against the FP64 path over the whole pass
    err64 = None
    fieldsmx = None
    if not args.no_check:
        fieldsmx, _, _ = wk.run(fields=("psi", "ex"))
        if fields64 is not None:
            err64 = {}
            for f in ("psi", "ex"):
                d = fieldsmx[f] - fields64[f]
                err64[f] = [float(np.abs(d).max() / np.abs(fields64[f]).max()),
                            float(np.linalg.norm(d) / np.linalg.norm(fields64[f]))]
            log(f"[bench] mixed vs fp64 GPU (max, rms rel): {err64}")
    wk.close()

    # ---- passes with every slice's eleven output planes handed to a host
    # callback (outside the timed region): inline vs streamed
    output_pass = None
    if not args.no_check:
        output_pass = {}
        for mode in (0, 1):
            wo = make_wake(capi, ctx, wl, capi.QP_PREC_MIXED, async_output=mode)
            wo.run(fields=FIELD_NAMES_ALL)
            t0 = time.perf_counter()
            wo.run(fields=FIELD_NAMES_ALL)
            output_pass["async" if mode else "inline"] = 1e3 * (time.perf_counter() - t0)
            wo.close()
        log(f"[bench] pass with all planes to a host callback: inline {output_pass['inline']:.1f} ms, "
            f"streamed {output_pass['async']:.1f} ms")

    # ---- laser-envelope advance on the same grid (outside the timed region)
    envelope = bench_envelope(capi, ctx, wl, max(args.steps, 3), use_graph=0 if args.no_graph else 1)

    cpu = None
    err_oracle = None
    if not args.no_cpu:
        cpu, err_oracle = cpu_baseline(wl, args.cpu_budget, fieldsmx)
        log(f"[bench] cpu baseline: {cpu['value']:.3e} pushes/s on {cpu['cores']} threads; err vs oracle "
            f"{err_oracle}")
    ctx.close()

    line = {
        "metric": "particle-slice pushes/s (mixed precision, 1 MI355X, <=1e-5 rel err in psi and Ez vs FP64)",
        "value": value,
        "unit": "particle-slice pushes/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "storage": "particles: u32 cell + u32x2 offsets, f64x2 momenta, f32 weight, f32x2 p* increments; grids and solves f64",
        "data": "synthetic",
        "config": wl.describe(),
        "solves_per_s": solves / (ms_per_step * 1e-3),
        "device_ms_per_step": float(np.median(dev_ms)),
        "kernel_classes": classes,
        "launch_shapes": tuning,
        "roofline": roofline,
        "fp64_gpu_pushes_per_s": fp64,
        "speedup_vs_fp64_gpu": (value / fp64) if fp64 else None,
        "err_vs_fp64_gpu": err64,
        "err_vs_oracle_sample": err_oracle,
        "cpu_baseline": cpu,
        "speedup_vs_cpu": (value / cpu["value"]) if cpu else None,
        "envelope_advance": envelope,
        "output_pass_ms": output_pass,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
