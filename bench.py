#!/usr/bin/env python3
"""Headline benchmark: RK4 grid-point-steps/s of the pseudospectral step at
4096^2 FP64 (N=4096, alpha=2, dt=5e-5, nu=1e-28 k^8, mu=0.02, ring (63,65],
initial field from seed 2 with spectrum k^4 exp(-(k/32)^2)).

A step is what the reference's driver loop does per iteration when no
diagnostics or output are due (src/main.f90:97-137): step4 + forcing, on a
state resident in HBM.  W warm-up steps, then K timed steps bracketed by
device synchronisation.

The roofline entry is the kernel class with the largest time per step,
priced at its algorithmic bytes: the bytes the pass must move on the
compact layout the kernels keep between boundary crossings (every spectral
field of the step is zero outside the 2/3 rectangle, so only that
rectangle is read and written; `compact_bytes`), over its average duration
measured live over the timed region with HIP events on the library's own
stream (only that class is event-timed there: every timed launch records
two events, which costs a step a few per cent).  Beside it: the HBM
traffic the counters measured for that kernel (FETCH_SIZE + WRITE_SIZE,
scaled by a plain copy, from the committed counter summary) as
`frac_measured`, and the survey's unfused byte model (A = 8 N^2 per field
pass, 468 B per grid-point-step; `model_bytes`) as `model_*`, which counts
bytes the fused kernels never move and so is not a rate of moved bytes.
The per-class breakdown ("kernels") comes from a separate sample of
--breakdown-steps steps that times every class; a rate of moved bytes
above the HBM peak stops the benchmark with an error.

The CPU baseline is the oracle (plain C, OpenMP) timed on this host on
bounded samples of the same workload, at every core the process may use
and at one core, with the 256^2 Euler workload (1000 steps) beside it; the
reference reports the mean wall time per iteration (src/main.f90:190-193).

Also measured, in the same line: the RK2 variant of the headline workload
("rk2"); the 256^2 Euler and 1024^2 SQG workloads' step loops ("configs");
the batched ensemble ("ensemble": 32 x 256^2, 24 x 512^2, 8 x 1024^2,
alpha cycling through 1/2, 1, 2) against the plain per-run loop of the same
members, both over the same number of steps; and the cost of one adjoint
gradient against the plain forward run ("adjoint": N=512, alpha=1, 200 RK4
steps, checkpoints every 20).

Prints ONE JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "RK4 grid-pt-steps/s at 4096² FP64; ensemble & adjoint vs plain loop"
HBM_PEAK_GBS = 8000.0


def model_bytes(n, rko):
    """Algorithmic HBM bytes per launch of each kernel class in the fixed byte
    model of the path (A = 8 N^2 bytes, one pass over an FP64 half-spectrum
    field): per nonlinear term 11A -- the inverse y transforms read q (1A)
    and write q, u, v (3A), the x pass reads 3A and writes 2A, the forward
    y transforms read 2A -- plus the RK4 epilogues 14.5A per step (stage 1
    3.5A, stages 2 and 3 4.5A each, the final update 2A); RK2: 27A per step.
    Distributed over the fused kernels: x pass 5A; y start 4A; y stage
    2A + epilogue + 4A; y final 2A + epilogue.  RK4: 58.5A = 468 B per
    grid point and step."""
    a = 8.0 * n * n
    if rko == 4:
        stage = 2 * a + (3.5 + 4.5 + 4.5) / 3 * a + 4 * a  # stages 1-3 averaged
        final = 2 * a + 2 * a
    else:  # RK2 27A: y start 4A, 2 x 5A, stage 1 2A + 2.5A + 4A, final 2A + 2.5A
        stage = 2 * a + 2.5 * a + 4 * a
        final = 2 * a + 2.5 * a
    # the fused step boundary (final stage + forcing + next y start)
    return {"y_start": 4 * a, "y_stage": stage, "y_final": final, "x_pass": 5 * a,
            "y_bound": final + 4 * a}


def compact_bytes(n, rect, rko):
    """HBM bytes each kernel class must move per launch on the compact layout
    the kernels use (a kept mode is 16 B; a y-buffer field holds N*kc complex
    values; emk is 8 B per kept mode) -- what the measured traffic is checked
    against."""
    kx, p, q = rect
    kc, ky = kx + 1, p + 1 + q
    Z = 16.0 * kc * ky          # one compact spectral field
    E = 8.0 * kc * ky           # the integrating-factor table
    Y = 16.0 * n * kc           # one field between the passes
    if rko == 4:
        stage = 5 * Y + E + (2 * Z + 3 * Z + 3 * Z) / 3.0  # stages 1-3 averaged
        final = 2 * Y + 2 * Z
    else:
        stage = 5 * Y + E + Z
        final = 2 * Y + E + 2 * Z
    # fused step boundary: the final stage's reads, z written once (forced),
    # the next stage input's three fields written
    bound = final + 3 * Y
    return {"y_start": Z + 3 * Y, "y_stage": stage, "y_final": final,
            "x_pass": 5 * Y, "y_bound": bound}


def read_traffic(kernel):
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE and
    WRITE_SIZE), from the committed summary, or None."""
    path = os.path.join(ROOT, "profiles", "pmc_euler4096_rk4.txt")
    if not os.path.exists(path):
        return None
    vals = {}
    with open(path) as f:
        for line in f:
            parts = line.split()
            if len(parts) >= 3 and parts[0] == "traffic_bytes_per_launch":
                vals[parts[1]] = float(parts[2])
    return vals.get(kernel)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_threads():
    """The cores this process may use: OMP_NUM_THREADS when set (the host's
    share of cores for one GPU), else the CPU affinity set."""
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        return int(env)
    return len(os.sched_getaffinity(0))


def cpu_rate(p, w0, seed, threads, budget_s, max_steps):
    """Oracle step loop (step + forcing, no diagnostics or output) on
    `threads` OpenMP threads: steps run until max_steps or until budget_s
    has passed (at least one)."""
    import oracle
    from gturbo_amd import io
    oracle.set_threads(threads)
    o = oracle.Oracle(p)
    z = io.field_to_ref(w0)
    work = [o.field() for _ in range(4)]
    o.fft_dir(z)
    o.trunc(z)
    steps = 0
    t0 = time.perf_counter()
    while steps < max_steps:
        if p.rko == 4:
            o.step4(z, *work)
        else:
            o.step(z, work[0], work[2], work[3])
        seed = o.forcing(z, seed)
        steps += 1
        if time.perf_counter() - t0 > budget_s:
            break
    el = time.perf_counter() - t0
    return {"value": p.n * p.n * steps / el, "unit": "grid-pt-steps/s", "cores": threads,
            "ms_per_step": el / steps * 1e3,
            "sample": f"{steps} step(s)+forcing of {p.n}^2 RK{p.rko} ({el:.1f} s wall, "
                      f"{threads} OpenMP thread{'s' if threads > 1 else ''})"}


def cpu_baseline(p, w0, seed):
    """The oracle on this host's cores: the headline workload at every core
    this process may use (`value`) and at one core, and the 256^2 Euler
    workload (all 1000 steps at every core; a bounded sample at one core)."""
    import oracle
    from gturbo_amd import ic, workloads
    oracle.build()
    nt = cpu_threads()
    full = cpu_rate(p, w0, seed, nt, 8.0, 40)
    one = cpu_rate(p, w0, seed, 1, 6.0, 40)
    w = workloads()["euler256"]
    w256 = ic.physical_field(w.params, w.ic)
    e_all = cpu_rate(w.params, w256, w.forcing_seed, nt, 15.0, w.params.npas)
    e_one = cpu_rate(w.params, w256, w.forcing_seed, 1, 5.0, w.params.npas)
    oracle.set_threads(nt)
    return {"value": full["value"], "unit": "grid-pt-steps/s", "cores": nt, "kind": "port",
            "sample": full["sample"], "ms_per_step": full["ms_per_step"],
            "nproc": os.cpu_count(), "affinity_cpus": len(os.sched_getaffinity(0)),
            "cpu_model": cpu_model(),
            "one_core": one,
            "euler256": {"workload": w.name, "all_cores": e_all, "one_core": e_one}}


def workload_rate(name, steps, warmup, n=None):
    """grid-pt-steps/s of the step loop (step or step4, then forcing; no
    diagnostics or output) of a named workload on the GPU."""
    from gturbo_amd import ic, io, workloads
    from gturbo_amd.solver import Solver
    w = workloads()[name]
    p = w.params if n is None else w.params.replace(n=n)
    with Solver(p) as s:
        z = io.field_to_ref(ic.physical_field(p, w.ic))
        arrs = [z] + [s.field() for _ in range(4)]
        s.enter_data(z)
        for a in arrs[1:]:
            s.enter_data(a, copyin=False)
        s.fft_dir(z)
        s.trunc(z)
        seed = w.forcing_seed
        for it in range(warmup + steps):
            if it == warmup:
                s.synchronize()
                t0 = time.perf_counter()
            if p.rko == 4:
                s.step4(*arrs)
            else:
                s.step(arrs[0], arrs[1], arrs[3], arrs[4])
            seed = s.forcing(z, seed)
        s.synchronize()
        el = time.perf_counter() - t0
        s.update_host(z)
        ok = bool(np.all(np.isfinite(z)))
    return {"workload": w.name, "value": p.n * p.n * steps / el, "unit": "grid-pt-steps/s",
            "ms_per_step": el / steps * 1e3, "steps": steps, "finite": ok}


def ensemble_bench(steps):
    """Batched ensemble (tuned schedule, graph-captured steps) against the
    plain per-run loop (each member alone: step4 + forcing per step through
    the single-run entry points), same members, same number of steps."""
    from gturbo_amd import ensemble as E, ic
    from gturbo_amd.params import ensemble_members
    from gturbo_amd.solver import Solver
    wl = ensemble_members()
    fields = [ic.physical_field(w.params, w.ic) for w in wl]
    ens = E.Ensemble([w.params for w in wl], seeds=[w.forcing_seed for w in wl])
    for m, f in enumerate(fields):
        ens.set_field(m, f)
    cfg, _ = ens.tune(3)
    ens.run(10)
    ens.synchronize()
    t0 = time.perf_counter()
    ens.run(steps)
    ens.synchronize()
    tb = time.perf_counter() - t0
    finite = all(bool(np.all(np.isfinite(ens.get_field(m, E.SPECTRAL)))) for m in range(len(wl)))
    # the same schedule on persistent grids (workgroups looping over the work
    # items, 1/2/4 waves per SIMD) against the tuned one
    persistent = {}
    for wps in (1, 2, 4):
        ens.schedule = [cfg[0], cfg[1], wps] + list(cfg[3:])
        ens.run(5)
        ens.synchronize()
        t1 = time.perf_counter()
        ens.run(max(steps // 5, 10))
        ens.synchronize()
        persistent[f"{wps}_waves_per_simd_ms_per_step"] = \
            (time.perf_counter() - t1) / max(steps // 5, 10) * 1e3
    ens.close()
    tp = 0.0
    for w, f in zip(wl, fields):
        p = w.params
        with Solver(p) as s:
            z = np.zeros((p.n, p.n + 2))
            z[:, :p.n] = f
            arrs = [z] + [np.zeros_like(z) for _ in range(4)]
            for a in arrs:
                s.enter_data(a)
            s.fft_dir(z)
            s.trunc(z)
            seed = w.forcing_seed
            for it in range(3 + steps):
                if it == 3:
                    s.synchronize()
                    t0 = time.perf_counter()
                s.step4(*arrs)
                seed = s.forcing(z, seed)
            s.synchronize()
            tp += time.perf_counter() - t0
            for a in arrs:
                s.exit_data(a, copyout=False)
    gp = sum(w.params.n ** 2 for w in wl)
    sizes = {}
    for w in wl:
        sizes[w.params.n] = sizes.get(w.params.n, 0) + 1
    return {"workload": "+".join(f"{c}x{n}^2" for n, c in sorted(sizes.items())) +
                        " RK4, alpha 1/2,1,2, forced, dt=0.25/N",
            "members": len(wl), "steps": steps, "schedule": cfg,
            "batched_ms_per_step": tb / steps * 1e3,
            "persistent_grid": persistent,
            "plain_loop_ms_per_step": tp / steps * 1e3,
            "speedup": tp / tb,
            "batched_grid_pt_steps_per_s": gp * steps / tb,
            "finite": finite}


def diag_bench(sizes=(1024, 4096), reps=20):
    """Wall time per call of the diagnostics entry points (spectrum, fluxes,
    glob; src/diagnostics.f90) on the spectral state of the 1024^2 SQG and
    4096^2 Euler workloads after fft_dir + trunc, results copied to the
    host; the nonlinear term the fluxes need is computed once, outside the
    timing."""
    from gturbo_amd import ic, io, workloads
    from gturbo_amd.solver import Solver
    names = {1024: "sqg1024", 4096: "euler4096_rk4"}
    out = {}
    for n in sizes:
        w = workloads()[names[n]]
        p = w.params
        with Solver(p) as s:
            z = io.field_to_ref(ic.physical_field(p, w.ic))
            znl, u, v = s.field(), s.field(), s.field()
            for a in (z, znl, u, v):
                s.enter_data(a)
            s.fft_dir(z)
            s.trunc(z)
            s.copy(znl, z)
            s.nlt(znl, u, v)
            res = {}
            for name, call in (("spectrum", lambda: s.spectrum(z)),
                               ("fluxes", lambda: s.fluxes(z, znl)),
                               ("glob", lambda: s.glob(z))):
                call()  # first call allocates the partial sums
                ts = []
                for _ in range(reps):
                    t0 = time.perf_counter()
                    call()
                    ts.append(time.perf_counter() - t0)
                res[name + "_ms"] = float(np.median(ts)) * 1e3
            out[f"{n}"] = {"workload": w.name, **res}
    return out


def adjoint_bench():
    """The adjoint workload (N=512, alpha=1, RK4, dt=5e-4, 200 steps, forcing
    off, checkpoints every 20 steps): time of gt_adj_forward + gt_adj_backward
    (one gradient) against the plain forward loop of the same steps."""
    from gturbo_amd import ic, workloads
    from gturbo_amd.adjoint import Adjoint
    from gturbo_amd.solver import Solver
    w = workloads()["adjoint512"]
    p = w.params
    n, steps = p.n, p.npas
    w0 = ic.physical_field(p, w.ic).astype(np.float64)
    with Solver(p) as s:
        z = np.zeros((n, n + 2))
        z[:, :n] = w0
        arrs = [z] + [s.field() for _ in range(4)]
        s.enter_data(z)
        for a in arrs[1:]:
            s.enter_data(a, copyin=False)
        s.fft_dir(z)
        s.trunc(z)
        seed = w.forcing_seed
        for it in range(5 + steps):
            if it == 5:
                s.synchronize()
                t0 = time.perf_counter()
            s.step4(*arrs)
            seed = s.forcing(z, seed)
        s.synchronize()
        t_plain = time.perf_counter() - t0
    with Adjoint(p, steps, kchk=20) as adj:
        adj.forward(w0, w.forcing_seed)  # warm-up (allocations)
        adj.backward(np.zeros((n, n)))
        t0 = time.perf_counter()
        wT, _ = adj.forward(w0, w.forcing_seed)
        t1 = time.perf_counter()
        g = adj.backward(wT)
        t2 = time.perf_counter()
    return {"workload": w.name + ", checkpoints every 20 steps",
            "plain_forward_ms": t_plain * 1e3, "adjoint_forward_ms": (t1 - t0) * 1e3,
            "adjoint_backward_ms": (t2 - t1) * 1e3,
            "gradient_cost_vs_forward": (t2 - t0) / t_plain,
            "finite": bool(np.all(np.isfinite(g)))}


def progress(msg):
    """One line per finished leg, on stderr (stdout carries the JSON line)."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--breakdown-steps", type=int, default=10,
                    help="steps of the separate sample that times every kernel class")
    ap.add_argument("--rko", type=int, default=4, choices=(2, 4))
    ap.add_argument("--n", type=int, default=4096)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--rk2-steps", type=int, default=50,
                    help="steps of the RK2 variant of the headline workload (0: skip)")
    ap.add_argument("--ens-steps", type=int, default=100,
                    help="steps of the batched-ensemble comparison (0: skip)")
    ap.add_argument("--no-configs", dest="config_steps", action="store_false",
                    help="skip the 256^2 and 1024^2 workloads' step loops")
    ap.add_argument("--no-adjoint", action="store_true",
                    help="skip the adjoint cost measurement")
    ap.add_argument("--calib-copy", action="store_true",
                    help="also run one device-to-device copy of a reference-layout "
                         "field (a plain copy of known size for HBM counter scaling)")
    ap.add_argument("--diag-only", action="store_true",
                    help="only time the diagnostics entry points at 1024^2 and 4096^2")
    args = ap.parse_args()
    if args.diag_only and args.gpus == 1:
        print(json.dumps({"diagnostics": diag_bench()}))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "gturbo_amd runs on one GPU: the 2D pseudospectral "
                          "step is not sharded across GPUs; "
                          f"--gpus {args.gpus} is not supported"}))
        sys.exit(2)

    from gturbo_amd import ic, io, workloads
    from gturbo_amd.solver import Solver
    w = workloads()[f"euler4096_rk{args.rko}"]
    p = w.params.replace(n=args.n)
    w0 = ic.physical_field(p, w.ic)
    s = Solver(p)
    z = io.field_to_ref(w0)
    znl, znew, u, v = (s.field() for _ in range(4))
    s.enter_data(z)
    for a in (znl, znew, u, v):
        s.enter_data(a, copyin=False)
    if args.calib_copy:
        # A plain copy of known size for scaling the HBM counters: (N+2)*N*8
        # bytes read and written.  The source is written first and three
        # more fields after it, so it is no longer held by the 256 MB
        # Infinity Cache when it is copied.
        cal = [s.field() for _ in range(4)]
        for a in cal:
            s.enter_data(a, copyin=False)
        s.synchronize()
        s.copy(cal[1], cal[0])
        s.synchronize()
        for a in cal:
            s.exit_data(a, copyout=False)
    s.fft_dir(z)
    s.trunc(z)
    seed = w.forcing_seed

    def one_step(seed):
        if p.rko == 4:
            s.step4(z, znl, znew, u, v)
        else:
            s.step(z, znl, u, v)
        return s.forcing(z, seed)

    for _ in range(args.warmup):
        seed = one_step(seed)
    ab = model_bytes(p.n, p.rko)
    cb = compact_bytes(p.n, s.rect, p.rko)

    def traffic_of(name):
        # the committed counter summary is of the 4096^2 RK4 workload only
        return read_traffic(name) if (args.n == 4096 and p.rko == 4) else None

    # Kernel breakdown: a sample of its own with every kernel class
    # event-timed (each timed launch records two events on the stream,
    # which costs a step a few per cent).
    s.synchronize()
    s.timing_read()
    s.timing_enable(True)
    for _ in range(args.breakdown_steps):
        seed = one_step(seed)
    s.synchronize()
    s.timing_enable(False)
    kernels = {}
    for name, (ms, cnt) in s.timing_read().items():
        if cnt == 0:
            continue
        avg_s = ms / cnt / 1e3
        ent = {"launches_per_step": cnt / args.breakdown_steps, "avg_us": avg_s * 1e6,
               "ms_per_step": ms / args.breakdown_steps}
        if name in cb:
            # moved bytes: what the pass reads and writes on the compact layout,
            # and what the counters measured (4096^2 RK4 only)
            ent["algorithmic_bytes"] = cb[name]
            ent["achieved_GBs"] = cb[name] / avg_s / 1e9
            tr = traffic_of(name)
            if tr is not None:
                ent["measured_bytes"] = tr
                ent["measured_GBs"] = tr / avg_s / 1e9
            ent["model_bytes"] = ab[name]
        kernels[name] = ent
    dom = max((k for k in kernels if k in ab), key=lambda k: kernels[k]["ms_per_step"])

    # Timed region: K steps; the dominant kernel class alone is event-timed,
    # live, for the roofline entry.
    s.timing_enable(True, classes=[dom])
    t0 = time.perf_counter()
    for _ in range(args.steps):
        seed = one_step(seed)
    s.synchronize()
    el = time.perf_counter() - t0
    s.timing_enable(False)
    dom_ms, dom_cnt = s.timing_read()[dom]
    # the state must stay finite
    s.update_host(z)
    if not np.all(np.isfinite(z)):
        print(json.dumps({"error": "non-finite state after the timed steps"}))
        sys.exit(3)

    rate = p.n * p.n * args.steps / el
    dom_avg_s = dom_ms / dom_cnt / 1e3
    ach = cb[dom] / dom_avg_s / 1e9
    traffic = traffic_of(dom)
    roof = {"bound": "hbm", "kernel": dom, "achieved": ach, "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": ach / HBM_PEAK_GBS, "traffic": traffic,
            "algorithmic_bytes": cb[dom], "avg_us": dom_avg_s * 1e6, "launches": dom_cnt,
            "achieved_measured": None if traffic is None else traffic / dom_avg_s / 1e9,
            "frac_measured": None if traffic is None else traffic / dom_avg_s / 1e9 / HBM_PEAK_GBS,
            # the survey's unfused model: bytes an unfused step would move
            "model_bytes": ab[dom], "model_achieved": ab[dom] / dom_avg_s / 1e9,
            "model_frac": ab[dom] / dom_avg_s / 1e9 / HBM_PEAK_GBS}
    over = [(k, v[f]) for k, v in kernels.items() for f in ("achieved_GBs", "measured_GBs")
            if f in v and v[f] > HBM_PEAK_GBS]
    if roof["achieved"] > HBM_PEAK_GBS or (roof["achieved_measured"] or 0) > HBM_PEAK_GBS:
        over.append((dom, roof["achieved"]))
    if over:
        print(json.dumps({"error": f"moved-byte rates above the HBM peak: {over}"}))
        sys.exit(4)
    step_bytes = sum(ab[k] * kernels[k]["launches_per_step"] for k in ab if k in kernels)
    step_compact = sum(cb[k] * kernels[k]["launches_per_step"] for k in cb if k in kernels)
    out = {
        "metric": METRIC,
        "value": rate,
        "unit": "grid-pt-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": el / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": w.name if args.n == 4096 else f"{w.name} at N={args.n}",
                   "n": p.n, "rko": p.rko, "dt": p.dt, "alpv": p.alpv,
                   "kept_rect": list(s.rect), "nf": s.nf},
        "roofline": roof,
        "step": {"algorithmic_bytes_per_step": step_compact,
                 "bytes_per_gp_step": step_compact / (p.n * p.n),
                 "GBs_at_measured_rate": step_compact * args.steps / el / 1e9,
                 "frac": step_compact * args.steps / el / 1e9 / HBM_PEAK_GBS,
                 "model_bytes_per_step": step_bytes,
                 "model_bytes_per_gp_step": step_bytes / (p.n * p.n),
                 "model_frac": step_bytes * args.steps / el / 1e9 / HBM_PEAK_GBS},
        "kernels": kernels,
        "kernels_sample_steps": args.breakdown_steps,
    }
    s.close()
    progress(f"RK{p.rko} at {p.n}^2: {out['ms_per_step']:.3f} ms per step")
    if args.rk2_steps > 0 and args.rko == 4:
        out["rk2"] = workload_rate("euler4096_rk2", args.rk2_steps, 5, n=args.n)
        progress("RK2 done")
    if args.config_steps:
        out["configs"] = {"euler256": workload_rate("euler256", 1000, 10),
                          "sqg1024": workload_rate("sqg1024", 2000, 10)}
        progress("256^2 and 1024^2 workloads done")
    if args.ens_steps > 0:
        out["ensemble"] = ensemble_bench(args.ens_steps)
        progress("ensemble done")
    if not args.no_adjoint:
        out["adjoint"] = adjoint_bench()
        progress("adjoint done")
    if args.config_steps and args.n == 4096:
        out["diagnostics"] = diag_bench()
        progress("diagnostics done")
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(p, w0, w.forcing_seed)
        out["cpu_baseline"]["gpu_over_cpu"] = rate / out["cpu_baseline"]["value"]
        progress("CPU baseline done")
    print(json.dumps(out))


if __name__ == "__main__":
    main()
