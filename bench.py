#!/usr/bin/env python3
"""Throughput of the MI355X quasilinear (QL) shearing-box time step.

One "step" is one RK3 step (three substeps) of the QL MHD model over every
(kx, ky) mode of the box, with the shearing-box remap after it, exactly the
unit of the reference's main loop (main.cpp:108-123).  The default workload is
the 64x128x128 box (4032 modes x 85 dealiased kz, Re = Rm = 8000, dt = 0.005)
with a seeded synthetic state already resident in HBM when timing starts.

Prints one JSON line.  Besides the throughput it reports
  * roofline: the dominant kernel (the fused per-mode kernel) against HBM
    bandwidth, from its algorithmic bytes per launch and its average launch
    duration measured with HIP events on the library's stream;
  * cpu_baseline: the CPU oracle (a C restatement of the reference algorithm)
    timed on this host's cores on a bounded sample of the same workload;
  * mixed: the FP32-transform path against FP64 at 128x128x256 (speed, and
    errors in mean fields, EMF / stress drives and energies at steps 100, 500
    and 2000);
  * ensemble: 256 uneven runs stepped as one run set against the same runs
    stepped one after another (aggregate mode-steps/s);
  * adjoint: gradient of the final mean-field energy at 32x32x64 (200 steps,
    checkpoint every 20) -- forward / backward time and the largest relative
    difference from central finite differences on 16 directions.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "QL RK3-CN steps/s at 64x128x128; mixed/batched vs FP64 speedup at EMF err<=1e-6"
ERROR_BAR = 1e-6  # mixed-precision bar on mean fields, EMF / stress drives and energies vs FP64


def metric_name(p, workload: str, equations: str, integrator: str, precision: str) -> str:
    """The headline metric string for the default workload (64x128x128,
    MHD_FullQlin, RK3CN_Elin as main.cpp:95 uses, FP64); a descriptive one
    naming the sizes, model, integrator and precision for any other run."""
    if (workload, equations, integrator, precision) == ("box64x128x128", "MHD_FullQlin", "RK3CN_Elin", "fp64"):
        return METRIC
    return f"QL {integrator} steps/s at {p.NX}x{p.NY}x{p.NZ} ({equations}, {precision})"


HBM_PEAK_GBS = 8000.0


def traffic_from_profiles(workload: str, equations: str = "MHD_FullQlin", precision: str = "fp64"):
    """HBM bytes per mode-kernel launch from the committed rocprofv3 PMC summary
    (measured for MHD_FullQlin in FP64; None for other models and precisions)."""
    if equations != "MHD_FullQlin" or precision != "fp64":
        return None
    path = os.path.join(ROOT, "profiles", f"k_mode_full_{workload}_traffic.txt")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        m = re.search(r"^bytes_per_launch:\s*([0-9.eE+]+)", f.read(), re.M)
    return float(m.group(1)) if m else None


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _time_oracle(library, p, lin, mf, cores, budget_s, max_steps):
    from oracle.oracle import Oracle
    o = Oracle.from_params(p, nthreads=cores, library=library)
    o.set_state(lin, mf)
    o.linear_ops_init(0.0)
    t = o.steps(2, 0.0)  # warm-up
    n = 0
    t0 = time.perf_counter()
    while n < max_steps:
        t = o.steps(1, t)
        n += 1
        if n >= 10 and time.perf_counter() - t0 > budget_s:
            break
    el = time.perf_counter() - t0
    o.close()
    return n, el


def cpu_baseline(p, lin, mf, budget_s: float = 12.0, max_steps: int = 40):
    """Times the CPU oracle (a C restatement of the reference algorithm: 32
    length-NZ transforms per mode per rhs, three rhs per step, contiguous mode
    blocks per thread as MPIdata::Split_NXY_Grid) on every core this process may
    run on: 2 warm-up steps, then at least 10 steps and about budget_s seconds.
    Built here, for this host, -O3 -march=native; the -Ofast -ffast-math build
    (the reference's own flags, SConstruct:6) is timed the same way and reported
    beside it."""
    import ctypes
    import shutil
    import subprocess
    import tempfile
    from oracle.oracle import bind
    odir = os.path.join(ROOT, "oracle")
    cores = len(os.sched_getaffinity(0))
    out = tempfile.mkdtemp(prefix="ql_oracle_timing_")
    subprocess.run(["make", "-s", "-C", odir, "timing", f"OUT={out}"], check=True)
    flags = subprocess.run(["make", "-s", "-C", odir, "flags"], check=True, capture_output=True,
                           text=True).stdout.strip().splitlines()
    res = {}
    for kind in ("timing", "fastmath"):
        L = bind(ctypes.CDLL(os.path.join(out, f"libql_oracle_{kind}.so")))
        res[kind] = _time_oracle(L, p, lin, mf, cores, budget_s, max_steps)
    shutil.rmtree(out, ignore_errors=True)
    n, el = res["timing"]
    nf, elf = res["fastmath"]
    return {"value": n / el, "unit": "steps/s", "cores": cores, "kind": "port",
            "sample": f"{n} steps of the same workload after 2 warm-up steps "
                      f"(oracle/ql_oracle.c, OpenMP contiguous mode blocks, radix-2 transforms)",
            "nproc": os.cpu_count(), "cpu_model": _cpu_model(), "flags": flags[0].split(": ", 1)[1],
            "fastmath": {"value": nf / elf, "unit": "steps/s", "steps": nf, "flags": flags[1].split(": ", 1)[1]}}


def bench_ensemble(steps: int = 500, warmup: int = 3, tune_steps: int = 2):
    """The 256-run parameter-study ensemble: one run set (all runs in the same
    launches, one replayed graph per 16 steps, launch shapes tuned first) against
    the same runs stepped one after another, each as its own single-run context.
    Throughput is aggregate mode-steps/s (modes of every run x steps / time)."""
    from ql_dns_amd.capi import QLBatch, QLContext
    from ql_dns_amd.ic import initial_state
    from ql_dns_amd.params import ensemble
    runs = ensemble()
    states = [initial_state(p) for p in runs]
    mode_steps = sum(p.nxy for p in runs) * steps
    b = QLBatch(runs)
    for r, (lin, mf) in enumerate(states):
        b.set_state(r, lin, mf)
    t_tune = time.perf_counter()
    b.tune(tune_steps)
    t_tune = time.perf_counter() - t_tune
    b.linear_ops_init(0.0)
    t = b.steps(warmup, 0.0)
    t0 = time.perf_counter()
    b.steps(steps, t)
    el_b = time.perf_counter() - t0
    finite = all(abs(b.get_state(r)[1]).max() < 1e300 for r in range(len(runs)))
    cfg = b.config()
    b.close()
    el_s = 0.0
    for p, (lin, mf) in zip(runs, states):
        c = QLContext(p)
        c.set_state(lin, mf)
        c.linear_ops_init(0.0)
        t = c.steps(warmup, 0.0)
        t0 = time.perf_counter()
        c.steps(steps, t)
        el_s += time.perf_counter() - t0
        c.close()
    return {"runs": len(runs), "modes": sum(p.nxy for p in runs), "steps": steps,
            "mode_steps_per_s": mode_steps / el_b, "run_steps_per_s": len(runs) * steps / el_b,
            "sequential_mode_steps_per_s": mode_steps / el_s, "speedup_vs_sequential": el_s / el_b,
            "seconds": el_b, "sequential_seconds": el_s, "tune_seconds": t_tune,
            "shapes": cfg, "state_finite": finite,
            "workload": "256 runs: Rm log-uniform [500, 20000], q uniform [1, 1.8], "
                        "NZ in {32, 64, 128}, NX, NY in {8, 16, 32}, Re 2000, dt 0.01"}


def bench_mixed(name: str = "box128x128x256", marks=(100, 500, 2000)):
    """FP32-transform path against the FP64 path on the same seeded state:
    speed (graph replay, steps/s) and, at each mark, the error of the mixed run
    against the FP64 run (norm-wise relative) in mean fields, EMF / stress
    drives and fluctuation energy."""
    import numpy as np
    from ql_dns_amd.capi import QLContext
    from ql_dns_amd.ic import initial_state
    from ql_dns_amd.params import PREC_MIXED, workload

    def rel(a, b):
        d = float(np.abs(b).max())
        return float(np.abs(a - b).max() / (d if d > 0 else 1.0))

    p = workload(name)
    lin, mf = initial_state(p)
    ctx = {}
    for prec in ("fp64", "mixed"):
        c = QLContext(p.replace(precision=PREC_MIXED) if prec == "mixed" else p)
        # 16 untimed steps capture the replay graphs; then the same start state
        c.set_state(lin, mf)
        c.linear_ops_init(0.0)
        c.steps(16, 0.0)
        c.set_state(lin, mf)
        c.linear_ops_init(0.0)
        ctx[prec] = c
    fp64_modes_start = ctx["mixed"].mixed_split()
    el = {"fp64": 0.0, "mixed": 0.0}
    t = {"fp64": 0.0, "mixed": 0.0}
    errs = []
    done = 0
    for m in marks:
        for prec, c in ctx.items():
            c.synchronize()
            t0 = time.perf_counter()
            t[prec] = c.steps(m - done, t[prec])
            el[prec] += time.perf_counter() - t0
        done = m
        l64, m64, _ = ctx["fp64"].get_state()
        l32, m32, _ = ctx["mixed"].get_state()
        d64, d32 = ctx["fp64"].drive(), ctx["mixed"].drive()
        e64, e32 = ctx["fp64"].diagnostics(t["fp64"]), ctx["mixed"].diagnostics(t["mixed"])
        e = {"step": m, "mean_fields": rel(m32, m64), "emf_stress_drive": rel(d32, d64),
             "fluct_energy_u": abs(e32[2] - e64[2]) / abs(e64[2]),
             "fluct_energy_b": abs(e32[3] - e64[3]) / abs(e64[3]),
             "fluct_state": rel(l32, l64)}
        e["within_1e-6"] = all(e[k] <= ERROR_BAR for k in ("mean_fields", "emf_stress_drive", "fluct_energy_u",
                                                            "fluct_energy_b"))
        errs.append(e)
    fp64_modes_end = ctx["mixed"].mixed_split()
    for c in ctx.values():
        c.close()
    n = marks[-1]
    return {"workload": name, "steps": n, "fp64_steps_per_s": n / el["fp64"],
            "mixed_steps_per_s": n / el["mixed"], "speedup": el["fp64"] / el["mixed"],
            "fp64_modes": [fp64_modes_start, fp64_modes_end], "modes": p.nxy,
            "errors_vs_fp64": errs, "error_bar": ERROR_BAR,
            "error_bar_met": all(e["within_1e-6"] for e in errs)}


def bench_adjoint(name: str = "box32x32x64", nsteps: int = 200, every: int = 20, ndir: int = 16,
                  eps: float = 1e-6):
    """Gradient of the final 0.5 <Bx^2 + By^2>_z w.r.t. the initial Bx(z),
    By(z), q and Rm through the adjoint (checkpoint every `every` steps):
    forward and backward wall time, and the largest relative difference from
    central finite differences of the FP64 forward on `ndir` seeded
    directions (each input scaled to its magnitude)."""
    import numpy as np
    from ql_dns_amd.adjoint import mf_energy
    from ql_dns_amd.capi import QLContext
    from ql_dns_amd.ic import initial_state
    from ql_dns_amd.params import workload
    p = workload(name)
    lin0, mf0 = initial_state(p)

    def forward(mf, q, Rm):
        c = QLContext(p.replace(q=q, eta=1.0 / Rm))
        c.set_state(lin0, mf)
        c.linear_ops_init(0.0)
        c.steps(nsteps, 0.0)
        L = mf_energy(p, c.get_state()[1])
        c.close()
        return L

    Rm0 = 1.0 / p.eta
    c = QLContext(p)
    c.set_state(lin0, mf0)
    c.linear_ops_init(0.0)
    c.steps(2, 0.0)  # warm-up of the launch path
    c.set_state(lin0, mf0)
    c.linear_ops_init(0.0)
    t0 = time.perf_counter()
    c.adjoint_forward(nsteps, every, 0.0)
    t_fwd = time.perf_counter() - t0
    mfT = c.get_state()[1]
    g = np.zeros_like(mfT)
    g[:2] = mfT[:2] / p.NZ ** 2
    t0 = time.perf_counter()
    _, gm0, gpar = c.adjoint_backward(None, g)
    t_bwd = time.perf_counter() - t0
    c.close()
    rng = np.random.default_rng(16)
    sm = float(np.sqrt(np.mean(np.abs(mf0[:2]) ** 2)))
    worst = 0.0
    for _ in range(ndir):
        vm = np.zeros_like(mf0)
        vm[:2] = (rng.standard_normal((2, p.nz)) + 1j * rng.standard_normal((2, p.nz))) * sm
        vq, vR = rng.standard_normal(), rng.standard_normal() * Rm0
        ad = float(np.sum(gm0.real * vm.real + gm0.imag * vm.imag)) + gpar[0] * vq - gpar[3] / Rm0 ** 2 * vR
        fd = (forward(mf0 + eps * vm, p.q + eps * vq, Rm0 + eps * vR) -
              forward(mf0 - eps * vm, p.q - eps * vq, Rm0 - eps * vR)) / (2 * eps)
        worst = max(worst, abs(ad - fd) / abs(fd))
    return {"workload": name, "steps": nsteps, "checkpoint_every": every,
            "forward_seconds": t_fwd, "backward_seconds": t_bwd, "backward_over_forward": t_bwd / t_fwd,
            "grad_q": float(gpar[0]), "grad_Rm": float(-gpar[3] / Rm0 ** 2),
            "fd_directions": ndir, "fd_eps": eps, "max_rel_err_vs_fd": worst}


def device_warm(p, lin, mf, shape, ms: float):
    """Sustained device work right before the warm-up steps.  The GPU raises
    its clocks over the first ~20 ms of sustained work after an idle spell of a
    few ms (the setup's uploads, the tuner's restore and frees): measured at
    64x128x128, 120 us per step falling to 114 us over ~100 steps, from any
    state.  A second context of the same workload steps for `ms` of wall time
    and is returned open (closing it frees memory, which idles the device); the
    timed context's state is untouched."""
    from ql_dns_amd.capi import QLContext
    w = QLContext(p)
    w.set_state(lin, mf)
    w.set_config(shape["NZ"], shape["line"], shape["target"])
    w.set_tail(shape["NZ"], shape["tail"])
    w.linear_ops_init(0.0)
    t = w.steps(16, 0.0)  # captures its graphs
    w.synchronize()
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < ms * 1e-3:
        t = w.steps(16, t)
    return w


def copy_calibration(nbytes: int = 1 << 30):
    """A plain device-to-device copy of known size, for scaling the HBM byte
    counters (FETCH_SIZE / WRITE_SIZE) that rocprofv3 reports per dispatch."""
    import torch
    n = nbytes // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": nbytes, "kib": nbytes / 1024}), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="box64x128x128")
    ap.add_argument("--integrator", choices=["RK3CN_Elin", "RK3CN"], default="RK3CN_Elin")
    ap.add_argument("--equations", choices=["MHD_FullQlin", "MHD_BQlin", "HD_fullU"], default="MHD_FullQlin",
                    help="model of the timed run (HD_fullU runs the velocity-only policy)")
    ap.add_argument("--profile-steps", type=int, default=50)
    ap.add_argument("--device-warm-ms", type=float, default=40.0,
                    help="sustained device work (a second context) right before the warm-up steps; 0: none")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the 256-run ensemble section")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision section")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint section")
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64",
                    help="arithmetic of the main workload (mixed: FP32 transforms; for profiling)")
    ap.add_argument("--no-tune", action="store_true",
                    help="keep the default launch shape (profiling runs: only steady-state launches)")
    ap.add_argument("--transform", choices=["fft", "dft"], default=None,
                    help="z transforms of the workgroup kernel: LDS radix passes or matrix-core products "
                         "(NZ 32 and 64, FP64; implies --no-tune and line 128)")
    ap.add_argument("--tail", choices=["auto", "two", "one"], default="auto",
                    help="substep tail: as the tuner (or the default) leaves it, the column reduction and "
                         "mean-field kernel as two launches, or one launch (k_mf_fused)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 1 GiB device copy (counter calibration) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py: this build runs the QL step on one GPU only; the multi-GPU mode "
              "sharding is a design (mode blocks + a 4*half complex all-reduce per substep) "
              "that is not built here.", file=sys.stderr)
        sys.exit(2)

    from ql_dns_amd.capi import QLContext
    from ql_dns_amd.ic import initial_state
    from ql_dns_amd.params import INTEGRATOR_NAMES, MODEL_NAMES, PREC_FP64, PREC_MIXED, workload

    p = workload(args.workload, integrator=INTEGRATOR_NAMES[args.integrator],
                 precision=PREC_MIXED if args.precision == "mixed" else PREC_FP64,
                 model=MODEL_NAMES[args.equations])
    lin, mf = initial_state(p)
    ctx = QLContext(p)
    ctx.set_state(lin, mf)
    if args.transform:
        ctx.set_config(p.NZ, 128, 1024)
        ctx.set_transform(p.NZ, 1 if args.transform == "dft" else 0)
    elif not args.no_tune:
        ctx.tune(3)  # launch shape; restores the state
    if args.tail != "auto":
        ctx.set_tail(p.NZ, 1 if args.tail == "one" else 0)
    ctx.linear_ops_init(0.0)
    warm = device_warm(p, lin, mf, ctx.config()[0], args.device_warm_ms) if args.device_warm_ms > 0 else None
    t = ctx.steps(args.warmup, 0.0)
    ctx.synchronize()
    t0 = time.perf_counter()
    t = ctx.steps(args.steps, t)
    ctx.synchronize()
    el = time.perf_counter() - t0
    steps_per_s = args.steps / el
    if warm is not None:
        warm.close()

    # per-kernel launch durations (HIP events on the library stream)
    _, avg_ms, counts = ctx.profile_kernels(args.profile_steps, t)
    S = p.state_bytes
    bytes_per_launch = 10.0 * S / 3.0  # u read+write, RHS write (s1), RHS read+write (s2), RHS read (s3)
    achieved = bytes_per_launch / (avg_ms[0] * 1e-3) / 1e9
    traffic = traffic_from_profiles(args.workload, args.equations, args.precision)
    shape = ctx.config()[0]
    lin2, mf2, _ = ctx.get_state()
    finite = bool(abs(lin2).max() < 1e300 and abs(mf2).max() < 1e300)
    ctx.close()

    # the mode kernel the tuned shape launches (line 64: one wavefront per mode)
    elin = 1 if args.integrator == "RK3CN_Elin" else 2
    bits = 64 if args.precision == "fp64" else 32
    pol = {"MHD_FullQlin": "ModelFull", "MHD_BQlin": "ModelBQ", "HD_fullU": "ModelHD"}[args.equations]
    # (mixed precision: the wavefront kernel that also runs the FP64-flagged modes)
    kname = ((f"k_wave_mix<{p.NZ},{pol},{elin}>" if bits == 32 else f"k_wave{bits}<{p.NZ},{pol},{elin}>")
             if shape.get("line") == 64
             else f"k_mode{bits}{'d' if shape.get('transform') else ''}<{p.NZ},{shape.get('modes_per_pass')},{pol},{elin}>")
    out = {
        "metric": metric_name(p, args.workload, args.equations, args.integrator, args.precision),
        "value": steps_per_s,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * el / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32 transforms, f64 state",
        "data": "synthetic: seeded divergence-free fluctuations and 4-harmonic mean fields",
        "config": {"workload": args.workload, "NX": p.NX, "NY": p.NY, "NZ": p.NZ,
                   "modes": p.nxy, "nz_dealiased": p.nz, "Re": 1.0 / p.nu, "Rm": 1.0 / p.eta,
                   "dt": p.dt, "equations": args.equations, "integrator": args.integrator,
                   "remap": bool(p.remap), "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": kname, "algorithmic_bytes_per_launch": bytes_per_launch,
                     "avg_launch_ms": avg_ms[0]},
        "kernels_ms": {kname: avg_ms[0], ("k_mf_fused" if shape.get("tail") else "k_reduce_cols+k_meanfield"): avg_ms[1],
                       "launches_profiled": counts},
        "state_finite": finite,
        "launch_shape": shape,
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(p, lin, mf)
        out["speedup_vs_cpu"] = steps_per_s / out["cpu_baseline"]["value"]
    if not args.no_mixed:
        out["mixed"] = bench_mixed()
    if not args.no_ensemble:
        out["ensemble"] = bench_ensemble()
    if not args.no_adjoint:
        out["adjoint"] = bench_adjoint()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
