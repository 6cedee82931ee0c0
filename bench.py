#!/usr/bin/env python3
"""Single-rank SPH hydro step throughput on one MI355X (mixed vs FP64).

A step is one force evaluation from unsorted particles resident in HBM: Hilbert keys,
radix sort, permutation of the fields, cornerstone octree build and link, node
centres, neighbor search (with the smoothing-length iteration), density, EOS, IAD
and momentum-energy, in the stage order of the reference's HydroProp::computeForces.

Workload: subsonic turbulence box, 256^3 = 16,777,216 particles on a jittered
periodic lattice (sigma = 0.1 spacing), velocities a Gaussian random field with
P(k) ~ k^(-11/3) for 1 <= |k| <= 3 at rms Mach 0.3, u = 1, h for ~100 neighbors.
`value` is the mixed-precision step's particles/s; the FP64 step is timed in the same
run and reported beside it.  Also reported: the momentum-energy adjoint with respect to
positions after a step (lists recomputed), the integration step (time step, positions and
energy, smoothing lengths) that completes a time step, the VE propagator's force
evaluation (the reference's default propagator) on the same input, the glass relaxation of a 100^3 jittered
lattice with the density adjoint (200 iterations), the roofline of the dominant kernel
(live HIP-event timing; HBM traffic from the committed PMC summary in profiles/), and
the reference CPU path timed on the host cores on the same input (one warm-up step, then
the median of three; threads bound one per core), and each output field's error of the mixed
step against the FP64 step.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--side 256]
Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# OpenMP thread placement of the CPU baseline: one thread per core, packed.  The OpenMP
# runtime reads these when it loads (with torch, before main runs), so they are set here.
os.environ.setdefault("OMP_PROC_BIND", "close")
os.environ.setdefault("OMP_PLACES", "cores")

METRIC = "hydro step particles/sec (nbr search+density+IAD+mom-energy), mixed vs FP64"

# published MI355X peaks
PEAK_HBM_GBS = 8000.0
PEAK_FP64_TFLOPS = 78.6
PEAK_FP32_TFLOPS = 157.3

# algorithmic work per unit, counted from the reference pair loops
# (mul/add/sub/div/sqrt = 1 flop, one table lookup = 7 flops)
FLOP_PER_PAIR = {"densityKernel": 19.0, "iadKernel": 37.0, "momentumEnergyKernel": 121.0}
FLOP_PER_TEST = 11.0  # neighbor candidate test, the reference's own convention
IAD_FLOP_PER_PARTICLE = 60.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def valu_peak(fp64, device=0):
    """Sustained VALU FMA rate of the device in TFLOP/s (bench_valu_probe.hip, built by
    __graft_entry__.build() next to this file): FP64, or scalar FP32."""
    import ctypes as C

    path = os.path.join(ROOT, "bench_valu_probe.so")
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} not built: run __graft_entry__.build()")
    lib = C.CDLL(path)
    lib.bench_valu_peak.restype = C.c_int
    lib.bench_valu_peak.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_double)]
    v = C.c_double()
    st = lib.bench_valu_peak(device, int(bool(fp64)), C.byref(v))
    if st != 0:
        raise RuntimeError(f"bench_valu_peak failed (HIP error {st})")
    return v.value


# the step's output fields compared between the mixed and the FP64 run (same sorted order)
OUTPUT_FIELDS = ("rho", "c11", "c12", "c13", "c22", "c23", "c33", "ax", "ay", "az", "du")


def field_errors(mixed, fp64):
    """Per output field, the mixed result against the FP64 one on the same particles:
    max |m - f| / max |f| (the field's scale) and the RMS relative error
    sqrt(sum (m - f)^2 / sum f^2); for rho, which is positive, also max |m - f| / f."""
    out = {}
    for k in OUTPUT_FIELDS:
        m, f = mixed[k], fp64[k]
        diff = (m - f).abs()
        e = {"max_abs_over_scale": float(diff.max() / f.abs().max()),
             "rms_rel": float(((m - f) ** 2).sum().sqrt() / (f ** 2).sum().sqrt())}
        if k == "rho":
            e["max_rel"] = float((diff / f.abs()).max())
        out[k] = e
    return out


def run_precision(ctx, hydro, inp64, box, n, precision, steps, warmup, **step_kw):
    import torch

    step = hydro.HydroStep(ctx, n, box, precision, **step_kw)
    dt = step.d.dtype
    inp = {k: (v if (k in ("x", "y", "z") or dt == torch.float64) else v.to(dt)) for k, v in inp64.items()}
    for w in range(warmup):
        step.compute_forces(inp)
        torch.cuda.synchronize()
        log(f"precision {precision} warmup {w + 1}/{warmup} done")
    # one instrumented step for the per-stage breakdown (outside the timed region)
    step.compute_forces(inp, timed=True)
    stages = step.stage_times_ms()
    # the timed steps record HIP events around the heavy launches only (the neighbor
    # candidates and the particle loops: the dominant kernel's live launch time for the
    # roofline); every kernel's time comes from as many instrumented steps afterwards
    ctx.set_timing("loops")
    ctx.timing_summary()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        step.compute_forces(inp)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    heavy = ctx.timing_summary()
    ctx.set_timing(True)
    for k in range(steps):
        step.compute_forces(inp)
    kernels = ctx.timing_summary()
    ctx.set_timing(False)
    ms = (t1 - t0) * 1e3 / steps
    log(f"precision {precision}: {ms:.2f} ms/step over {steps} steps")
    ncs = step.d.nc.to(torch.int64)
    pairs = int((ncs - 1).clamp(min=0, max=step.d.ngmax).sum().item())
    tests = ctx.candidate_tests()
    return step, ms, stages, kernels, heavy, pairs, tests


def time_me_adjoint(step, reps=3):
    """Momentum-energy adjoint w.r.t. positions after a step, with the neighbor lists
    recomputed first (as a backward pass that does not keep the forward's lists)."""
    import torch

    g = [torch.randn(step.n, dtype=step.d.dtype, device=step.d.x.device) for _ in range(4)]
    step.momentum_energy_vjp(*g)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        step.momentum_energy_vjp(*g)
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) * 1e3 / reps


def run_ve(ctx, hydro, inp64, box, n, precision, steps, warmup):
    """The VE propagator's force evaluation (HydroVeProp::computeForces, avClean off) on the
    same input: sync, tree, groups, search, xmass, veDefGradh, EOS, iadDivvCurlv, AVswitches,
    VE momentum-energy.  Returns ms per step, stage and kernel times."""
    import torch

    step = hydro.VeHydroStep(ctx, n, box, precision)
    dt = step.d.dtype
    inp = {k: (v if (k in ("x", "y", "z") or dt == torch.float64) else v.to(dt)) for k, v in inp64.items()}
    for _ in range(warmup):
        step.compute_forces(inp)
    torch.cuda.synchronize()
    step.compute_forces(inp, timed=True)
    stages = step.stage_times_ms()
    # timed steps with events around the heavy launches only, then instrumented steps
    ctx.set_timing("loops")
    ctx.timing_summary()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step.compute_forces(inp)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    ctx.set_timing(True)
    ctx.timing_summary()
    for _ in range(steps):
        step.compute_forces(inp)
    kernels = ctx.timing_summary()
    ctx.set_timing(False)
    log(f"VE precision {precision}: {ms:.2f} ms/step over {steps} steps")
    out = {"ms_per_step": ms, "particles_per_s": n / (ms * 1e-3), "stages_ms": {k: round(v, 3) for k, v in stages.items()},
           "kernels_ms_per_step": {k: round(v[1] / steps, 3) for k, v in kernels.items()}}
    rho = step.d.rho.double().clone()
    del step
    torch.cuda.empty_cache()
    return out, rho


def run_ve_both(ctx, hydro, inp64, box, n, precs, steps, warmup):
    """run_ve for each precision; with both, the mixed VE density's error against FP64."""
    from sphexa_amd._lib import MIXED

    res, rho = {}, {}
    for p in precs:
        key = "mixed" if p == MIXED else "fp64"
        res[key], rho[key] = run_ve(ctx, hydro, inp64, box, n, p, steps, warmup)
    if len(rho) == 2:
        res["rho_max_rel_err_mixed_vs_fp64"] = float(((rho["mixed"] - rho["fp64"]).abs() / rho["fp64"].abs()).max())
    return res


def time_integrate(step, reps=5):
    """HydroProp::integrate after a force evaluation: computeTimestep (acceleration limit
    included, g = 1), computePositions and updateSmoothingLength on the step's SFC-ordered
    fields, with zero x_m1 / du_m1 state.  Each call moves the particles, so it runs after
    every other measurement on this step."""
    import torch

    from sphexa_amd import hydro

    d = step.d
    d.allocate(d.x.device, integration=True)
    ts = hydro.TimestepState(minDt=1e-6, minDt_m1=1e-6, g=1.0)
    step.integrate(ts)
    torch.cuda.synchronize()
    step.ctx.set_timing(True)
    step.ctx.timing_summary()
    t0 = time.perf_counter()
    for _ in range(reps):
        step.integrate(ts)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / reps
    kernels = step.ctx.timing_summary()
    step.ctx.set_timing(False)
    return {"ms": ms, "kernels_ms": {k: round(v[1] / reps, 4) for k, v in kernels.items()}, "minDt": ts.minDt,
            "note": "computeTimestep (acceleration limit, one host sync) + computePositions + "
                    "updateSmoothingLength (one host sync)"}


def glass_relaxation(ctx, side=100, iterations=200):
    """Differentiable density: gradient descent on sum (rho - mean rho)^2 over a side^3
    periodic lattice with Gaussian jitter 0.3 spacing, fixed h, FP64 density adjoint."""
    import numpy as np
    import torch

    from sphexa_amd import autodiff, hydro
    from sphexa_amd._lib import FP64
    from sphexa_amd.workloads import jittered_lattice

    n = side ** 3
    f = jittered_lattice(side, 0.3, seed=4)[0]
    dev = ctx.device
    X = [torch.tensor(f[k], device=dev) for k in ("x", "y", "z")]
    h = torch.full((n,), 0.5 * (300 / (4 * np.pi * n)) ** (1 / 3), dtype=torch.float64, device=dev)
    m = torch.full((n,), 1.0 / n, dtype=torch.float64, device=dev)
    box = hydro.make_box(0.0, 1.0, 1)
    autodiff.relax_glass(ctx, box, *X, h, m, iterations=2, precision=FP64)  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _, _, _, hist = autodiff.relax_glass(ctx, box, *X, h, m, iterations=iterations, precision=FP64)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    return {"particles": n, "iterations": iterations, "ms_per_iteration": dt * 1e3 / iterations,
            "loss_first": hist[0], "loss_last": hist[-1],
            "note": "per iteration: keys, sort, tree, groups, fixed-h search, density (forward) and the same "
                    "with the density adjoint (backward, lists recomputed), FP64"}


def roofline(kernels, pairs, n, steps, precision, tests):
    """Achieved flop rate of the dominant kernel from live HIP-event timings."""
    name, (launches, total) = max(kernels.items(), key=lambda kv: kv[1][1])
    per_launch_ms = total / max(1, launches)
    if name in FLOP_PER_PAIR:
        flop = FLOP_PER_PAIR[name] * pairs + (IAD_FLOP_PER_PARTICLE * n if name == "iadKernel" else 0.0)
        unit_note = f"{FLOP_PER_PAIR[name]:.0f} flop/pair x {pairs} pairs"
    elif name == "nsCandidates" and tests:
        flop = FLOP_PER_TEST * tests
        unit_note = f"{FLOP_PER_TEST:.0f} flop/candidate test x {tests} tests"
    else:
        flop = None
        unit_note = "no algorithmic flop count"
    # the neighbor predicate is FP64 in both precisions; the particle loops compute in
    # the precision of the fields
    fp64 = precision == 0 or name.startswith("ns")
    peak = PEAK_FP64_TFLOPS if fp64 else PEAK_FP32_TFLOPS
    achieved = (flop / (per_launch_ms * 1e-3) / 1e12) if flop else None
    return {"kernel": name, "bound": "valu", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
            "frac": (achieved / peak) if achieved else None, "traffic": None, "per_launch_ms": per_launch_ms,
            "launches": launches, "work": unit_note,
            "peak_note": "MI355X published vector peak for the kernel's arithmetic (FP64 78.6 for the "
                         "neighbor search and the FP64 loops, FP32 157.3 for the mixed loops); MFMA is not used"}


def measured_traffic(kernel, workload_tag):
    """HBM bytes per launch of `kernel` from the committed PMC summary
    profiles/traffic_<workload>.txt (FETCH_SIZE + WRITE_SIZE per dispatch, each scaled by
    the same counter's reading for a copy of known size); None if absent."""
    path = os.path.join(ROOT, "profiles", f"traffic_{workload_tag}.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(f"bytes_per_launch {kernel}:"):
            return float(line.split(":", 1)[1])
    return None


def calibrate_copy(nbytes):
    """A plain device copy of known size (torch's copy kernel, nbytes read + nbytes written),
    run under the PMC counters to scale FETCH_SIZE and WRITE_SIZE to bytes."""
    import torch

    a = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": nbytes, "copy_bytes_written": nbytes, "launches": 3}), flush=True)


def host_cpu():
    """Model, sockets, physical cores and logical CPUs of the host (/proc/cpuinfo).  The
    process's affinity mask is not reported: with OMP_PROC_BIND set, the OpenMP runtime
    binds the initial thread to one core when it loads."""
    model, sockets, cores = "unknown", set(), set()
    phys = None
    try:
        for line in open("/proc/cpuinfo"):
            k, _, v = line.partition(":")
            k, v = k.strip(), v.strip()
            if k == "model name":
                model = v
            elif k == "physical id":
                phys = v
                sockets.add(v)
            elif k == "core id":
                cores.add((phys, v))
    except OSError:
        pass
    return {"model": model, "sockets": len(sockets) or 1, "physical_cores": len(cores) or os.cpu_count(),
            "logical_cpus": os.cpu_count()}


def cpu_baseline(f, side, steps):
    """The reference CPU path (oracle/_ref, compiled from the reference sources) on the
    host cores, on the same input as the GPU (host copy of f): one untimed warm-up step,
    then the median of `steps` timed steps, OpenMP threads bound one per core
    (OMP_PROC_BIND=close, OMP_PLACES=cores); the oracle port if the reference build is
    absent."""
    import ctypes as C

    import numpy as np

    import oracle

    kind = "reference" if oracle.available("ref") else "port"
    n = len(f["x"])
    cpu = host_cpu()
    threads = int(os.environ.get("OMP_NUM_THREADS") or cpu["physical_cores"])
    box = oracle.Box.cube(0.0, 1.0, oracle.BC_PERIODIC)
    p = lambda a: np.ascontiguousarray(a).ctypes.data_as(C.c_void_p)
    times = []
    stages = []
    if kind == "reference":
        r = oracle.Cpu("ref")
        rho, ax, du = np.zeros(n), np.zeros(n), np.zeros(n)
        nc = np.zeros(n, np.uint32)
        st = np.zeros(11)

        def one():
            r._hydro_step(n, p(f["x"]), p(f["y"]), p(f["z"]), p(f["h"]), p(f["m"]), p(f["vx"]), p(f["vy"]),
                          p(f["vz"]), p(f["u"]), C.byref(box), 64, 100, 150, 6.0, 5.0 / 3.0, 0.2, p(rho), p(ax),
                          p(du), p(nc), p(st))
            return [round(float(v), 3) for v in st]
    else:
        from tests.helpers import oracle_step

        def one():
            oracle_step(f, box)
            return None
    for s in range(steps + 1):
        t0 = time.perf_counter()
        stg = one()
        t = time.perf_counter() - t0
        log(f"cpu baseline ({kind}) {'warm-up' if s == 0 else f'step {s}/{steps}'}: {t:.2f} s")
        if s:
            times.append(t)
            stages.append(stg)
    k = sorted(range(steps), key=lambda q: times[q])[steps // 2]
    t = times[k]
    out = {"value": n / t, "unit": "particles/s", "cores": threads, "kind": kind,
           "sample": f"turbulence {side}^3 = {n} particles, the GPU run's own input copied to the host (FP64), "
                     f"one warm-up step then the median of {steps} full steps of the reference CPU path (keys, "
                     f"sort, tree, findNeighborsSph, density, EOS, IAD, momentum-energy), {threads} OpenMP "
                     f"threads, OMP_PROC_BIND={os.environ.get('OMP_PROC_BIND')}, "
                     f"OMP_PLACES={os.environ.get('OMP_PLACES')}",
           "seconds_per_step": t, "seconds_all_steps": [round(v, 3) for v in times], "host": cpu}
    if stages[k] is not None:
        out["stages_s"] = dict(zip(("keys", "sort_gather", "octree", "link_centers", "neighbors", "density", "eos",
                                    "iad", "momentum_energy"), stages[k]))
    return out


def run_other_workload(args, ctx):
    """The evrard and uniform800 configurations: the same step and timing, mixed and FP64, with
    the neighbor-count distribution (a parity configuration, not the metric's)."""
    import numpy as np
    import torch

    from sphexa_amd import hydro, workloads
    from sphexa_amd._lib import BC_OPEN, BC_PERIODIC, FP64, MIXED

    if args.workload == "evrard":
        f, (lo, hi, per) = workloads.evrard(4_194_304, seed=105)
        kw, desc = dict(ng0=100, ngmax=512), "Evrard-like 1/r sphere, 4,194,304 particles, open box, ~100 neighbors"
    else:
        f, (lo, hi, per) = workloads.uniform_random(1_000_000, seed=20261202)
        kw, desc = dict(ng0=0, ngmax=1024), "uniform random 1,000,000 particles, periodic, h as written (~800 " \
                                            "neighbors), fixed h"
    n = len(f["x"])
    box = hydro.make_box(lo, hi, BC_PERIODIC if per else BC_OPEN)
    inp64 = {k: torch.tensor(v, dtype=torch.float64, device=ctx.device) for k, v in f.items()}
    res = {}
    for prec in (MIXED, FP64):
        step, ms, stages, kernels, _, pairs, tests = run_precision(ctx, hydro, inp64, box, n, prec, args.steps,
                                                                   args.warmup, **kw)
        nc = step.d.nc.to(torch.int64).cpu().numpy() - 1
        res[prec] = dict(ms=ms, pairs=pairs, tests=tests, kernels={k: round(v[1] / args.steps, 3)
                                                                 for k, v in kernels.items()},
                         nc={"min": int(nc.min()), "median": float(np.median(nc)), "p99": float(np.percentile(nc, 99)),
                             "max": int(nc.max())},
                         fields={k: getattr(step.d, k).double().clone() for k in OUTPUT_FIELDS})
        del step
        torch.cuda.empty_cache()
    R, F = res[MIXED], res[FP64]
    out = {"metric": METRIC, "value": n / (R["ms"] * 1e-3), "unit": "particles/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": R["ms"], "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
           "config": {"workload": desc, "particles": n, **kw, "parallelism": "single GPU"},
           "pairs_per_s": R["pairs"] / (R["ms"] * 1e-3), "candidate_tests_per_pair": R["tests"] / max(1, R["pairs"]),
           "neighbors": R["nc"], "kernels_ms_per_step": R["kernels"],
           "fp64": {"value": n / (F["ms"] * 1e-3), "ms_per_step": F["ms"], "kernels_ms_per_step": F["kernels"]},
           "mixed_over_fp64": F["ms"] / R["ms"],
           "field_errors_mixed_vs_fp64": field_errors(R["fields"], F["fields"])}
    out["rho_max_rel_err_mixed_vs_fp64"] = out["field_errors_mixed_vs_fp64"]["rho"]["max_rel"]
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--side", type=int, default=256, help="lattice side; N = side^3")
    ap.add_argument("--cpu-side", type=int, default=0, help="CPU baseline lattice side (default: --side)")
    ap.add_argument("--cpu-steps", type=int, default=3, help="timed CPU baseline steps after one warm-up")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--only", choices=["mixed", "fp64"], default=None,
                    help="time one precision only (for profiling runs)")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint timings")
    ap.add_argument("--no-ve", action="store_true", help="skip the VE propagator timings")
    ap.add_argument("--ve-only", action="store_true", help="time only the VE propagator (profiling runs)")
    ap.add_argument("--workload", choices=["turbulence", "evrard", "uniform800"], default="turbulence",
                    help="turbulence: the metric's 256^3 box (default); evrard: 4M-particle 1/r sphere in an open "
                         "box (ngmax 512); uniform800: 1M uniform random particles, h as written (~800 "
                         "neighbors, fixed h, ngmax 1024)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                    help="only run a device copy of BYTES (PMC byte-counter calibration) and exit")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              f"(the single-rank hydro path has no multi-GPU decomposition).", file=sys.stderr)
        sys.exit(2)

    import torch

    from sphexa_amd import hydro
    from sphexa_amd._lib import BC_PERIODIC, FP64, MIXED
    from sphexa_amd.workloads import turbulence

    ctx = hydro.Context(0)
    if args.workload != "turbulence":
        run_other_workload(args, ctx)
        return
    n = args.side ** 3
    log(f"generating turbulence {args.side}^3 = {n} particles")
    inp64, _ = turbulence(args.side, device=ctx.device)
    torch.cuda.synchronize()
    box = hydro.make_box(0.0, 1.0, BC_PERIODIC)

    results = {}
    precs = [MIXED, FP64] if args.only is None else ([MIXED] if args.only == "mixed" else [FP64])
    if args.ve_only:
        ve = run_ve_both(ctx, hydro, inp64, box, n, precs, args.steps, args.warmup)
        print(json.dumps({"ve_hydro": ve}), flush=True)
        return
    for prec in precs:
        step, ms, stages, kernels, heavy, pairs, tests = run_precision(ctx, hydro, inp64, box, n, prec, args.steps,
                                                                       args.warmup)
        results[prec] = dict(ms=ms, stages=stages, kernels=kernels, heavy=heavy, pairs=pairs, tests=tests,
                             fields={k: getattr(step.d, k).double().clone() for k in OUTPUT_FIELDS})
        if not args.no_adjoint:
            results[prec]["me_adjoint_ms"] = time_me_adjoint(step)
            log(f"precision {prec}: momentum-energy adjoint {results[prec]['me_adjoint_ms']:.2f} ms")
        results[prec]["integrate"] = time_integrate(step)
        log(f"precision {prec}: integrate {results[prec]['integrate']['ms']:.3f} ms")
        del step
        torch.cuda.empty_cache()

    main_prec = MIXED if MIXED in results else FP64
    R = results[main_prec]
    out = {
        "metric": METRIC,
        "value": n / (R["ms"] * 1e-3),
        "unit": "particles/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": R["ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if main_prec == MIXED else "f64",
        "data": "synthetic",
        "config": {"workload": f"subsonic turbulence {args.side}^3 jittered periodic lattice ({n} particles), "
                               f"~100 neighbors, rms Mach 0.3, mixed precision (FP64 positions/offsets, "
                               f"FP32 weights and Kahan sums)",
                   "particles": n, "bucket": 64, "ng0": 100, "ngmax": 150, "parallelism": "single GPU"},
        "pairs_per_s": R["pairs"] / (R["ms"] * 1e-3),
        "candidate_tests_per_pair": (R["tests"] or 0) / max(1, R["pairs"]),
        "stages_ms": {k: round(v, 3) for k, v in R["stages"].items()},
        "kernels_ms_per_step": {k: round(v[1] / args.steps, 3) for k, v in R["kernels"].items()},
    }
    # whole-step algorithmic rates: pair flops of the three loops (19 + 37 + 121 per pair, 60 per
    # particle for the IAD inverse), 11 flops per candidate test of the search, and the
    # compulsory HBM bytes per particle of the step (keys, sort, permutation, loops)
    def step_totals(res, prec):
        flop = (19.0 + 37.0 + 121.0) * res["pairs"] + IAD_FLOP_PER_PARTICLE * n + FLOP_PER_TEST * (res["tests"] or 0)
        nbytes = (532.0 if prec == MIXED else 716.0) * n
        t = res["ms"] * 1e-3
        return {"algorithmic_TFLOPs": flop / t / 1e12, "compulsory_GBs": nbytes / t / 1e9,
                "hbm_frac": nbytes / t / 1e9 / PEAK_HBM_GBS,
                "note": "step flops = 177 per pair + 60 per particle + 11 per candidate test; compulsory bytes "
                        f"{'532' if prec == MIXED else '716'} per particle"}
    out["step_totals"] = step_totals(R, main_prec)
    if FP64 in results and MIXED in results:
        F = results[FP64]
        out["fp64"] = {"value": n / (F["ms"] * 1e-3), "ms_per_step": F["ms"],
                       "stages_ms": {k: round(v, 3) for k, v in F["stages"].items()},
                       "kernels_ms_per_step": {k: round(v[1] / args.steps, 3) for k, v in F["kernels"].items()}}
        out["fp64"]["integrate_ms"] = F["integrate"]["ms"]
        out["fp64"]["step_totals"] = step_totals(F, FP64)
        out["mixed_over_fp64"] = F["ms"] / R["ms"]
        errs = field_errors(results[MIXED]["fields"], results[FP64]["fields"])
        out["rho_max_rel_err_mixed_vs_fp64"] = errs["rho"]["max_rel"]
        out["field_errors_mixed_vs_fp64"] = errs
        out["field_errors_mixed_vs_fp64"]["note"] = (
            "the mixed step's outputs against the FP64 step's on the same input, same sorted order; "
            "max_abs_over_scale = max|mixed - fp64| / max|fp64|, rms_rel = ||mixed - fp64|| / ||fp64||")
    out["integrate"] = R["integrate"]
    out["time_step_ms"] = R["ms"] + R["integrate"]["ms"]
    if "me_adjoint_ms" in R:
        out["momentum_energy_adjoint"] = {
            "ms": R["me_adjoint_ms"], "wrt": "x, y, z",
            "note": "neighbor lists recomputed (search with the final h), records packed, adjoint kernel; "
                    "same precision as value"}
        if FP64 in results and "me_adjoint_ms" in results[FP64] and main_prec != FP64:
            out["momentum_energy_adjoint"]["fp64_ms"] = results[FP64]["me_adjoint_ms"]
    if not args.no_adjoint:
        log("glass relaxation (density adjoint), 100^3")
        out["density_adjoint_relaxation"] = glass_relaxation(ctx)
    if not args.no_ve:
        ve_steps = max(1, min(args.steps, 5))
        out["ve_hydro"] = run_ve_both(ctx, hydro, inp64, box, n, precs, ve_steps, 1)
        out["ve_hydro"]["note"] = ("VE propagator force evaluation (xmass, veDefGradh, EOS, iadDivvCurlv, "
                                   "AVswitches, VE momentum-energy; avClean off) on the same input, "
                                   f"{ve_steps} timed steps")
    out["roofline"] = roofline(R["heavy"], R["pairs"], n, args.steps, main_prec, R["tests"])
    # the device's sustained FMA rates, measured beside the published peaks
    out["roofline"]["valu_measured_TFLOPs"] = {"fp32_scalar": valu_peak(False), "fp64": valu_peak(True)}
    out["roofline"]["traffic"] = measured_traffic(out["roofline"]["kernel"],
                                                  f"turbulence{args.side}_{'mixed' if main_prec == MIXED else 'fp64'}")
    if out["roofline"]["traffic"]:
        # the same kernel against the HBM roofline: measured bytes per launch / launch time
        gbs = out["roofline"]["traffic"] / (out["roofline"]["per_launch_ms"] * 1e-3) / 1e9
        out["roofline"]["hbm"] = {"achieved": gbs, "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": gbs / PEAK_HBM_GBS,
                                  "note": "FETCH_SIZE + WRITE_SIZE per launch (profiles/traffic_*.txt) over the "
                                          "live launch time: the kernel is VALU-bound, not HBM-bound"}
    if not args.no_cpu:
        log("timing the CPU baseline")
        cs = args.cpu_side or args.side
        src = inp64 if cs == args.side else turbulence(cs, device=ctx.device)[0]
        host = {k: v.double().cpu().numpy() for k, v in src.items()}
        del src
        out["cpu_baseline"] = cpu_baseline(host, cs, args.cpu_steps)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
