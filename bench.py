#!/usr/bin/env python3
"""Benchmark of the fused Active Model B+ step on one MI355X.

Default workload: a single 4096 x 4096 periodic field, FP32 state, the roofline
configuration (a=-0.25, b=0, c=0.25, kappa=1, lambda=0.5, zeta=-1, T=0.05,
dt=5e-3, dx=dy=1, phi = 0.05 N(0,1) from numpy seed 3).  A "step" is one explicit
Euler step of the whole field; the field stays resident in HBM.

Prints ONE JSON line.  ``value`` is site-updates per second over the timed
region (wall clock around exactly --steps steps, after --warmup untimed steps).
``roofline`` prices the step kernel against HBM: algorithmic bytes per launch
(read phi^n + write phi^{n+1} = 8 B/site FP32, 16 B/site FP64) over the average
launch duration from HIP events on the integrator's own stream.  ``cpu_baseline``
times the CPU FP64 oracle (oracle/, the checker) on this host on a bounded sample.

Other workloads (--workload): c1 (256^2 FP64 spinodal), c2 (1024^2 FP32 noisy
active), c3 (default), c3t0 (the C3 model without noise: the deterministic path), c3f64
(4096^2 FP64), c3f64t0 (the same without noise), c3x8192 (the C3 model on an 8192^2 FP32
field: phi^n and phi^(n+1) together are 512 MiB, past the 256 MB MALL), c4 (batched ensemble of 512
realisations, sides 64..512, vs sequential plain runs on the <=256^2 subset), c5
(learned surrogate with Euler audits vs plain Euler, 256^2; --steps = coarse steps of
100 dt each, default 20000 = t 10^4, the horizon the speedup is quoted at).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "active-model-bplus_amd"))
sys.path.insert(0, REPO)

HBM_PEAK_GBS = 8000.0
TRAFFIC_PROFILE = {"c3": "step_c3_fp32_4096", "c3x8192": "step_c3_fp32_8192"}
VALU_PROFILE = {"c3": "step_c3_fp32_4096"}
# Non-packed FP32 vector issue rate of MI355X: 256 CUs x 4 SIMDs x 16 lanes at 2.4 GHz.
VALU_ISSUE_PEAK = 256 * 64 * 2.4e9
METRIC = "site-updates/s, % HBM peak; batched vs plain; surrogate speedup at ≤2% S(k) err"
MFMA_BF16_PEAK_TFLOPS = 2500.0


def surrogate_flops(h, w):
    """Algorithmic MFMA FLOPs of one inference step of the flux-form U-Net at h x w:
    18 h w cin cout per 3x3 convolution and 2 h w cin cout per 1x1 (c1 counted with its 5
    real input channels); the FP32 head (32 -> 2, on the vector ALUs) is not included."""
    layers = [(1, 5, 32, 9), (2, 32, 64, 9), (4, 64, 128, 9), (4, 128, 64, 1), (2, 64, 64, 9), (2, 64, 32, 1),
              (1, 32, 32, 9)]  # (down-sampling factor, cin, cout, taps)
    return sum(2 * taps * (h // f) * (w // f) * cin * cout for f, cin, cout, taps in layers)

WORKLOADS = {
    "c1": dict(n=256, precision=64, seed=1, stencil=(1e-3, 1, 1), model=(-0.25, 0, 0.25, 1, 0, 0, 0),
               init="spinodal", label="C1 passive spinodal 256x256 FP64"),
    "c2": dict(n=1024, precision=32, seed=2, stencil=(5e-3, 1, 1), model=(-0.25, 0, 0.25, 1, -1, -4, 0.1),
               init=("normal", -0.4, 0.01), label="C2 AMB+ noisy 1024x1024 FP32"),
    "c3": dict(n=4096, precision=32, seed=3, stencil=(5e-3, 1, 1), model=(-0.25, 0, 0.25, 1, 0.5, -1, 0.05),
               init=("normal", 0.0, 0.05), label="C3 AMB+ noisy 4096x4096 FP32 (roofline)"),
    "c3t0": dict(n=4096, precision=32, seed=3, stencil=(5e-3, 1, 1), model=(-0.25, 0, 0.25, 1, 0.5, -1, 0.0),
                 init=("normal", 0.0, 0.05), label="C3 model at T = 0 (noiseless) 4096x4096 FP32"),
    "c3f64": dict(n=4096, precision=64, seed=3, stencil=(5e-3, 1, 1), model=(-0.25, 0, 0.25, 1, 0.5, -1, 0.05),
                  init=("normal", 0.0, 0.05), label="C3 AMB+ noisy 4096x4096 FP64"),
    "c3f64t0": dict(n=4096, precision=64, seed=3, stencil=(5e-3, 1, 1), model=(-0.25, 0, 0.25, 1, 0.5, -1, 0.0),
                    init=("normal", 0.0, 0.05), label="C3 model at T = 0 (noiseless) 4096x4096 FP64"),
    "c3x8192": dict(n=8192, precision=32, seed=3, stencil=(5e-3, 1, 1), model=(-0.25, 0, 0.25, 1, 0.5, -1, 0.05),
                    init=("normal", 0.0, 0.05), label="C3 AMB+ noisy 8192x8192 FP32 (beyond the MALL)"),
}


def initial_field(w):
    rng = np.random.default_rng(w["seed"])
    n = w["n"]
    if w["init"] == "spinodal":
        phi = 0.1 * rng.random((n, n))
        return phi - phi.mean()
    _, mean, std = w["init"]
    return mean + std * rng.normal(size=(n, n))


def host_cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(w, phi, budget_s=15.0, budget_1core_s=8.0):
    """Times the FP64 CPU oracle (OpenMP over rows) on this host: with up to 16 threads
    (``value``, ``cores``) and on one thread (``value_1core``)."""
    from oracle import oracle as o

    o.lib()
    threads = max(1, min(16, os.cpu_count() or 1, o.max_threads()))
    n = phi.shape[0]
    # bounded sample: a sub-grid sized so one step takes a fraction of the budget
    side = n
    while side > 256 and (side * side) > 4e6:
        side //= 2
    sub = np.ascontiguousarray(phi[:side, :side])

    def rate(nthreads, budget):
        t0 = time.perf_counter()
        o.run(sub, w["stencil"], w["model"], 1, seed=w["seed"], nthreads=nthreads)
        one = time.perf_counter() - t0
        steps = max(1, int(budget / max(one, 1e-6)))
        t0 = time.perf_counter()
        o.run(sub, w["stencil"], w["model"], steps, seed=w["seed"], nthreads=nthreads)
        dt = time.perf_counter() - t0
        return side * side * steps / dt, steps, dt

    value, steps, dt = rate(threads, budget_s)
    value1, steps1, dt1 = rate(1, budget_1core_s)
    return {"value": value, "unit": "site-updates/s", "cores": threads, "kind": "port",
            "sample": f"oracle FP64, {side}x{side} periodic sub-grid of the same workload, {steps} steps, "
                      f"{threads} OpenMP threads, {dt:.1f} s; 1 thread: {steps1} steps, {dt1:.1f} s",
            "value_1core": value1, "host_cpu": f"{host_cpu_model()}, {os.cpu_count()} logical CPUs visible"}


def read_profile_value(path, tile, tile_key, value_key):
    """The value of `value_key:` in a profile summary under profiles/, or None when the file
    is missing or its `tile_key:` line names another tile shape than `tile`."""
    if not os.path.exists(path):
        return None
    fields = {}
    for line in open(path):
        if ":" in line and line.split(":")[0].strip() in (tile_key, value_key):
            k, v = line.split(":", 1)
            fields[k.strip()] = v.strip()
    if fields.get(tile_key) != "x".join(str(t) for t in tile) or value_key not in fields:
        return None
    return float(fields[value_key])


def read_traffic(name, tile):
    """HBM bytes per launch measured with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE and
    calibrated against a plain copy kernel (profiles/<name>_traffic.md), or None when
    no measurement exists for this workload and tile shape."""
    v = read_profile_value(os.path.join(REPO, "profiles", f"{name}_traffic.md"), tile, "traffic_tile",
                           "traffic_bytes_per_launch")
    return None if v is None else int(v)


def read_valu(name, tile):
    """VALU lane-instructions per site measured with rocprofv3 --pmc SQ_INSTS_VALU
    (profiles/<name>_sq_counters.md), or None when no measurement exists for this
    workload and tile shape."""
    return read_profile_value(os.path.join(REPO, "profiles", f"{name}_sq_counters.md"), tile, "valu_tile",
                              "valu_lane_instructions_per_site")


def bench_single(args, key):
    from activemodelbplus.integrator import Integrator, Model, Stencil

    w = WORKLOADS[key]
    phi = initial_field(w)
    sim = Integrator(phi, Stencil(*w["stencil"]), Model(*w["model"]), seed=w["seed"], precision=w["precision"],
                     tile=args.tile)
    tile = sim.autotune(4) if args.autotune else sim.options["tile"] - 1
    if args.warmup:
        sim.run(args.warmup)
    t0 = time.perf_counter()
    device_ms = sim.run_timed(args.steps)
    wall = time.perf_counter() - t0
    sites = phi.size
    value = sites * args.steps / wall
    esize = 4 if w["precision"] == 32 else 8
    bytes_per_launch = 2 * esize * sites
    launch_s = device_ms / 1e3 / args.steps
    achieved = bytes_per_launch / launch_s / 1e9
    mass, bad = sim.mass()
    out = {
        "metric": METRIC, "value": value, "unit": "site-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall * 1e3 / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32" if w["precision"] == 32 else "f64",
        "data": "synthetic (seeded numpy normal initial field)",
        "config": {"workload": w["label"], "grid": [w["n"], w["n"]], "stencil": w["stencil"],
                   "model": dict(zip(["a", "b", "c", "kappa", "lamb", "zeta", "T"], w["model"])),
                   "tile": dict(zip(["cols", "rows", "threads"], sim_tile_shape(w["precision"], tile)))},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": read_traffic(TRAFFIC_PROFILE.get(key, ""), sim_tile_shape(w["precision"], tile))},
        "device_ms_per_step": device_ms / args.steps,
        "mass_fp64": mass, "nonfinite": bad,
    }
    valu = read_valu(VALU_PROFILE.get(key, ""), sim_tile_shape(w["precision"], tile))
    if valu is not None:
        # the bound that holds: vector-ALU issue (the HBM fraction above is the honest one
        # against 8 TB/s, but the step runs out of VALU issue slots first)
        issued = valu * sites / launch_s
        out["roofline"]["valu"] = {"lane_instructions_per_site": valu, "achieved": issued, "peak": VALU_ISSUE_PEAK,
                                   "unit": "lane-instructions/s", "frac": issued / VALU_ISSUE_PEAK}
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(w, phi)
    return out


def sim_tile_shape(precision, index):
    from activemodelbplus.integrator import tile_shapes
    return tile_shapes(precision)[index]


def bench_batched(args):
    """C4: 512 realisations, side in {64,128,256,512}, lambda~U[-2,2], zeta~U[-5,0],
    phi0~U[-0.6,0.6], T in {0,0.05,0.1}; batched vs sequential plain runs, on the <=256^2
    subset and on all realisations (plain times extrapolated from --steps/10 steps each)."""
    from activemodelbplus.ensembles import random_ensemble
    from activemodelbplus.integrator import Ensemble, Integrator

    count = args.realisations
    fields, models, seeds, st, params = random_ensemble(count)
    sides = params["side"]
    small = [i for i in range(count) if sides[i] <= 256]
    large = [i for i in range(count) if sides[i] > 256]

    def batched(idx, steps):
        ens = Ensemble([fields[i] for i in idx], st, [models[i] for i in idx], [seeds[i] for i in idx], precision=32)
        ens.autotune(2)
        ens.run(args.warmup)
        t0 = time.perf_counter()
        ens.run(steps)
        return time.perf_counter() - t0, ens

    def sequential(idx, steps):
        total = 0.0
        for i in idx:
            sim = Integrator(fields[i], st, models[i], seed=seeds[i], precision=32)
            sim.run(min(args.warmup, 32))
            t0 = time.perf_counter()
            sim.run(steps)
            total += time.perf_counter() - t0
        return total

    steps = args.steps
    t_all, ens = batched(list(range(count)), steps)
    if args.batched_only:
        # one ensemble, nothing else (a profiler run of the batched kernels alone)
        sites_all = int(sum(s * s for s in sides))
        return {"metric": METRIC, "value": sites_all * steps / t_all, "unit": "site-updates/s", "n_gpus": 1,
                "steps": steps, "warmup": args.warmup, "ms_per_step": t_all * 1e3 / steps, "higher_is_better": True,
                "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded numpy)",
                "config": {"workload": f"batched ensemble of {count} realisations (batched leg only)",
                           "sites": sites_all, "resident_members": ens.resident_count}}
    del ens
    t_small, _ = batched(small, steps)
    seq_steps = max(32, steps // 10)
    t_seq_small = sequential(small, seq_steps) * steps / seq_steps
    t_seq_all = t_seq_small + sequential(large, seq_steps) * steps / seq_steps
    sites_all = int(sum(s * s for s in sides))
    sites_small = int(sum(sides[i] ** 2 for i in small))
    value = sites_all * steps / t_all
    achieved = 8 * sites_all * steps / t_all / 1e9
    return {
        "metric": METRIC, "value": value, "unit": "site-updates/s", "n_gpus": 1, "steps": steps,
        "warmup": args.warmup, "ms_per_step": t_all * 1e3 / steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded numpy)",
        "config": {"workload": f"C4 batched ensemble of {count} realisations", "sites": sites_all},
        "batched_vs_plain_le256": t_seq_small / t_small,
        "batched_vs_plain_all": t_seq_all / t_all,
        "plain_sequential_le256_s_extrapolated_from": seq_steps,
        "batched_le256_site_updates_per_s": sites_small * steps / t_small,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "note": "8 B per site-update over the wall clock of the graph-replayed batched steps"},
    }


def bench_copy(args):
    """Plain device copy of a side x side FP32 array (--copy-side, default 4096: read
    64 MiB + write 64 MiB per launch): the calibration kernel for the rocprofv3
    FETCH_SIZE / WRITE_SIZE counters."""
    import torch

    side = args.copy_side
    src = torch.randn(side, side, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(max(1, args.warmup)):
        dst.copy_(src)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        dst.copy_(src)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    nbytes = 2 * src.numel() * 4
    return {"metric": "copy GB/s (counter calibration)", "value": nbytes * args.steps / wall / 1e9, "unit": "GB/s",
            "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall * 1e3 / args.steps,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
            "data": "synthetic", "config": {"workload": f"plain copy {side}x{side} FP32 (torch)",
                                              "bytes_per_launch": nbytes}}


def bench_surrogate(args):
    """C5: the learned flux-form surrogate with its Euler judge on one 256 x 256 FP32
    field at T = 0 (lambda = 1, zeta = -2, phi0 = 0, phi = 0.01 N(0,1) from numpy seed 5),
    Delta = 100 Euler steps of dt = 5e-3 per coarse step.  --steps coarse steps are
    timed for the hybrid integrator (audit every --audit-every coarse steps, tolerance
    2%; the interval doubles after audits within half the tolerance, up to
    --max-audit-every, and resets on a failure) and the same span is timed for plain
    explicit Euler; the Euler end state is the reference for the structure-factor
    first moment <k> and the L2 error."""
    from activemodelbplus import surrogate as SG
    from activemodelbplus.integrator import Integrator, Model, Stencil, Surrogate

    side, delta = 256, SG.DELTA_STEPS
    rng = np.random.default_rng(5)
    phi = 0.01 * rng.normal(size=(side, side))
    model, st = Model(-0.25, 0, 0.25, 1, 1.0, -2.0, 0.0), Stencil(SG.DT, 1, 1)
    sur = Surrogate(side, side, max_batch=1, seed=0)
    weights = SG.default_weights_path()
    trained = os.path.exists(weights)
    if trained:
        SG.load_weights(sur, weights)
    coarse = args.steps
    # warm-up: graphs, tiles and the audit path
    w = SG.HybridIntegrator(sur, phi, model, st, audit_every=args.audit_every, max_audit_every=args.max_audit_every,
                            autotune=True)
    w.run_coarse(min(args.warmup, 2 * args.audit_every))
    sim = Integrator(phi, st, model, seed=0, precision=32)
    sim.autotune(4)  # plain Euler at its fastest compiled tile for this grid (kept by autotune)
    sim.run(min(args.warmup, 2) * delta)
    sim.set_field(phi)

    h = SG.HybridIntegrator(sur, phi, model, st, audit_every=args.audit_every, max_audit_every=args.max_audit_every,
                            autotune=True)
    t0 = time.perf_counter()
    h.run_coarse(coarse)
    f_h = h.field
    t_hybrid = time.perf_counter() - t0
    t0 = time.perf_counter()
    sim.run(coarse * delta)
    f_e = sim.field
    t_euler = time.perf_counter() - t0
    only = Integrator(phi, st, model, seed=0, precision=32)
    sur.advance(only, 2, [1.0, -2.0, 0.0], delta)
    only.set_field(phi)
    dev_ms = sur.advance(only, coarse, [1.0, -2.0, 0.0], delta)
    sites = side * side
    k_e, k_h = SG.mean_wavenumber(f_e), SG.mean_wavenumber(f_h)
    s = h.stats
    tflops = surrogate_flops(side, side) * coarse / (dev_ms / 1e3) / 1e12
    return {
        "metric": METRIC, "value": sites * delta * coarse / t_hybrid, "unit": "site-updates/s (Euler-equivalent)",
        "n_gpus": 1, "steps": coarse, "warmup": args.warmup, "ms_per_step": t_hybrid * 1e3 / coarse,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "bf16 (fp32 accumulate)",
        "data": "synthetic (seeded numpy)" + ("" if trained else "; untrained weights (no weight file)"),
        "config": {"workload": f"C5 surrogate {side}x{side} T=0, Delta={delta} dt, t={coarse * delta * SG.DT:g}",
                   "audit_every": [args.audit_every, args.max_audit_every], "tol": 0.02},
        "speedup_vs_euler": t_euler / t_hybrid,
        "euler_site_updates_per_s": sites * delta * coarse / t_euler,
        "surrogate_only_us_per_step": dev_ms * 1e3 / coarse,
        "fallback_coarse_steps": s["fallback_steps"], "audits": s["audits"],
        "max_audit_error": s["max_audit_error"],
        "k_mean_euler": k_e, "k_mean_rel_err": k_h / k_e - 1,
        "k_peak_smooth_rel_err": SG.smoothed_peak_wavenumber(f_h) / SG.smoothed_peak_wavenumber(f_e) - 1,
        "l2_rel_err_vs_euler": float(np.linalg.norm(f_h - f_e) / np.linalg.norm(f_e)),
        "mass_drift": float(f_h.sum() - phi.astype(np.float32).sum()),
        "roofline": {"bound": "mfma", "achieved": tflops, "peak": MFMA_BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": tflops / MFMA_BF16_PEAK_TFLOPS, "traffic": None,
                     "note": f"{surrogate_flops(side, side) / 1e9:.2f} GFLOP per inference step (four launches), "
                             "HIP events over the surrogate-only rollout"},
    }


def main():
    p = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=None, help="timed steps (default 2000; c5: 20000 coarse steps)")
    p.add_argument("--warmup", type=int, default=300,
                   help="untimed steps first (the GPU clock ramps up over a process's first ~300 step launches)")
    p.add_argument("--workload", default="c3", choices=sorted(WORKLOADS) + ["c4", "c5", "copy"])
    p.add_argument("--audit-every", type=int, default=10, help="c5: initial coarse steps per Euler audit")
    p.add_argument("--max-audit-every", type=int, default=200, help="c5: longest audit interval")
    p.add_argument("--realisations", type=int, default=512)
    p.add_argument("--batched-only", action="store_true", help="c4: time the batched ensemble only")
    p.add_argument("--copy-side", type=int, default=4096, help="copy: side of the square FP32 array")
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    p.add_argument("--autotune", action="store_true", help="pick the tile shape by timing")
    p.add_argument("--tile", type=int, default=0, help="0 = default tile, k > 0 = tile index k-1")
    args = p.parse_args()
    if args.steps is None:
        args.steps = 20000 if args.workload == "c5" else 2000
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the multi-GPU slab "
              "decomposition is a design, not a measured path.", file=sys.stderr)
        sys.exit(2)
    if args.workload == "c4":
        out = bench_batched(args)
    elif args.workload == "c5":
        out = bench_surrogate(args)
    elif args.workload == "copy":
        out = bench_copy(args)
    else:
        out = bench_single(args, args.workload)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
