"""Learned flux-form surrogate: data generation, training, evaluation, and a hybrid
integrator that uses the network with the explicit-Euler path as judge and fallback.

No counterpart in the reference.  The network (``Surrogate``, HIP kernels in
``csrc/amb_unet.hip``) maps (phi, lambda, zeta, phi0) to a current J and advances
phi by phi - div J over a coarse step Delta = ``delta_steps`` Euler steps, so the
mass sum(phi) is conserved exactly.  Training pairs (phi(t), phi(t + Delta)) come
from the batched Euler solver (``Ensemble``) at T = 0 on periodic grids.

Command line (GPU):
    python -m activemodelbplus.surrogate train --out weights.npz
    python -m activemodelbplus.surrogate evaluate --weights weights.npz
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

from . import integrator as _core
from .integrator import Ensemble, Integrator, Model, Stencil, Surrogate

DT = 5e-3
DELTA_STEPS = 100


# ---- observables --------------------------------------------------------------------
# The structure factor follows examples/structureFactor.ipynb cell 7 (see trajectory.py).
from .trajectory import (BIN_EDGES, mean_wavenumber, peak_wavenumber, smoothed_peak_wavenumber,  # noqa: E402,F401
                         structure_factor)


# ---- data -----------------------------------------------------------------------------

def sample_params(rng, n):
    """lambda ~ U[-2, 2], zeta ~ U[-5, 0], phi0 ~ U[-0.6, 0.6] (the ensemble configuration)."""
    return np.stack([rng.uniform(-2, 2, n), rng.uniform(-5, 0, n), rng.uniform(-0.6, 0.6, n)], axis=1)


def initial_fields(rng, params, side):
    return [p[2] + 0.01 * rng.normal(size=(side, side)) for p in params]


def generate_pairs(n_real=64, side=256, t_max=1e4, n_times=312, seed=5, log=print):
    """Snapshot pairs (phi(t), phi(t + Delta)) from the batched Euler solver at T = 0,
    n_times start times per realisation up to t_max, half geometrically and half
    uniformly spaced.  Returns float32 arrays
    phi_in, phi_out (S, side, side), params (S, 3) and the sample times (S,)."""
    rng = np.random.default_rng(seed)
    params = sample_params(rng, n_real)
    fields = initial_fields(rng, params, side)
    models = [Model(-0.25, 0, 0.25, 1, p[0], p[1], 0.0) for p in params]
    ens = Ensemble(fields, Stencil(DT, 1, 1), models, list(range(n_real)), precision=32)
    ens.autotune(2)
    last = int(round(t_max / DT)) - DELTA_STEPS
    # half geometric (early, fast dynamics), half uniform in time (late coarsening);
    # consecutive starts at least Delta apart (one trajectory), shifted, not dropped
    grid = np.sort(np.concatenate([np.geomspace(1, last, n_times - n_times // 2),
                                   np.linspace(1, last, n_times // 2)]))
    starts = []
    for g in grid:
        st = int(round(g)) if not starts else max(int(round(g)), starts[-1] + DELTA_STEPS)
        if st > last:
            break
        starts.append(st)
    phi_in, phi_out, par, times = [], [], [], []
    cur, t0 = 0, time.time()
    for i, st in enumerate(starts):
        ens.run(st - cur)
        a = np.stack([ens.field(r) for r in range(n_real)]).astype(np.float32)
        ens.run(DELTA_STEPS)
        b = np.stack([ens.field(r) for r in range(n_real)]).astype(np.float32)
        cur = st + DELTA_STEPS
        phi_in.append(a)
        phi_out.append(b)
        par.append(params.astype(np.float32))
        times += [st * DT] * n_real
        if i % 20 == 0 or i == len(starts) - 1:
            log(f"  data: sample time {st * DT:9.1f} ({i + 1}/{len(starts)}), {time.time() - t0:6.1f} s")
    return (np.concatenate(phi_in), np.concatenate(phi_out), np.concatenate(par), np.asarray(times))


# ---- training --------------------------------------------------------------------------

def validation_loss(sur, phi_in, phi_out, params, batch=16):
    """Mean per-sample loss over a held-out set (no parameter update)."""
    total, n = 0.0, phi_in.shape[0]
    for i in range(0, n, batch):
        j = min(n, i + batch)
        loss, _ = sur.loss_grad(phi_in[i:j].astype(np.float64), params[i:j].astype(np.float64),
                                phi_out[i:j].astype(np.float64))
        total += loss * (j - i)
    return total / n


def train(sur, phi_in, phi_out, params, schedule=((4000, 1e-3), (3000, 3e-4), (1000, 1e-4)), batch=16,
          clip=1.0, chunk=200, log=print, noise=0.0, noise_seed=0, val=None, val_every=1000):
    """Adam on random mini-batches; ``schedule`` = ((iterations, learning rate), ...).
    ``noise``: relative training-input noise (``Surrogate.set_train_noise``).  With a
    held-out set ``val`` = (phi_in, phi_out, params) the validation loss is computed
    every ``val_every`` iterations and the best parameters are kept.
    Returns the list of (iteration, mean training loss of the chunk, validation loss)."""
    sur.set_dataset(phi_in, phi_out, params)
    sur.set_train_noise(noise, noise_seed)
    hist, it, t0 = [], 0, time.time()
    best = (np.inf, None)
    for n, lr in schedule:
        done = 0
        while done < n:
            k = min(chunk, n - done)
            loss = sur.train(k, batch, lr, clip)
            done += k
            it += k
            v = None
            if val is not None and (it % val_every == 0 or it == sum(m for m, _ in schedule)):
                v = validation_loss(sur, *val)
                if v < best[0]:
                    best = (v, sur.get_params().copy())
            hist.append((it, loss, v))
            log(f"  train: it {it:6d}  lr {lr:.1e}  loss {loss:.4e}" + (f"  val {v:.4e}" if v is not None else "")
                + f"  {time.time() - t0:6.1f} s")
    sur.set_train_noise(0.0)
    if best[1] is not None:
        sur.set_params(best[1])
        log(f"  kept the parameters of the best validation loss {best[0]:.4e}")
    return hist


def default_weights_path():
    """Weights trained by ``python -m activemodelbplus.surrogate train`` (256 x 256, T = 0)."""
    return os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "surrogate_256.npz")


def save_weights(sur, path, meta=None):
    np.savez(path, params=sur.get_params(), meta=json.dumps(meta or {}))


def load_weights(sur, path):
    with np.load(path, allow_pickle=False) as z:
        sur.set_params(z["params"])
        return json.loads(str(z["meta"]))


# ---- hybrid integrator ----------------------------------------------------------------

class HybridIntegrator:
    """Advances an FP32 field in coarse steps of ``delta_steps`` Euler steps with the
    surrogate, auditing it against the Euler path.

    Audit (``Surrogate.audit_begin`` / ``audit_end``, on the device): from the same
    state, one surrogate step and ``delta_steps`` Euler steps on a second integrator,
    the judge, which runs on its own stream while the surrogate carries on; error =
    |phi_S - phi_E| / |phi_E| (L2).  If the error exceeds ``tol`` (or a field is not
    finite) the Euler result is kept and the coarse steps up to the next audit run on
    the Euler path (fallback); otherwise they run on the surrogate.  The audit interval
    starts at ``audit_every`` coarse steps, doubles after an audit with error at most
    tol / 2, up to ``max_audit_every`` (default: fixed interval), and returns to
    ``audit_every`` after a failed audit.  Mass is conserved exactly on both paths.
    """

    def __init__(self, sur, field, model, stencil=None, delta_steps=DELTA_STEPS, audit_every=20, tol=0.02,
                 seed=0, max_audit_every=None, main=None, judge=None, autotune=False):
        """``main`` / ``judge``: FP32 integrators to use (created from ``field`` if None);
        ``autotune``: pick their tile shape by timing (``Integrator.autotune``)."""
        stencil = stencil or Stencil(DT, 1, 1)
        self.sur = sur
        self.model, self.stencil = model, stencil
        self.params3 = [model.lamb, model.zeta, float(np.mean(field))]
        self.main = main if main is not None else Integrator(field, stencil, model, seed=seed, precision=32)
        self.judge = judge if judge is not None else Integrator(field, stencil, model, seed=seed, precision=32)
        if autotune:
            self.judge.set_tile(self.main.autotune(4))  # autotune keeps the fastest for main
        self.delta, self.tol = delta_steps, tol
        self.audit_min = max(1, int(audit_every))
        self.audit_max = max(self.audit_min, int(max_audit_every or audit_every))
        self.interval = self.audit_min
        # surrogate steps taken while the judge runs, before the audit is settled: enough
        # to hide the judge's delta_steps Euler steps, few enough that a failed audit
        # wastes little
        self.speculate = 16
        self.stats = dict(coarse_steps=0, surrogate_steps=0, fallback_steps=0, discarded_steps=0, audits=0,
                          audit_failures=0, max_audit_error=0.0, audit_errors=[])

    @property
    def field(self):
        return self.main.field

    @property
    def time(self):
        return self.main.time

    def _record(self, err, ok):
        st = self.stats
        st["audits"] += 1
        st["audit_errors"].append(err)
        st["max_audit_error"] = max(st["max_audit_error"], err) if np.isfinite(err) else float("inf")
        if not ok:
            st["audit_failures"] += 1
            self.interval = self.audit_min
        elif err <= 0.5 * self.tol:
            self.interval = min(2 * self.interval, self.audit_max)

    def run_coarse(self, n):
        """n coarse steps (n * delta_steps Euler-equivalent steps).  Each segment of
        ``interval`` coarse steps starts with an audit whose Euler judge runs while the
        surrogate continues (``Surrogate.audit_begin``).  If the audit fails, the
        segment's speculative surrogate steps are discarded (``audit_end`` restores
        the Euler field of the audited step), the interval returns to its minimum and
        the rest of that shorter segment runs on the Euler path.  At most
        ``speculate`` steps run before the audit is settled; the rest of a passing
        segment follows it."""
        st = self.stats
        done = 0
        while done < n:
            k = min(self.interval, n - done)
            self.sur.audit_begin(self.main, self.judge, self.params3, self.delta)
            spec = min(k - 1, self.speculate)
            if spec > 0:
                self.sur.advance(self.main, spec, self.params3, self.delta)
            err, ok = self.sur.audit_end(self.tol)
            self._record(err, ok)
            if not ok:
                st["discarded_steps"] += spec
                k = min(self.audit_min, k)
                if k > 1:
                    self.main.run((k - 1) * self.delta)
                st["fallback_steps"] += k
            else:
                if k - 1 - spec > 0:
                    self.sur.advance(self.main, k - 1 - spec, self.params3, self.delta)
                st["surrogate_steps"] += k
            st["coarse_steps"] += k
            done += k

    def run_for_time(self, t):
        n = int(round(t / (self.delta * self.stencil.dt)))
        self.run_coarse(n)


# ---- evaluation ------------------------------------------------------------------------

def evaluate(sur, n_eval=8, side=256, t_max=1e4, seed=1234, audit_every=10, tol=0.02, checkpoints=(10, 100, 1000, 1e4),
             log=print, euler_timing=True, max_audit_every=200, fields_out=None):
    """Held-out parameter draws: Euler reference vs hybrid vs surrogate-only vs coarse Euler.

    Returns a dict of metrics: S(k) first moment, argmax bin and smoothed peak
    (trajectory.smoothed_peak_wavenumber) at each checkpoint time, mass drift, wall times
    and the speedup of the hybrid over plain Euler to t_max.  A second Euler realisation per
    member (same parameters, other initial noise) gives the natural scatter of <k> and of
    the peak against which the surrogate's error is read, and an FP64 Euler run of the same
    initial field the divergence that a change of arithmetic alone produces by t_max.
    ``fields_out`` (a path) keeps the fields at t_max in an .npz."""
    rng = np.random.default_rng(seed)
    params = sample_params(rng, n_eval)
    # keep held-out draws inside the spinodal region so that they phase-separate
    params[:, 2] = rng.uniform(-0.4, 0.4, n_eval)
    fields = initial_fields(rng, params, side)
    st = Stencil(DT, 1, 1)
    models = [Model(-0.25, 0, 0.25, 1, p[0], p[1], 0.0) for p in params]
    cps = [c for c in checkpoints if c <= t_max]
    res = {"params": params.tolist(), "checkpoints": cps, "euler": [], "hybrid": [], "surrogate_only": []}

    # Euler reference (batched for speed; each member equals a plain run bit for bit)
    twins = initial_fields(np.random.default_rng(seed + 1), params, side)
    ens = Ensemble(fields + twins, st, models + models, list(range(2 * n_eval)), precision=32)
    ens.autotune(2)
    cur = 0
    ref_fields, twin_fields = {}, {}
    for c in cps:
        target = int(round(c / DT))
        ens.run(target - cur)
        cur = target
        ref_fields[c] = [ens.field(i) for i in range(n_eval)]
        twin_fields[c] = [ens.field(n_eval + i) for i in range(n_eval)]
        log(f"  euler reference at t={c:g}")
    del ens
    # the same initial fields in FP64 (the pointwise divergence of two arithmetics of one
    # deterministic, chaotic coarsening: the floor for any per-member comparison)
    ens64 = Ensemble(fields, st, models, list(range(n_eval)), precision=64)
    ens64.run(int(round(cps[-1] / DT)))
    fp64_fields = [ens64.field(i) for i in range(n_eval)]
    del ens64
    res["euler_fp64"] = [dict(k_mean=mean_wavenumber(f), k_peak_smooth=smoothed_peak_wavenumber(f)) for f in fp64_fields]
    log("  euler FP64 reference done")
    res["euler_twin_k_mean"] = [{str(c): mean_wavenumber(twin_fields[c][i]) for c in cps} for i in range(n_eval)]
    res["euler_twin_k_peak_smooth"] = [{str(c): smoothed_peak_wavenumber(twin_fields[c][i]) for c in cps}
                                       for i in range(n_eval)]
    for i in range(n_eval):
        res["euler"].append({str(c): dict(k_mean=mean_wavenumber(ref_fields[c][i]),
                                         k_peak=peak_wavenumber(ref_fields[c][i]),
                                         k_peak_smooth=smoothed_peak_wavenumber(ref_fields[c][i]),
                                         mass=float(ref_fields[c][i].sum())) for c in cps})

    # hybrid and surrogate-only rollouts
    t_hybrid = []
    finals = {}
    for i in range(n_eval):
        for mode in ("hybrid", "surrogate_only"):
            if mode == "hybrid":
                h = HybridIntegrator(sur, fields[i], models[i], st, audit_every=audit_every, tol=tol,
                                     max_audit_every=max_audit_every, autotune=True)
            else:
                h = HybridIntegrator(sur, fields[i], models[i], st, audit_every=10**9, tol=1e9)
            m0 = float(h.field.sum())
            out, prev, wall = {}, 0.0, 0.0
            for c in cps:
                t0 = time.perf_counter()
                h.run_for_time(c - prev)
                f = h.field
                wall += time.perf_counter() - t0  # rollout only, not the analysis below
                prev = c
                out[str(c)] = dict(k_mean=mean_wavenumber(f), k_peak=peak_wavenumber(f),
                                   k_peak_smooth=smoothed_peak_wavenumber(f),
                                   mass_drift=float(f.sum() - m0), l2_vs_euler=float(
                                       np.linalg.norm(f - ref_fields[c][i]) / np.linalg.norm(ref_fields[c][i])))
            finals.setdefault(mode, []).append(h.field.copy())
            stats = {k: v for k, v in h.stats.items() if k != "audit_errors"}
            stats["median_audit_error"] = float(np.median(h.stats["audit_errors"])) if h.stats["audit_errors"] else None
            res[mode].append(dict(metrics=out, wall_s=wall, stats=stats))
            if mode == "hybrid":
                t_hybrid.append(wall)
            log(f"  {mode} member {i}: {wall:6.2f} s, fallback {stats['fallback_steps']}/{stats['coarse_steps']}, "
                f"<k> err at t={cps[-1]:g}: {out[str(cps[-1])]['k_mean'] / res['euler'][i][str(cps[-1])]['k_mean'] - 1:+.3f}")

    # plain Euler wall time to t_max for one member (the baseline the hybrid replaces),
    # with its tile shape autotuned
    if euler_timing:
        sim = Integrator(fields[0], st, models[0], seed=0, precision=32)
        sim.autotune(4)  # keeps the fastest compiled tile for this grid
        sim.run(200)
        n = int(round(t_max / DT))
        t0 = time.perf_counter()
        sim.run(n)
        res["euler_wall_s"] = time.perf_counter() - t0
        res["speedup_hybrid_vs_euler"] = res["euler_wall_s"] / float(np.mean(t_hybrid))

    # ablation: the coarse step without the network (explicit Euler with dt = Delta)
    coarse = Integrator(fields[0], Stencil(DT * DELTA_STEPS, 1, 1), models[0], seed=0, precision=32)
    try:
        coarse.run(int(round(min(t_max, 100.0) / (DT * DELTA_STEPS))))
        res["coarse_euler"] = dict(finite=True, max_abs=float(np.abs(coarse.field).max()))
    except _core.CudaException as e:
        res["coarse_euler"] = dict(finite=False, error=str(e))

    last = str(cps[-1])
    errs = [res["hybrid"][i]["metrics"][last]["k_mean"] / res["euler"][i][last]["k_mean"] - 1 for i in range(n_eval)]
    res["k_mean_rel_err_hybrid"] = errs
    res["k_mean_rel_err_hybrid_ensemble"] = float(
        np.mean([res["hybrid"][i]["metrics"][last]["k_mean"] for i in range(n_eval)])
        / np.mean([res["euler"][i][last]["k_mean"] for i in range(n_eval)]) - 1)
    errs_s = [res["surrogate_only"][i]["metrics"][last]["k_mean"] / res["euler"][i][last]["k_mean"] - 1
              for i in range(n_eval)]
    res["k_mean_rel_err_surrogate_only"] = errs_s
    res["k_mean_rel_diff_euler_twin"] = [res["euler_twin_k_mean"][i][last] / res["euler"][i][last]["k_mean"] - 1
                                         for i in range(n_eval)]
    res["k_peak_rel_err_hybrid"] = [res["hybrid"][i]["metrics"][last]["k_peak"] / res["euler"][i][last]["k_peak"] - 1
                                    for i in range(n_eval)]
    # the structure-factor peak (trajectory.smoothed_peak_wavenumber), per member, for the
    # hybrid, the surrogate alone, the second Euler realisation and the FP64 Euler run of the
    # same initial field (the floor that different arithmetic alone reaches)
    ref_pk = [res["euler"][i][last]["k_peak_smooth"] for i in range(n_eval)]
    for key, vals in (("hybrid", [res["hybrid"][i]["metrics"][last]["k_peak_smooth"] for i in range(n_eval)]),
                      ("surrogate_only",
                       [res["surrogate_only"][i]["metrics"][last]["k_peak_smooth"] for i in range(n_eval)]),
                      ("euler_twin", [res["euler_twin_k_peak_smooth"][i][last] for i in range(n_eval)]),
                      ("euler_fp64", [res["euler_fp64"][i]["k_peak_smooth"] for i in range(n_eval)])):
        errs = [v / r - 1 for v, r in zip(vals, ref_pk)]
        res[f"k_peak_smooth_rel_err_{key}"] = errs
        res[f"k_peak_smooth_rel_err_{key}_mean"] = float(np.mean(errs))
        res[f"k_peak_smooth_rel_err_{key}_rms"] = float(np.sqrt(np.mean(np.square(errs))))
        res[f"k_peak_smooth_rel_err_{key}_max_abs"] = float(np.max(np.abs(errs)))
    res["k_mean_rel_err_euler_fp64"] = [res["euler_fp64"][i]["k_mean"] / res["euler"][i][last]["k_mean"] - 1
                                        for i in range(n_eval)]
    if fields_out:
        # the fields at t_max, for analysis of other observables
        np.savez_compressed(fields_out, params=params, euler=np.asarray(ref_fields[cps[-1]]),
                            euler_twin=np.asarray(twin_fields[cps[-1]]), euler_fp64=np.asarray(fp64_fields),
                            hybrid=np.asarray(finals["hybrid"]),
                            surrogate_only=np.asarray(finals["surrogate_only"]))
    return res


def _main(argv=None):
    p = argparse.ArgumentParser(prog="python -m activemodelbplus.surrogate")
    sub = p.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--out", required=True)
    t.add_argument("--realisations", type=int, default=64)
    t.add_argument("--times", type=int, default=312)
    t.add_argument("--t-max", type=float, default=1e4)
    t.add_argument("--iterations", type=int, default=8000)
    t.add_argument("--batch", type=int, default=16)
    t.add_argument("--seed", type=int, default=5)
    t.add_argument("--noise", type=float, default=0.3, help="relative training-input noise")
    t.add_argument("--val-realisations", type=int, default=8)
    f = sub.add_parser("profile", help="inference steps and training iterations on synthetic data, for a profiler")
    f.add_argument("--steps", type=int, default=1000)
    f.add_argument("--warmup", type=int, default=3000, help="untimed inference steps first (GPU clock ramp)")
    f.add_argument("--train-iterations", type=int, default=50)
    e = sub.add_parser("evaluate")
    e.add_argument("--weights", required=True)
    e.add_argument("--members", type=int, default=8)
    e.add_argument("--t-max", type=float, default=1e4)
    e.add_argument("--audit-every", type=int, default=10)
    e.add_argument("--max-audit-every", type=int, default=200)
    e.add_argument("--tol", type=float, default=0.02)
    e.add_argument("--out")
    e.add_argument("--fields", help="write the fields at t_max to this .npz")
    a = p.parse_args(argv)

    def log(msg):
        print(msg, flush=True)

    if a.cmd == "profile":
        rng = np.random.default_rng(0)
        sim = Integrator(0.1 * rng.normal(size=(256, 256)), Stencil(DT, 1, 1), Model(-0.25, 0, 0.25, 1, 0.5, -1, 0),
                         seed=0, precision=32)
        sur = Surrogate(256, 256, max_batch=1, seed=1)
        if os.path.exists(default_weights_path()):
            load_weights(sur, default_weights_path())
        sur.advance(sim, max(a.warmup, 1), [0.5, -1.0, 0.0], DELTA_STEPS)
        ms = sur.advance(sim, a.steps, [0.5, -1.0, 0.0], DELTA_STEPS)
        log(f"inference: {ms / a.steps * 1e3:.1f} us per step (batch 1, 256^2)")
        tr = Surrogate(256, 256, max_batch=16, seed=1)
        phi = rng.normal(0, 0.3, (32, 256, 256)).astype(np.float32)
        tr.set_dataset(phi, phi + 0.01, np.zeros((32, 3), np.float32))
        tr.set_train_noise(0.3, 1)
        t0 = time.perf_counter()
        tr.train(a.train_iterations, 16, 1e-4, 1.0)
        log(f"training: {(time.perf_counter() - t0) / a.train_iterations * 1e3:.2f} ms per iteration (batch 16, 256^2)")
    elif a.cmd == "train":
        t0 = time.time()
        data = generate_pairs(a.realisations, 256, a.t_max, a.times, a.seed, log=log)
        log(f"dataset: {data[0].shape[0]} pairs in {time.time() - t0:.1f} s")
        val = None
        if a.val_realisations > 0:
            v = generate_pairs(a.val_realisations, 256, a.t_max, a.times // 4, a.seed + 1000, log=log)
            val = v[:3]
            log(f"validation: {v[0].shape[0]} pairs")
        sur = Surrogate(256, 256, max_batch=a.batch, seed=a.seed)
        it = a.iterations
        sched = ((int(it * 0.5), 1e-3), (int(it * 0.375), 3e-4), (it - int(it * 0.5) - int(it * 0.375), 1e-4))
        t1 = time.time()
        hist = train(sur, *data[:3], schedule=sched, batch=a.batch, log=log, noise=a.noise, noise_seed=a.seed,
                     val=val)
        vals = [h[2] for h in hist if h[2] is not None]
        meta = dict(pairs=int(data[0].shape[0]), realisations=a.realisations, t_max=a.t_max, iterations=it,
                    batch=a.batch, train_s=time.time() - t1, final_loss=hist[-1][1], delta_steps=DELTA_STEPS, dt=DT,
                    noise=a.noise, best_val_loss=min(vals) if vals else None)
        save_weights(sur, a.out, meta)
        log(json.dumps(meta))
    else:
        sur = Surrogate(256, 256, max_batch=1, seed=0)
        meta = load_weights(sur, a.weights)
        res = evaluate(sur, a.members, 256, a.t_max, audit_every=a.audit_every, tol=a.tol, log=log,
                       max_audit_every=a.max_audit_every, fields_out=a.fields)
        res["weights_meta"] = meta
        text = json.dumps(res)
        if a.out:
            with open(a.out, "w") as f:
                f.write(text)
        log(text[:2000])


if __name__ == "__main__":
    _main()
