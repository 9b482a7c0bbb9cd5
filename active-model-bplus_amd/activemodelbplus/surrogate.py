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
BIN_EDGES = np.linspace(0, 3, 300)[1:]


# ---- observables --------------------------------------------------------------------

def structure_factor(phi, dx=1.0, dy=1.0, bin_edges=BIN_EDGES):
    """Radially binned S(k) = <|phi_k|^2> of a periodic field (the reference notebook's
    examples/structureFactor.ipynb cell 7 convention: phi_k = FFT(phi - mean) dx dy,
    bins ``linspace(0, 3, 300)[1:]``).  Returns (k centres, S) with NaN for empty bins."""
    phi = np.asarray(phi, dtype=np.float64)
    ny, nx = phi.shape
    phiq = np.abs(np.fft.fft2(phi - phi.mean()) * dx * dy) ** 2
    kx = 2 * np.pi * np.fft.fftfreq(nx, d=dx)
    ky = 2 * np.pi * np.fft.fftfreq(ny, d=dy)
    K = np.sqrt(kx[None, :] ** 2 + ky[:, None] ** 2)
    idx = np.digitize(K, bin_edges, right=True)
    nb = len(bin_edges) - 1
    counts = np.bincount(idx.ravel(), minlength=nb + 2)[1:nb + 1]
    sums = np.bincount(idx.ravel(), weights=phiq.ravel(), minlength=nb + 2)[1:nb + 1]
    with np.errstate(invalid="ignore", divide="ignore"):
        S = sums / counts
    return 0.5 * (bin_edges[1:] + bin_edges[:-1]), S


def mean_wavenumber(phi, **kw):
    """First moment <k> = sum k S(k) / sum S(k) over the occupied bins."""
    k, S = structure_factor(phi, **kw)
    ok = np.isfinite(S)
    return float(np.sum(k[ok] * S[ok]) / np.sum(S[ok]))


def peak_wavenumber(phi, **kw):
    k, S = structure_factor(phi, **kw)
    return float(k[np.nanargmax(S)])


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

def train(sur, phi_in, phi_out, params, schedule=((4000, 1e-3), (3000, 3e-4), (1000, 1e-4)), batch=16,
          clip=1.0, chunk=200, log=print):
    """Adam on random mini-batches; ``schedule`` = ((iterations, learning rate), ...).
    Returns the list of (iteration, mean loss of the chunk)."""
    sur.set_dataset(phi_in, phi_out, params)
    hist, it, t0 = [], 0, time.time()
    for n, lr in schedule:
        done = 0
        while done < n:
            k = min(chunk, n - done)
            loss = sur.train(k, batch, lr, clip)
            done += k
            it += k
            hist.append((it, loss))
            log(f"  train: it {it:6d}  lr {lr:.1e}  loss {loss:.4e}  {time.time() - t0:6.1f} s")
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
    surrogate, auditing it against the Euler path every ``audit_every`` coarse steps.

    Audit (``Surrogate.audit``, on the device): from the same state, one surrogate
    step and ``delta_steps`` Euler steps on a second integrator, the judge; error =
    |phi_S - phi_E| / |phi_E| (L2).  If the error exceeds ``tol`` (or a field is not
    finite) the Euler result is kept and the next ``audit_every - 1`` coarse steps run
    on the Euler path (fallback); otherwise the run continues with the surrogate.  Mass
    is conserved exactly on both paths.
    """

    def __init__(self, sur, field, model, stencil=None, delta_steps=DELTA_STEPS, audit_every=20, tol=0.02,
                 seed=0):
        stencil = stencil or Stencil(DT, 1, 1)
        self.sur = sur
        self.model, self.stencil = model, stencil
        self.params3 = [model.lamb, model.zeta, float(np.mean(field))]
        self.main = Integrator(field, stencil, model, seed=seed, precision=32)
        self.judge = Integrator(field, stencil, model, seed=seed, precision=32)
        self.delta, self.audit_every, self.tol = delta_steps, audit_every, tol
        self.stats = dict(coarse_steps=0, surrogate_steps=0, fallback_steps=0, audits=0, audit_failures=0,
                          max_audit_error=0.0, audit_errors=[])

    @property
    def field(self):
        return self.main.field

    @property
    def time(self):
        return self.main.time

    def _audit(self):
        err, ok = self.sur.audit(self.main, self.judge, self.params3, self.delta, self.tol)
        st = self.stats
        st["audits"] += 1
        st["audit_errors"].append(err)
        st["max_audit_error"] = max(st["max_audit_error"], err) if np.isfinite(err) else float("inf")
        if not ok:
            st["audit_failures"] += 1
        return ok

    def run_coarse(self, n):
        """n coarse steps (n * delta_steps Euler-equivalent steps)."""
        st = self.stats
        done = 0
        while done < n:
            ok = self._audit()
            done += 1
            st["coarse_steps"] += 1
            if ok:
                st["surrogate_steps"] += 1
            else:
                st["fallback_steps"] += 1
            k = min(self.audit_every - 1, n - done)
            if k <= 0:
                continue
            if ok:
                self.sur.advance(self.main, k, self.params3, self.delta)
                st["surrogate_steps"] += k
            else:
                self.main.run(k * self.delta)
                st["fallback_steps"] += k
            st["coarse_steps"] += k
            done += k

    def run_for_time(self, t):
        n = int(round(t / (self.delta * self.stencil.dt)))
        self.run_coarse(n)


# ---- evaluation ------------------------------------------------------------------------

def evaluate(sur, n_eval=4, side=256, t_max=1e4, seed=1234, audit_every=20, tol=0.02, checkpoints=(10, 100, 1000, 1e4),
             log=print, euler_timing=True):
    """Held-out parameter draws: Euler reference vs hybrid vs surrogate-only vs coarse Euler.

    Returns a dict of metrics: S(k) first moment and peak at each checkpoint time, mass
    drift, wall times and the speedup of the hybrid over plain Euler to t_max."""
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
    ens = Ensemble(fields, st, models, list(range(n_eval)), precision=32)
    ens.autotune(2)
    cur = 0
    ref_fields = {}
    for c in cps:
        target = int(round(c / DT))
        ens.run(target - cur)
        cur = target
        ref_fields[c] = [ens.field(i) for i in range(n_eval)]
        log(f"  euler reference at t={c:g}")
    for i in range(n_eval):
        res["euler"].append({str(c): dict(k_mean=mean_wavenumber(ref_fields[c][i]),
                                         k_peak=peak_wavenumber(ref_fields[c][i]),
                                         mass=float(ref_fields[c][i].sum())) for c in cps})

    # hybrid and surrogate-only rollouts
    t_hybrid = []
    for i in range(n_eval):
        for mode in ("hybrid", "surrogate_only"):
            h = HybridIntegrator(sur, fields[i], models[i], st, audit_every=audit_every if mode == "hybrid" else 10**9,
                                 tol=tol if mode == "hybrid" else 1e9)
            m0 = float(h.field.sum())
            out, prev, t0 = {}, 0.0, time.perf_counter()
            for c in cps:
                h.run_for_time(c - prev)
                prev = c
                f = h.field
                out[str(c)] = dict(k_mean=mean_wavenumber(f), k_peak=peak_wavenumber(f),
                                   mass_drift=float(f.sum() - m0), l2_vs_euler=float(
                                       np.linalg.norm(f - ref_fields[c][i]) / np.linalg.norm(ref_fields[c][i])))
            wall = time.perf_counter() - t0
            stats = {k: v for k, v in h.stats.items() if k != "audit_errors"}
            stats["median_audit_error"] = float(np.median(h.stats["audit_errors"])) if h.stats["audit_errors"] else None
            res[mode].append(dict(metrics=out, wall_s=wall, stats=stats))
            if mode == "hybrid":
                t_hybrid.append(wall)
            log(f"  {mode} member {i}: {wall:6.2f} s, fallback {stats['fallback_steps']}/{stats['coarse_steps']}, "
                f"<k> err at t={cps[-1]:g}: {out[str(cps[-1])]['k_mean'] / res['euler'][i][str(cps[-1])]['k_mean'] - 1:+.3f}")

    # plain Euler wall time to t_max for one member (the baseline the hybrid replaces)
    if euler_timing:
        sim = Integrator(fields[0], st, models[0], seed=0, precision=32)
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
    e = sub.add_parser("evaluate")
    e.add_argument("--weights", required=True)
    e.add_argument("--members", type=int, default=4)
    e.add_argument("--t-max", type=float, default=1e4)
    e.add_argument("--audit-every", type=int, default=20)
    e.add_argument("--tol", type=float, default=0.02)
    e.add_argument("--out")
    a = p.parse_args(argv)

    def log(msg):
        print(msg, flush=True)

    if a.cmd == "train":
        t0 = time.time()
        data = generate_pairs(a.realisations, 256, a.t_max, a.times, a.seed, log=log)
        log(f"dataset: {data[0].shape[0]} pairs in {time.time() - t0:.1f} s")
        sur = Surrogate(256, 256, max_batch=a.batch, seed=a.seed)
        it = a.iterations
        sched = ((int(it * 0.5), 1e-3), (int(it * 0.375), 3e-4), (it - int(it * 0.5) - int(it * 0.375), 1e-4))
        t1 = time.time()
        hist = train(sur, *data[:3], schedule=sched, batch=a.batch, log=log)
        meta = dict(pairs=int(data[0].shape[0]), realisations=a.realisations, t_max=a.t_max, iterations=it,
                    batch=a.batch, train_s=time.time() - t1, final_loss=hist[-1][1], delta_steps=DELTA_STEPS, dt=DT)
        save_weights(sur, a.out, meta)
        log(json.dumps(meta))
    else:
        sur = Surrogate(256, 256, max_batch=1, seed=0)
        meta = load_weights(sur, a.weights)
        res = evaluate(sur, a.members, 256, a.t_max, audit_every=a.audit_every, tol=a.tol, log=log)
        res["weights_meta"] = meta
        text = json.dumps(res)
        if a.out:
            with open(a.out, "w") as f:
                f.write(text)
        log(text[:2000])


if __name__ == "__main__":
    _main()
