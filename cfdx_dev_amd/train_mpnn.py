"""Trains the learned preconditioner stage on the GPU with libcfdxhip's own
forward, backward and Adam kernels.

Training data are solver-generated: for each training system (a box of the
16.8M-cell benchmark family at a reduced size, new seed per system: wall
clustered layers, jittered and permuted cells, log-normal 1/a_P) the FP64
AMG-PCG solves the system and the pairs (r_k = b - A x_k, e_k = x* - x_k) are
kept at a fixed set of iterations k. Each system's pairs then drive `steps`
Adam steps on that system's graph. The plain mixed method's residual
reduction per iteration on the same systems is recorded with the weights.

Usage: python -m cfdx_dev_amd.train_mpnn [--n 48] [--systems 12] [--steps 120]
                                         [--validate-n 64 --validate-systems 80]
                                         [--out cfdx_dev_amd/weights/mpnn_box.npy]
Writes the float32 parameters (.npy) and a JSON sidecar (.json) with the
reference rate and the training log.
"""
from __future__ import annotations

import argparse
import json
import os
import time

import numpy as np

from . import LearnedStage, NativeAMGPreconditioner
from . import workloads as W

PAIR_ITERATIONS = (0, 1, 2, 3, 4, 6, 8, 12)
HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_OUT = os.path.join(HERE, "weights", "mpnn_box.npy")


def amg_for(s, device=0, cheb_degree=3, cheb_ratio=30.0, cheb_degree_coarse=0):
    h = NativeAMGPreconditioner(precision="fp32", prolongator_filter=0.2, filter_keeps_aggregate=True,
                                smoother_fp32="chebyshev", chebyshev_degree=cheb_degree, chebyshev_ratio=cheb_ratio,
                                chebyshev_degree_coarse=cheb_degree_coarse, device=device)
    if not h.setup(s.n, s.rowp, s.cols, s.vals):
        raise RuntimeError(h.last_error())
    return h


def train(n=48, systems=12, steps=120, lr=2e-3, seed=0, device=0, tol=1e-8, cheb_degree=3, cheb_ratio=30.0,
          log=print, cheb_degree_coarse=0):
    """Trains against the FP32 V-cycle it will run with (the Chebyshev degree
    and ratio matter: a stage trained for one V-cycle can hurt another)."""
    net = LearnedStage(device=device, params=LearnedStage.initial_params(seed))
    rates, history = [], []
    t0 = time.perf_counter()
    for k in range(systems):
        s = W.c5_box(n=n, seed=1000 + k, device=device)
        h = amg_for(s, device, cheb_degree, cheb_ratio, cheb_degree_coarse)
        x = np.zeros(s.n)
        plain = h.solve_fgmres(s.b, x, 30, 1000, tol)
        if plain.status == 0 and plain.iterations > 0:
            rates.append(-np.log10(plain.residual_relative) / plain.iterations)
        try:
            got, res = net.collect_pairs(h, s.b, PAIR_ITERATIONS, tol=1e-10)
        except RuntimeError as e:
            # The FP64 PCG can stop NOT_APPLICABLE on a system (as the
            # reference's does); such a system yields no pairs.
            rec = dict(system=k, seed=1000 + k, skipped=str(e))
            history.append(rec)
            log(json.dumps(rec))
            continue
        losses = net.train(h, steps, lr)
        rec = dict(system=k, seed=1000 + k, pairs=got, pcg_iterations=int(res.iterations),
                   plain_fgmres_iterations=int(plain.iterations), loss_first=float(losses[:8].mean()),
                   loss_last=float(losses[-8:].mean()), seconds=round(time.perf_counter() - t0, 2))
        history.append(rec)
        log(json.dumps(rec))
    trained = [r for r in history if "skipped" not in r]
    meta = dict(n=n, systems=systems, systems_trained=len(trained), pairs=int(sum(r["pairs"] for r in trained)),
                steps=steps, lr=lr, seed=seed, pair_iterations=list(PAIR_ITERATIONS),
                cheb_degree=cheb_degree, cheb_ratio=cheb_ratio, cheb_degree_coarse=cheb_degree_coarse,
                plain_rate=float(np.mean(rates)) if rates else 0.0, history=history)
    return net.get_params(), meta


def validate(params, n=64, systems=80, seed0=5000, device=0, tol=1e-8, cheb_degree=3, cheb_ratio=30.0,
             log=print, cheb_degree_coarse=0):
    """Mixed FGMRES(30) + one FP32 V-cycle on unseen boxes (seeds seed0..),
    without the stage and with it (fallback off): per-system iterations and
    solve seconds, and their means."""
    net = LearnedStage(device=device, params=params)
    net.set_fallback(-1.0, 3)
    rows = []
    for k in range(systems):
        s = W.c5_box(n=n, seed=seed0 + k, device=device)
        h = amg_for(s, device, cheb_degree, cheb_ratio, cheb_degree_coarse)
        x = np.zeros(s.n)
        plain = h.solve_fgmres(s.b, x, 30, 1000, tol)
        x = np.zeros(s.n)
        learned = h.solve_fgmres_learned(s.b, x, net, 30, 1000, tol)
        rec = dict(seed=seed0 + k, plain_iterations=int(plain.iterations), learned_iterations=int(learned.iterations),
                   plain_s=plain.solve_seconds, learned_s=learned.solve_seconds,
                   ok=bool(plain.status == 0 and learned.status == 0))
        rows.append(rec)
        log(json.dumps(rec))
    pi = np.array([r["plain_iterations"] for r in rows], float)
    li = np.array([r["learned_iterations"] for r in rows], float)
    return dict(n=n, systems=systems, all_converged=all(r["ok"] for r in rows),
                mean_plain_iterations=float(pi.mean()), mean_learned_iterations=float(li.mean()),
                systems_fewer_iterations=int((li < pi).sum()), systems_more_iterations=int((li > pi).sum()),
                mean_plain_s=float(np.mean([r["plain_s"] for r in rows])),
                mean_learned_s=float(np.mean([r["learned_s"] for r in rows])), rows=rows)


def save(params, meta, out):
    os.makedirs(os.path.dirname(out), exist_ok=True)
    np.save(out, params.astype(np.float32))
    with open(os.path.splitext(out)[0] + ".json", "w") as f:
        json.dump(meta, f, indent=1)


def load(path=DEFAULT_OUT):
    """(params, meta) of a saved stage, or None when absent."""
    if not os.path.exists(path):
        return None
    p = np.load(path, allow_pickle=False)
    with open(os.path.splitext(path)[0] + ".json") as f:
        meta = json.load(f)
    return p, meta


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=48)
    ap.add_argument("--systems", type=int, default=12)
    ap.add_argument("--steps", type=int, default=120)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--cheb-degree", type=int, default=3)
    ap.add_argument("--cheb-ratio", type=float, default=30.0)
    ap.add_argument("--cheb-degree-coarse", type=int, default=0, help="levels below the finest (0: --cheb-degree)")
    ap.add_argument("--validate-n", type=int, default=0, help="validate on unseen boxes of this size (0: skip)")
    ap.add_argument("--validate-systems", type=int, default=80)
    ap.add_argument("--out", default=DEFAULT_OUT)
    a = ap.parse_args()
    params, meta = train(a.n, a.systems, a.steps, a.lr, a.seed, cheb_degree=a.cheb_degree, cheb_ratio=a.cheb_ratio,
                         cheb_degree_coarse=a.cheb_degree_coarse)
    if a.validate_n:
        v = validate(params, a.validate_n, a.validate_systems, cheb_degree=a.cheb_degree, cheb_ratio=a.cheb_ratio,
                     cheb_degree_coarse=a.cheb_degree_coarse)
        meta["validation"] = {k: val for k, val in v.items() if k != "rows"}
        meta["validation_rows"] = v["rows"]
        print(json.dumps(meta["validation"]))
    save(params, meta, a.out)
    print(f"saved {a.out} (plain mixed method: {meta['plain_rate']:.4f} decades/iteration on the training boxes)")


if __name__ == "__main__":
    main()
