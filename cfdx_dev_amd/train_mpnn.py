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


def amg_for(s, device=0, cheb_degree=3, cheb_ratio=30.0):
    h = NativeAMGPreconditioner(precision="fp32", prolongator_filter=0.2, filter_keeps_aggregate=True,
                                smoother_fp32="chebyshev", chebyshev_degree=cheb_degree, chebyshev_ratio=cheb_ratio,
                                device=device)
    if not h.setup(s.n, s.rowp, s.cols, s.vals):
        raise RuntimeError(h.last_error())
    return h


def train(n=48, systems=12, steps=120, lr=2e-3, seed=0, device=0, tol=1e-8, cheb_degree=3, cheb_ratio=30.0,
          log=print):
    """Trains against the FP32 V-cycle it will run with (the Chebyshev degree
    and ratio matter: a stage trained for one V-cycle can hurt another)."""
    net = LearnedStage(device=device, params=LearnedStage.initial_params(seed))
    rates, history = [], []
    t0 = time.perf_counter()
    for k in range(systems):
        s = W.c5_box(n=n, seed=1000 + k, device=device)
        h = amg_for(s, device, cheb_degree, cheb_ratio)
        x = np.zeros(s.n)
        plain = h.solve_fgmres(s.b, x, 30, 1000, tol)
        if plain.status == 0 and plain.iterations > 0:
            rates.append(-np.log10(plain.residual_relative) / plain.iterations)
        got, res = net.collect_pairs(h, s.b, PAIR_ITERATIONS, tol=1e-10)
        losses = net.train(h, steps, lr)
        rec = dict(system=k, seed=1000 + k, pairs=got, pcg_iterations=int(res.iterations),
                   plain_fgmres_iterations=int(plain.iterations), loss_first=float(losses[:8].mean()),
                   loss_last=float(losses[-8:].mean()), seconds=round(time.perf_counter() - t0, 2))
        history.append(rec)
        log(json.dumps(rec))
    meta = dict(n=n, systems=systems, steps=steps, lr=lr, seed=seed, pair_iterations=list(PAIR_ITERATIONS),
                cheb_degree=cheb_degree, cheb_ratio=cheb_ratio,
                plain_rate=float(np.mean(rates)) if rates else 0.0, history=history)
    return net.get_params(), meta


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
    ap.add_argument("--out", default=DEFAULT_OUT)
    a = ap.parse_args()
    params, meta = train(a.n, a.systems, a.steps, a.lr, a.seed)
    save(params, meta, a.out)
    print(f"saved {a.out} (plain mixed method: {meta['plain_rate']:.4f} decades/iteration on the training boxes)")


if __name__ == "__main__":
    main()
