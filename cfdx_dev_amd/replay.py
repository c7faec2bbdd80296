"""Replays a captured linear system (.cfdxlin, written by the reference's
SIMPLE driver capture hook or by this library) through the GPU solver.

Usage: python -m cfdx_dev_amd.replay FILE [--solver cg|fgmres|mixed]
                                          [--method sa|boomer] [--filter F]
Prints one JSON line: status, iterations, FP64 true relative residual, solve
and setup seconds. The system's own max_iterations and tolerance are used.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import NativeAMGPreconditioner, read_capture


def replay(path, solver="cg", method="sa", prolongator_filter=0.0, keep_aggregate=False, device=0):
    cap = read_capture(path)
    mixed = solver == "mixed"
    h = NativeAMGPreconditioner(method=method, precision="fp32" if mixed else "fp64",
                                prolongator_filter=prolongator_filter, filter_keeps_aggregate=keep_aggregate,
                                smoother_fp32="chebyshev" if mixed else "jacobi", chebyshev_degree=9,
                                chebyshev_degree_coarse=1, chebyshev_ratio=240.0, device=device)
    t0 = time.perf_counter()
    if not h.setup(cap.n, cap.rowp, cap.cols, cap.vals):
        return dict(file=str(path), status="setup_failed", error=h.last_error())
    t_setup = time.perf_counter() - t0
    x = cap.initial_solution.copy()
    if solver == "cg":
        r = h.solve_cg(cap.rhs, x, cap.max_iterations, cap.tolerance)
    else:
        r = h.solve_fgmres(cap.rhs, x, 30, cap.max_iterations, cap.tolerance)
    return dict(file=str(path), n=cap.n, nnz=int(len(cap.vals)), problem=cap.problem, solver=solver,
                status=int(r.status), iterations=int(r.iterations), residual_relative=r.residual_relative,
                solve_s=r.solve_seconds, setup_s=t_setup, checksum=f"{cap.checksum:x}",
                solution_norm=float(np.linalg.norm(x)))


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("file")
    ap.add_argument("--solver", choices=["cg", "fgmres", "mixed"], default="cg")
    ap.add_argument("--method", choices=["sa", "boomer"], default="sa")
    ap.add_argument("--filter", type=float, default=0.0)
    ap.add_argument("--keep-aggregate", action="store_true")
    a = ap.parse_args(argv)
    out = replay(a.file, a.solver, a.method, a.filter, a.keep_aggregate)
    print(json.dumps(out))
    return 0 if out.get("status") == 0 else 1


if __name__ == "__main__":
    sys.exit(main())
