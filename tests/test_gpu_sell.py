"""SELL-64 storage of the device SpMV.

Matrices with at least 4096 rows whose 64-row slices pad at most 25 % are
streamed from a sliced-ELL copy. The FP64 copy keeps each row's CSR entry
order, so y = A x must stay bit-identical to the reference's scalar CSR loop
(sparse_matrix.h:182-197, restated by the oracle); the FP32 copy sorts each
row's slots by column (a different summation order), held to an FP32
rounding bound. Cases: a pressure Laplacian whose row count is not a
multiple of 64 (partial last slice) with ragged boundary rows, a matrix with
a few long rows (too much padding: stays CSR), and a numeric refresh that
must rewrite the SELL values.
"""
import ctypes as C

import numpy as np
import pytest

import cfdx_dev_amd as cx
from systems import box_mesh, pressure_system

pytestmark = pytest.mark.gpu


def _hip():
    try:
        hip = C.CDLL("libamdhip64.so.7")
    except OSError:
        hip = C.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipDeviceSynchronize.argtypes = []
    return hip


def _device_spmv(h, x, precision):
    """y = A x with the handle's finest-level kernel, caller ordering."""
    hip = _hip()
    dt = np.float64 if precision == 0 else np.float32
    xs = np.ascontiguousarray(x, dtype=dt)
    y = np.empty_like(xs)
    dx, dy = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(dx), xs.nbytes) == 0
    assert hip.hipMalloc(C.byref(dy), y.nbytes) == 0
    try:
        assert hip.hipMemcpy(dx, xs.ctypes.data_as(C.c_void_p), xs.nbytes, 1) == 0
        sec = C.c_double()
        st = h._lib.cfdxhip_amg_spmv_device(h._h, precision, dx, dy, 1, C.byref(sec))
        assert st == 0, h.last_error()
        assert hip.hipDeviceSynchronize() == 0
        assert hip.hipMemcpy(y.ctypes.data_as(C.c_void_p), dy, y.nbytes, 2) == 0
    finally:
        hip.hipFree(dx)
        hip.hipFree(dy)
    return y


def _check(oracle, h, rowp, cols, vals, seed, fp32=True):
    rng = np.random.default_rng(seed)
    n = len(rowp) - 1
    x = rng.normal(size=n)
    y = oracle.spmv(rowp, cols, vals, x)
    y64 = _device_spmv(h, x, 0)
    assert np.array_equal(y, y64), f"FP64 max |dy| = {np.abs(y - y64).max()}"
    if not fp32:
        return
    # FP32: values and x rounded to FP32, any summation order. Bound per row:
    # (len + 2) * 2^-24 * sum |a_ij x_j| covers the two input roundings and
    # the len - 1 additions of an arbitrary order.
    x32 = x.astype(np.float32)
    y32 = _device_spmv(h, x32, 1).astype(np.float64)
    absrow = oracle.spmv(rowp, cols, np.abs(vals), np.abs(x32.astype(np.float64)))
    lens = np.diff(rowp).astype(np.float64)
    bound = (lens + 2.0) * 2.0 ** -24 * absrow * 1.01 + 1e-300
    err = np.abs(y32 - y)
    assert np.all(err <= bound), f"FP32 worst err/bound = {(err / bound).max()}"


def _handle(s, precision="fp32"):
    h = cx.NativeAMGPreconditioner(method="sa", precision=precision)
    assert h.setup(s["n"], s["rowp"], s["cols"], s["vals"]), h.last_error()
    return h


def test_sell_partial_slice_ragged_rows(oracle):
    # 22 x 20 x 19 cells = 8360, one pinned cell eliminated: 8359 rows =
    # 130 full slices + 39 rows; boundary rows hold 4-6 entries, interior 7.
    s = pressure_system(oracle, box_mesh(22, 20, 19, jitter=0.15, seed=11), sigma=1.0, seed=11)
    assert s["n"] % 64 != 0
    h = _handle(s)
    fmt, stored, slices = h.level_spmv_format(0)
    assert fmt == "sell64"
    assert slices == (s["n"] + 63) // 64
    assert len(s["cols"]) <= stored <= 1.25 * len(s["cols"])
    # value + 4 B column a slot, 4 B a slice offset
    assert h.level_spmv_bytes(0, "fp64") == stored * 12 + 4 * (slices + 1)
    assert h.level_spmv_bytes(0, "fp32") == stored * 8 + 4 * (slices + 1)
    _check(oracle, h, s["rowp"], s["cols"], s["vals"], seed=1)


def _with_hubs(s, hubs, width, seed):
    """Adds `hubs` rows coupled (symmetrically, weakly) to `width` other rows
    each: their slices pad far beyond 25 %."""
    n = s["n"]
    rng = np.random.default_rng(seed)
    A = {}
    rp, cl, vl = s["rowp"], s["cols"], s["vals"]
    for i in range(n):
        for k in range(rp[i], rp[i + 1]):
            A[(i, int(cl[k]))] = float(vl[k])
    for hub in rng.choice(n, size=hubs, replace=False):
        others = rng.choice(n, size=width, replace=False)
        for j in others:
            j = int(j)
            if j == hub or (int(hub), j) in A:
                continue
            w = 1e-3 * abs(A[(int(hub), int(hub))])
            A[(int(hub), j)] = -w
            A[(j, int(hub))] = -w
            A[(int(hub), int(hub))] += w
            A[(j, j)] += w
    rows = [[] for _ in range(n)]
    for (i, j), v in A.items():
        rows[i].append((j, v))
    rowp = np.zeros(n + 1, np.uint32)
    cols, vals = [], []
    for i in range(n):
        rows[i].sort()
        cols += [j for j, _ in rows[i]]
        vals += [v for _, v in rows[i]]
        rowp[i + 1] = len(cols)
    return dict(s, rowp=rowp, cols=np.array(cols, np.uint32), vals=np.array(vals))


def test_heavy_padding_keeps_csr(oracle):
    base = pressure_system(oracle, box_mesh(18, 18, 16, jitter=0.1, seed=12), sigma=0.5, seed=12)
    s = _with_hubs(base, hubs=80, width=120, seed=3)
    h = _handle(s)
    fmt, stored, slices = h.level_spmv_format(0)
    assert fmt == "csr" and stored == len(s["cols"]) and slices == 0
    # CSR stream: value + 4 B column a nonzero, 4 B a row offset
    assert h.level_spmv_bytes(0, "fp64") == stored * 12 + 4 * (s["n"] + 1)
    assert h.level_spmv_bytes(0, "fp32") == stored * 8 + 4 * (s["n"] + 1)
    _check(oracle, h, s["rowp"], s["cols"], s["vals"], seed=2)


def test_refresh_rewrites_sell_values(oracle):
    """update_values on a SELL level: the V-cycle stays bit-identical to the
    oracle's and y = A x uses the new values."""
    s = pressure_system(oracle, box_mesh(20, 18, 16, jitter=0.15, seed=13), sigma=1.0, seed=13)
    ao = oracle.amg("sa")
    assert ao.setup(s["n"], s["rowp"], s["cols"], s["vals"])
    ap = _handle(s, "fp64")   # FP64 V-cycle: comparable bit for bit
    am = _handle(s, "fp32")   # FP32 hierarchy: both SELL copies
    assert ap.level_spmv_format(0)[0] == "sell64" and am.level_spmv_format(0)[0] == "sell64"
    vals2 = s["vals"] * 2.0
    assert ao.update_values(vals2)
    assert ap.update_values(vals2), ap.last_error()
    assert am.update_values(vals2), am.last_error()
    r = np.random.default_rng(4).normal(size=s["n"])
    assert np.array_equal(ao.apply(r), ap.apply(r))
    _check(oracle, ap, s["rowp"], s["cols"], vals2, seed=3, fp32=False)
    _check(oracle, am, s["rowp"], s["cols"], vals2, seed=3)
