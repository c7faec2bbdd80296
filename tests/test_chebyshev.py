"""Chebyshev smoother variant (no reference counterpart inside the V-cycle).

Pinned without the reference:
  * the recurrence coefficients against the closed-form error polynomial
    T_k((theta - lam)/delta) / T_k(theta/delta) on the damped interval;
  * a two-level V-cycle with an exact (LU) coarse solve against its closed
    form (I - E) A^-1 r, E = p(D^-1 A) (I - P Ac^-1 P^T A) p(D^-1 A);
  * the lambda_max estimate (20 Lanczos steps) as a lower bound within 3 %
    of the true largest eigenvalue of D^-1 A;
  * the product's host estimate equal to the oracle's, bit for bit;
  * a four-level V-cycle with a separate degree below the finest level
    against its recursive closed form.
"""
import numpy as np
import pytest

import cfdx_dev_amd as cx
from systems import box_mesh, csr_to_dense, pressure_system


def cheb_poly(lam, hi, lo, k):
    """T_k((theta - lam)/delta) / T_k(theta/delta), T_k from numpy's
    Chebyshev series (valid inside and outside [-1, 1])."""
    from numpy.polynomial import chebyshev as T

    theta, delta = (hi + lo) / 2, (hi - lo) / 2
    c = [0.0] * k + [1.0]
    return T.chebval((theta - lam) / delta, c) / T.chebval(theta / delta, c)


@pytest.mark.parametrize("degree", [1, 2, 3, 5])
def test_coefficients_realise_chebyshev_polynomial(oracle, degree):
    """Run the recurrence on scalar modes x' = lam x: the error factor after k
    steps must be the scaled Chebyshev polynomial on [hi/ratio, hi]."""
    lam_max, ratio = 1.9, 30.0
    keep, scale = oracle.chebyshev_coefficients(lam_max, ratio, degree)
    hi = 1.1 * lam_max
    lo = hi / ratio
    lam = np.concatenate([np.linspace(lo, hi, 41), np.linspace(1e-3, lo, 7), [1.2 * hi]])
    # error recursion e_{k+1} = e_k - d_k, d_k = keep_k d_{k-1} + scale_k lam e_k
    e = np.ones_like(lam)
    d = np.zeros_like(lam)
    for k in range(degree):
        d = (keep[k] * d if k else 0.0) + scale[k] * lam * e
        e = e - d
    np.testing.assert_allclose(e, cheb_poly(lam, hi, lo, degree), rtol=0, atol=1e-13)
    # minimax bound on the damped interval
    assert np.max(np.abs(e[:41])) <= 1.0 / np.cosh(degree * np.arccosh((hi + lo) / (hi - lo))) + 1e-13


def small_system(oracle):
    m = box_mesh(4, 4, 3, jitter=0.15, seed=3)
    return pressure_system(oracle, m, sigma=1.0, seed=3)


def test_two_level_vcycle_closed_form(oracle):
    s = small_system(oracle)
    a = oracle.amg("sa", smoother="chebyshev", chebyshev_degree=3)
    assert a.setup(s["n"], s["rowp"], s["cols"], s["vals"])
    assert a.num_levels() == 2 and a.level_info(1)[0] <= 16
    n = s["n"]
    A = csr_to_dense(s["rowp"], s["cols"], s["vals"], n)
    _, prp, pc, pv = a.level_transfer(0)
    nc = a.level_info(1)[0]
    P = np.zeros((n, nc))
    for i in range(n):
        P[i, pc[prp[i]:prp[i + 1]]] = pv[prp[i]:prp[i + 1]]
    Ac = P.T @ A @ P
    d = np.diag(A)
    w, V = np.linalg.eigh(A / np.sqrt(np.outer(d, d)))
    lam = a.lambda_max(0)
    hi = 1.1 * lam
    pk = V @ np.diag(cheb_poly(w, hi, hi / 30.0, 3)) @ V.T
    S = np.diag(d ** -0.5) @ pk @ np.diag(d ** 0.5)  # p(D^-1 A)
    C = np.eye(n) - P @ np.linalg.solve(Ac, P.T @ A)
    E = S @ C @ S
    rng = np.random.default_rng(0)
    for _ in range(3):
        r = rng.normal(size=n)
        z = a.apply(r)
        expect = (np.eye(n) - E) @ np.linalg.solve(A, r)
        assert np.linalg.norm(z - expect) <= 1e-10 * np.linalg.norm(expect)


def _dense_level(a, l):
    n = a.level_info(l)[0]
    rp, cols, vals = a.level_matrix(l)
    return csr_to_dense(rp, cols, vals, n)


def _vcycle_operator(a, l, degrees, ratio):
    """Matrix M_l of the V-cycle from level l down (z = M_l r): the coarsest
    level (<= 16 rows) solved exactly; above it
    M_l = (I - E_l) A_l^-1, E_l = S_l (I - P_l M_{l+1} P_l^T A_l) S_l,
    S_l = p_l(D_l^-1 A_l) with p_l the scaled Chebyshev polynomial of degree
    degrees[l] on [1.1 lambda_l / ratio, 1.1 lambda_l]."""
    A = _dense_level(a, l)
    n = A.shape[0]
    if l + 1 == a.num_levels():
        assert n <= 16
        return np.linalg.inv(A)
    _, prp, pc, pv = a.level_transfer(l)
    nc = a.level_info(l + 1)[0]
    P = np.zeros((n, nc))
    for i in range(n):
        P[i, pc[prp[i]:prp[i + 1]]] = pv[prp[i]:prp[i + 1]]
    Mc = _vcycle_operator(a, l + 1, degrees, ratio)
    d = np.diag(A)
    w, V = np.linalg.eigh(A / np.sqrt(np.outer(d, d)))
    hi = 1.1 * a.lambda_max(l)
    S = np.diag(d ** -0.5) @ (V @ np.diag(cheb_poly(w, hi, hi / ratio, degrees[l])) @ V.T) @ np.diag(d ** 0.5)
    E = S @ (np.eye(n) - P @ Mc @ P.T @ A) @ S
    return (np.eye(n) - E) @ np.linalg.inv(A)


@pytest.mark.parametrize("degree,coarse", [(3, 1), (5, 2), (9, 1)])
def test_coarse_degree_vcycle_closed_form(oracle, degree, coarse):
    """A four-level V-cycle with `degree` Chebyshev steps on the finest level
    and `coarse` on the two levels below it, against the recursive closed
    form (a degree applied to the wrong level changes the operator)."""
    s = pressure_system(oracle, box_mesh(8, 8, 6, jitter=0.15, seed=3), sigma=1.0, seed=3)
    ratio = 30.0
    a = oracle.amg("sa", smoother="chebyshev", chebyshev_degree=degree, chebyshev_ratio=ratio,
                   chebyshev_degree_coarse=coarse)
    assert a.setup(s["n"], s["rowp"], s["cols"], s["vals"])
    L = a.num_levels()
    assert L == 4 and a.level_info(L - 1)[0] <= 16
    M = _vcycle_operator(a, 0, [degree] + [coarse] * (L - 1), ratio)
    M_uniform = _vcycle_operator(a, 0, [degree] * L, ratio)
    rng = np.random.default_rng(1)
    for _ in range(3):
        r = rng.normal(size=s["n"])
        z = a.apply(r)
        expect = M @ r
        assert np.linalg.norm(z - expect) <= 1e-10 * np.linalg.norm(expect)
        assert np.linalg.norm(z - M_uniform @ r) > 1e-6 * np.linalg.norm(expect)


def test_lambda_estimate_bounds(oracle):
    s = pressure_system(oracle, box_mesh(8, 8, 6, seed=4), sigma=1.0, seed=4)
    a = oracle.amg("sa", smoother="chebyshev")
    assert a.setup(s["n"], s["rowp"], s["cols"], s["vals"])
    for l in range(a.num_levels()):
        rp, c, v = a.level_matrix(l)
        n = len(rp) - 1
        A = csr_to_dense(rp, c, v, n)
        d = np.diag(A)
        true = np.linalg.eigvalsh(A / np.sqrt(np.outer(d, d)))[-1]
        est = a.lambda_max(l)
        assert est <= true * (1 + 1e-12), (l, est, true)
        assert est >= 0.97 * true, (l, est, true)


def test_host_lambda_matches_oracle(oracle):
    # 12479 unknowns: two 8192-row reduction blocks on the finest level
    s = pressure_system(oracle, box_mesh(26, 24, 20, seed=5), sigma=1.5, seed=5)
    a = oracle.amg("sa", smoother="chebyshev")
    assert a.setup(s["n"], s["rowp"], s["cols"], s["vals"])
    for threads in (1, 3):
        h = cx.NativeAMGPreconditioner(device=-1, smoother_fp32="chebyshev", host_threads=threads)
        assert h.setup(s["n"], s["rowp"], s["cols"], s["vals"]), h.last_error()
        assert h.num_levels() == a.num_levels()
        for l in range(a.num_levels()):
            assert h.level_lambda_max(l) == a.lambda_max(l), (threads, l)


def test_invalid_chebyshev_options():
    with pytest.raises(RuntimeError):
        cx.NativeAMGPreconditioner(device=-1, smoother="chebyshev", chebyshev_degree=0)
    with pytest.raises(RuntimeError):
        cx.NativeAMGPreconditioner(device=-1, smoother_fp32="chebyshev", chebyshev_ratio=1.0)
    with pytest.raises(RuntimeError):
        cx.NativeAMGPreconditioner(device=-1, smoother_fp32="chebyshev", chebyshev_degree_coarse=17)
    h = cx.NativeAMGPreconditioner(device=-1)
    with pytest.raises(RuntimeError):
        h.level_lambda_max(0)
