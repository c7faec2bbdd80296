"""Pins the CPU restatement of the learned stage (no reference counterpart):
its backward pass against central finite differences of its own forward pass,
its equivariances, and Adam against the closed-form first step."""
import numpy as np

import mpnn_oracle as M
from systems import box_mesh, pressure_system


def _graph(oracle, dims=(5, 4, 3), seed=2):
    s = pressure_system(oracle, box_mesh(*dims, jitter=0.15, seed=seed), sigma=1.0, seed=seed)
    return s, M.Graph(s["rowp"], s["cols"], s["vals"])


def test_vjp_matches_finite_differences(oracle):
    s, g = _graph(oracle)
    rng = np.random.default_rng(0)
    p = M.init_params(1, out_scale=1.0).astype(np.float64)
    # biases away from zero so relu kinks are not hit by the probe steps
    p += 0.05 * rng.standard_normal(M.NPARAMS)
    r = rng.standard_normal(g.n)
    ge = rng.standard_normal(g.n)
    grad = M.vjp(g, p, r, ge)
    idx = np.concatenate([np.arange(0, 64), rng.choice(M.NPARAMS, 120, replace=False),
                          np.arange(M.OFF_WOUT, M.NPARAMS)])
    h = 1e-6
    for k in idx:
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        fd = (ge @ M.forward(g, pp, r) - ge @ M.forward(g, pm, r)) / (2 * h)
        assert abs(fd - grad[k]) <= 1e-5 * max(1.0, abs(fd)), (k, fd, grad[k])


def test_scale_equivariance(oracle):
    """e(c r) = c e(r) for c > 0 (rho normalisation), and e(0) = 0."""
    s, g = _graph(oracle)
    p = M.init_params(3, out_scale=1.0)
    r = np.random.default_rng(1).standard_normal(g.n)
    np.testing.assert_allclose(M.forward(g, p, 7.5 * r), 7.5 * M.forward(g, p, r), rtol=1e-12, atol=1e-14)
    assert np.all(M.forward(g, p, np.zeros(g.n)) == 0.0)


def test_message_passing_is_local(oracle):
    """Four layers see exactly four hops: flipping the sign of r on a node
    more than four hops away (rho unchanged) leaves e_0 bit-identical, while
    a flip two hops away changes it."""
    import scipy.sparse as sp
    from scipy.sparse import csgraph

    s, g = _graph(oracle, dims=(12, 3, 3))
    p = M.init_params(4, out_scale=1.0)
    r = np.random.default_rng(2).standard_normal(g.n)
    A = sp.csr_matrix((np.ones(len(s["cols"])), s["cols"], s["rowp"]), shape=(g.n, g.n))
    dist = csgraph.shortest_path(A, unweighted=True, indices=0)
    e1 = M.forward(g, p, r)
    far, near = np.where(dist > 4)[0], np.where(dist == 2)[0]
    assert len(far) and len(near)
    r2 = r.copy()
    r2[far[0]] *= -1.0
    assert M.forward(g, p, r2)[0] == e1[0]
    r3 = r.copy()
    r3[near[0]] *= -1.0
    assert M.forward(g, p, r3)[0] != e1[0]


def test_adam_first_step_closed_form():
    rng = np.random.default_rng(3)
    p = rng.standard_normal(50)
    g = rng.standard_normal(50)
    lr, eps = 1e-3, 1e-8
    p1, m, v = M.adam(p, np.zeros(50), np.zeros(50), g, 1, lr, eps=eps)
    np.testing.assert_allclose(p1, p - lr * g / (np.abs(g) + eps), rtol=0, atol=1e-15)
    # second step with the same gradient: m_hat = g, v_hat = g^2 again
    p2, _, _ = M.adam(p1, m, v, g, 2, lr, eps=eps)
    np.testing.assert_allclose(p2, p1 - lr * g / (np.abs(g) + eps), rtol=0, atol=1e-15)
