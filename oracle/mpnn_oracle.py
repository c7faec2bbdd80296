"""CPU FP64 restatement of the learned preconditioner stage (test infrastructure).

The oracle is the checker for the HIP message-passing network: forward pass,
vector-Jacobian product (backward) and the Adam update, in float64 numpy.
Only tests, the smoke check and the benchmark's CPU leg may use it.

The reference has no learned component, so parity of this stage is
"unpinned" with respect to the reference. The oracle itself is pinned by
central finite differences of its own forward pass (tests/test_mpnn_oracle.py)
and Adam by its closed-form first step.

Model (hidden width 32, 4 message-passing layers), on the graph of a
symmetric matrix A with positive diagonal d:
    s_i  = 1 / sqrt(d_i),  w_ij = -a_ij s_i s_j  (j != i)
    u_i  = r_i s_i / rho,  rho = sqrt(mean_i (r_i s_i)^2)
    h0_i = relu(u_i w_in + b_in)
    layer l = 1..4:
        m_i  = sum_j w_ij h_j
        z_i  = W_l^T [h_i ; m_i] + b_l          (W_l: 64 x 32)
        h_i <- relu(z_i) + h_i
    y_i  = h_i . w_out + b_out
    e_i  = rho s_i y_i
Flat parameter layout (float32 on the device, 8417 values):
    w_in[32] @0, b_in[32] @32, layer l: W_l[64][32] row-major @64 + 2080 l,
    b_l[32] @64 + 2080 l + 2048, w_out[32] @8384, b_out @8416.
"""
from __future__ import annotations

import numpy as np

H = 32
LAYERS = 4
LAYER_STRIDE = 2 * H * H + H
OFF_WOUT = 2 * H + LAYERS * LAYER_STRIDE
NPARAMS = OFF_WOUT + H + 1


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    w_in, b_in = p[0:H], p[H:2 * H]
    Ws, bs = [], []
    for l in range(LAYERS):
        o = 2 * H + l * LAYER_STRIDE
        Ws.append(p[o:o + 2 * H * H].reshape(2 * H, H))
        bs.append(p[o + 2 * H * H:o + LAYER_STRIDE])
    return w_in, b_in, Ws, bs, p[OFF_WOUT:OFF_WOUT + H], p[OFF_WOUT + H]


def init_params(seed=0, out_scale=0.1):
    """Seeded initialisation: w_in ~ N(0,1), W_l ~ U(+-sqrt(6/96)) / 2,
    w_out ~ out_scale N(0, 1/32), biases 0."""
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAMS)
    p[0:H] = rng.standard_normal(H)
    lim = np.sqrt(6.0 / (3 * H)) / 2
    for l in range(LAYERS):
        o = 2 * H + l * LAYER_STRIDE
        p[o:o + 2 * H * H] = rng.uniform(-lim, lim, 2 * H * H)
    p[OFF_WOUT:OFF_WOUT + H] = out_scale * rng.standard_normal(H) / np.sqrt(H)
    return p.astype(np.float32)


class Graph:
    def __init__(self, rowp, cols, vals):
        self.n = len(rowp) - 1
        rows = np.repeat(np.arange(self.n), np.diff(rowp))
        cols = np.asarray(cols, dtype=np.int64)
        vals = np.asarray(vals, dtype=np.float64)
        diag = np.zeros(self.n)
        on = cols == rows
        diag[rows[on]] = vals[on]
        self.s = 1.0 / np.sqrt(diag)
        off = ~on
        self.rows, self.cols = rows[off], cols[off]
        self.w = -vals[off] * self.s[self.rows] * self.s[self.cols]

    def mix(self, h):
        """m_i = sum_j w_ij h_j (rows of h are nodes)."""
        out = np.zeros_like(h)
        np.add.at(out, self.rows, self.w[:, None] * h[self.cols])
        return out


def bf16(a):
    """Round to the nearest BF16 value (ties to even), returned as float64."""
    x = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    x = (x + 0x7FFF + ((x >> 16) & 1)) & 0xFFFF0000
    return x.astype(np.uint32).view(np.float32).astype(np.float64)


def forward(g: Graph, p, r, keep=False, emulate_bf16=False):
    """e = N(r). With emulate_bf16 the values the device keeps in BF16 are
    rounded at the same points: hidden states, the message half of each
    layer input, and the layer weights; everything else stays float64."""
    R = bf16 if emulate_bf16 else (lambda a: a)
    w_in, b_in, Ws, bs, w_out, b_out = unpack(p)
    rs = np.asarray(r, dtype=np.float64) * g.s
    rho = np.sqrt(np.mean(rs * rs))
    if rho == 0.0:
        return (np.zeros(g.n), None) if keep else np.zeros(g.n)
    u = rs / rho
    z0 = u[:, None] * w_in[None, :] + b_in[None, :]
    h = R(np.maximum(z0, 0.0))
    tape = dict(u=u, rho=rho, z0=z0, hs=[h], zs=[], ms=[], R=R)
    for l in range(LAYERS):
        m = R(g.mix(h))
        z = np.concatenate([h, m], axis=1) @ R(Ws[l]) + bs[l][None, :]
        h = R(np.maximum(z, 0.0) + h)
        tape["ms"].append(m)
        tape["zs"].append(z)
        tape["hs"].append(h)
    y = h @ w_out + b_out
    e = rho * g.s * y
    return (e, tape) if keep else e


def vjp(g: Graph, p, r, ge, emulate_bf16=False):
    """Gradient of sum_i ge_i e_i(p) with respect to the flat parameters.
    With emulate_bf16 the upstream gradient of each layer's pre-activation is
    rounded to BF16 where the device feeds it to the matrix cores; the exact
    derivative of the BF16-rounded network is otherwise not defined, so this
    mode is for comparing with the device, the plain mode for finite
    differences."""
    w_in, b_in, Ws, bs, w_out, b_out = unpack(p)
    e, t = forward(g, p, r, keep=True, emulate_bf16=emulate_bf16)
    grad = np.zeros(NPARAMS)
    if t is None:
        return grad
    R = t["R"]
    gy = np.asarray(ge, dtype=np.float64) * t["rho"] * g.s
    hL = t["hs"][-1]
    grad[OFF_WOUT:OFF_WOUT + H] = hL.T @ gy
    grad[OFF_WOUT + H] = gy.sum()
    gh = gy[:, None] * w_out[None, :]
    for l in reversed(range(LAYERS)):
        o = 2 * H + l * LAYER_STRIDE
        hp, m, z = t["hs"][l], t["ms"][l], t["zs"][l]
        gz = gh * (z > 0.0)
        GZ = R(gz)
        X = np.concatenate([hp, m], axis=1)
        grad[o:o + 2 * H * H] = (X.T @ GZ).ravel()
        grad[o + 2 * H * H:o + LAYER_STRIDE] = gz.sum(0)
        gX = GZ @ R(Ws[l]).T
        gh = gh + gX[:, :H] + g.mix(gX[:, H:])  # A symmetric: mix^T = mix
    gz0 = gh * (t["z0"] > 0.0)
    grad[0:H] = t["u"] @ gz0
    grad[H:2 * H] = gz0.sum(0)
    return grad


def adam(p, m, v, g, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (Kingma & Ba 2015, Algorithm 1); t counts from 1."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v
