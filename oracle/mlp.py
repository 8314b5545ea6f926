"""CPU restatement of the surrogate network and its training loss (test
infrastructure only: the checker of the device kernels, never the product).

Network: inputs x (n x 8, five used) -> 128 -> 128 -> 128 -> 32 (three used),
SiLU after the three hidden layers; parameters in one flat vector laid out as
W1[128][8] b1[128] W2[128][128] b2 W3[128][128] b3 W4[32][128] b4[32]
(row = output unit). Loss of a batch of n rows with targets t (n x 8:
y_nu, v, theta, mask_v, mask_theta, tau, -, -): (o0 - y_nu)^2 +
mask_v (o1 - v)^2 + w_t mask_theta pinball_tau(theta - o2), summed over rows
(the gradient is of the mean).

`emulate_bf16` rounds what the device rounds to BF16 before a matrix product:
the weights W2..W4, the hidden activations, the output gradient and the
hidden-layer deltas (layer 1 and all accumulations stay FP32). With it off and
float64, the analytic gradient is pinned against central finite differences
(tests/test_mlp.py).

This network has no counterpart in the reference; the training targets it
learns come from nonlinear_compton/lp/mod.rs:478-676 and
nonlinear_compton/cp/mod.rs:181-234 (tables/mod.rs:46-121).
"""
import numpy as np

IN, H, OUT = 8, 128, 32
W1 = 0
B1 = W1 + H * IN
W2 = B1 + H
B2 = W2 + H * H
W3 = B2 + H
B3 = W3 + H * H
W4 = B3 + H
B4 = W4 + OUT * H
NPARAM = B4 + OUT


def bf16(a):
    """Round to BF16 (nearest even) and back to float32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def unpack(p):
    return (p[W1:B1].reshape(H, IN), p[B1:W2], p[W2:B2].reshape(H, H), p[B2:W3], p[W3:B3].reshape(H, H), p[B3:W4],
            p[W4:B4].reshape(OUT, H), p[B4:NPARAM])


def _sig(z):
    return 1.0 / (1.0 + np.exp(-z))


def silu(z):
    return z * _sig(z)


def silu_grad(z):
    s = _sig(z)
    return s * (1.0 + z * (1.0 - s))


def forward_backward(p, x, t, w_t=1.0, emulate_bf16=True, dtype=np.float32):
    """Returns (outputs n x 32, gradient of the mean loss, loss sums [3])."""
    r = bf16 if emulate_bf16 else (lambda a: a)
    p = np.asarray(p, dtype=dtype)
    x = np.asarray(x, dtype=dtype)
    t = np.asarray(t, dtype=dtype)
    w1, b1, w2, b2, w3, b3, w4, b4 = unpack(p)
    n = x.shape[0]
    z1 = x @ w1.T + b1
    a1 = r(silu(z1)).astype(dtype)
    z2 = a1 @ r(w2).astype(dtype).T + b2
    a2 = r(silu(z2)).astype(dtype)
    z3 = a2 @ r(w3).astype(dtype).T + b3
    a3 = r(silu(z3)).astype(dtype)
    o = a3 @ r(w4).astype(dtype).T + b4
    # loss
    e0 = o[:, 0] - t[:, 0]
    e1 = (o[:, 1] - t[:, 1]) * t[:, 3]
    rr = t[:, 2] - o[:, 2]
    tau, mt = t[:, 5], t[:, 4] * w_t
    loss = np.array([np.sum(e0 * e0), np.sum(e1 * e1),
                     np.sum(mt * np.where(rr >= 0, tau * rr, (tau - 1.0) * rr))])
    d_o = np.zeros((n, OUT), dtype=dtype)
    d_o[:, 0] = 2.0 * e0 / n
    d_o[:, 1] = 2.0 * e1 / n
    d_o[:, 2] = mt * np.where(rr > 0, -tau, 1.0 - tau) / n
    d_ob = r(d_o).astype(dtype)
    g = np.zeros(NPARAM, dtype=dtype)
    g[W4:B4] = (d_ob.T @ a3).ravel()
    g[B4:NPARAM] = d_o.sum(0)
    d3 = (d_ob @ r(w4).astype(dtype)) * silu_grad(z3)
    g[B3:W4] = d3.sum(0)
    d3b = r(d3).astype(dtype)
    g[W3:B3] = (d3b.T @ a2).ravel()
    d2 = (d3b @ r(w3).astype(dtype)) * silu_grad(z2)
    g[B2:W3] = d2.sum(0)
    d2b = r(d2).astype(dtype)
    g[W2:B2] = (d2b.T @ a1).ravel()
    d1 = (d2b @ r(w2).astype(dtype)) * silu_grad(z1)
    g[B1:W2] = d1.sum(0)
    g[W1:B1] = (d1.T @ x).ravel()
    return o, g, loss


def mean_loss(p, x, t, w_t=1.0):
    """The mean loss in float64 without rounding (for finite differences)."""
    _, _, loss = forward_backward(p, x, t, w_t, emulate_bf16=False, dtype=np.float64)
    return loss.sum() / x.shape[0]


def init_params(seed=0):
    """A random parameter vector shaped like the device's initialisation."""
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAM)
    p[W1:B1] = rng.uniform(-1, 1, H * IN) * np.sqrt(6 / 5)
    p[W1:B1].reshape(H, IN)[:, 5:] = 0.0
    p[B1:W2] = rng.normal(0, 0.1, H)
    p[W2:B2] = rng.uniform(-1, 1, H * H) * np.sqrt(6 / H)
    p[B2:W3] = rng.normal(0, 0.1, H)
    p[W3:B3] = rng.uniform(-1, 1, H * H) * np.sqrt(6 / H)
    p[B3:W4] = rng.normal(0, 0.1, H)
    p[W4:W4 + 3 * H] = rng.uniform(-1, 1, 3 * H) * np.sqrt(6 / H)
    p[B4:B4 + 3] = rng.normal(0, 0.1, 3)
    return p
