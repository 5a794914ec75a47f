"""ORACLE — test infrastructure only.  Nothing in the product imports this.

Float64 NumPy restatement of the learned cost model's network (include/bh_costmodel.h):
features standardised as (x - mean) * inv_std, two ReLU hidden layers and a linear
output, loss = mean((y - t)^2).  The flat parameter layout is the engine's:
W1[F][H] | b1[H] | W2[H][H] | b2[H] | w3[H] | b3.
The analytic gradient here is checked against central finite differences in the
tests, and the HIP kernels are checked against it.  The engine's cost model has no
counterpart in the reference (its fuser weight is analytic, core/jitk/graph.cpp:136-170),
so it is pinned by calculus, not by reference vectors.
"""
import numpy as np


def layout(F, H):
    o = {}
    o["W1"] = 0
    o["b1"] = o["W1"] + F * H
    o["W2"] = o["b1"] + H
    o["b2"] = o["W2"] + H * H
    o["w3"] = o["b2"] + H
    o["b3"] = o["w3"] + H
    o["n"] = o["b3"] + 1
    return o


def unpack(p, F, H):
    o = layout(F, H)
    p = np.asarray(p, np.float64)
    return (p[o["W1"]:o["b1"]].reshape(F, H), p[o["b1"]:o["W2"]], p[o["W2"]:o["b2"]].reshape(H, H),
            p[o["b2"]:o["w3"]], p[o["w3"]:o["b3"]], p[o["b3"]])


def forward(p, x, F, H, mean=None, inv_std=None):
    """Predictions and the activations (xn, a1, h1, a2, h2) for n x F inputs."""
    W1, b1, W2, b2, w3, b3 = unpack(p, F, H)
    x = np.asarray(x, np.float64)
    mean = np.zeros(F) if mean is None else np.asarray(mean, np.float64)
    inv_std = np.ones(F) if inv_std is None else np.asarray(inv_std, np.float64)
    xn = (x - mean) * inv_std
    a1 = xn @ W1 + b1
    h1 = np.maximum(a1, 0.0)
    a2 = h1 @ W2 + b2
    h2 = np.maximum(a2, 0.0)
    y = h2 @ w3 + b3
    return y, (xn, a1, h1, a2, h2)


def loss_and_gradient(p, x, t, F, H, mean=None, inv_std=None):
    """mean((y - t)^2) and its gradient in the flat layout (float64)."""
    W1, b1, W2, b2, w3, b3 = unpack(p, F, H)
    y, (xn, a1, h1, a2, h2) = forward(p, x, F, H, mean, inv_std)
    t = np.asarray(t, np.float64)
    n = len(t)
    e = y - t
    loss = float(np.mean(e * e))
    dy = 2.0 * e / n
    dh2 = np.outer(dy, w3) * (a2 > 0)
    dh1 = (dh2 @ W2.T) * (a1 > 0)
    g = np.concatenate([(xn.T @ dh1).ravel(), dh1.sum(0), (h1.T @ dh2).ravel(), dh2.sum(0), h2.T @ dy, [dy.sum()]])
    return loss, g


def finite_difference_gradient(p, x, t, F, H, mean=None, inv_std=None, eps=1e-6):
    p = np.asarray(p, np.float64)
    g = np.zeros_like(p)
    for k in range(len(p)):
        q = p.copy()
        q[k] += eps
        lp, _ = loss_and_gradient(q, x, t, F, H, mean, inv_std)
        q[k] -= 2 * eps
        lm, _ = loss_and_gradient(q, x, t, F, H, mean, inv_std)
        g[k] = (lp - lm) / (2 * eps)
    return g


def adam_steps(p, x, t, F, H, steps, lr, mean=None, inv_std=None):
    """Full-batch Adam (beta1 0.9, beta2 0.999, eps 1e-8) as the engine applies it; returns (params, losses)."""
    p = np.asarray(p, np.float64).copy()
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    losses = []
    for k in range(1, steps + 1):
        loss, g = loss_and_gradient(p, x, t, F, H, mean, inv_std)
        m = 0.9 * m + 0.1 * g
        v = 0.999 * v + 0.001 * g * g
        p -= lr * (m / (1 - 0.9 ** k)) / (np.sqrt(v / (1 - 0.999 ** k)) + 1e-8)
        losses.append(loss)
    return p, losses
