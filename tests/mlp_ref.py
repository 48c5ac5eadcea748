"""A numpy restatement of the bucket-size MLP (cuttlefish-private_amd/csrc/
mlp.hip): N_FEAT -> 64 -> 64 -> 1 with ReLU, inputs standardised by the
training set's mean and population std, He-uniform weights from a seeded
hash, zero biases but the output bias (the mean target), full-batch mean
squared error, Adam with decoupled weight decay on the weight matrices.
Test infrastructure: the checker of the device kernels, in float64 unless
asked otherwise; its backward pass is pinned by finite differences
(tests/test_mlp_cpu.py)."""
import numpy as np

H = 64
M64 = (1 << 64) - 1


def fmix64(h):
    h ^= h >> 33
    h = (h * 0xFF51AFD7ED558CCD) & M64
    h ^= h >> 33
    h = (h * 0xC4CEB9FE1A85EC53) & M64
    h ^= h >> 33
    return h


def urand(seed, i):
    return float(fmix64(((seed << 32) ^ i ^ 0x9E3779B97F4A7C15) & M64) >> 40) / 16777216.0


def layout(F):
    o = {}
    o["W1"] = 0
    o["b1"] = o["W1"] + H * F
    o["W2"] = o["b1"] + H
    o["b2"] = o["W2"] + H * H
    o["w3"] = o["b2"] + H
    o["b3"] = o["w3"] + H
    o["mean"] = o["b3"] + 1
    o["istd"] = o["mean"] + F
    o["end"] = o["istd"] + F
    return o


def unpack(P, F):
    o = layout(F)
    return (P[o["W1"]:o["b1"]].reshape(H, F), P[o["b1"]:o["W2"]], P[o["W2"]:o["b2"]].reshape(H, H),
            P[o["b2"]:o["w3"]], P[o["w3"]:o["b3"]], P[o["b3"]], P[o["mean"]:o["istd"]], P[o["istd"]:o["end"]])


def init(F, seed, dtype=np.float64):
    o = layout(F)
    P = np.zeros(o["end"], dtype)
    for i in range(o["b1"]):
        P[i] = (2 * urand(seed, i) - 1) * np.sqrt(6.0 / F)
    for i in range(o["W2"], o["b2"]):
        P[i] = (2 * urand(seed, i) - 1) * np.sqrt(6.0 / H)
    for i in range(o["w3"], o["b3"]):
        P[i] = (2 * urand(seed, i) - 1) * np.sqrt(3.0 / H)
    return P


def standardise(X):
    m = X.mean(axis=0)
    sd = np.sqrt(((X - m) ** 2).mean(axis=0))
    return m, np.where(sd > 1e-6, 1.0 / np.where(sd > 1e-6, sd, 1.0), 0.0)


def forward(P, F, X):
    """Predictions for raw feature rows X (standardised with P's mean / std)."""
    W1, b1, W2, b2, w3, b3, m, istd = unpack(P, F)
    Z = (X - m) * istd
    h1 = np.maximum(Z @ W1.T + b1, 0)
    h2 = np.maximum(h1 @ W2.T + b2, 0)
    return h2 @ w3 + b3


def loss_and_grad(P, F, Z, T):
    """Mean squared error on standardised rows Z and its gradient w.r.t. the
    network parameters (the standardisation entries get zero)."""
    W1, b1, W2, b2, w3, b3, _, _ = unpack(P, F)
    n = len(T)
    z1 = Z @ W1.T + b1
    h1 = np.maximum(z1, 0)
    z2 = h1 @ W2.T + b2
    h2 = np.maximum(z2, 0)
    y = h2 @ w3 + b3
    d = y - T
    dy = 2 * d / n
    G = np.zeros_like(P)
    o = layout(F)
    G[o["w3"]:o["b3"]] = dy @ h2
    G[o["b3"]] = dy.sum()
    dz2 = np.outer(dy, w3) * (z2 > 0)
    G[o["W2"]:o["b2"]] = (dz2.T @ h1).reshape(-1)
    G[o["b2"]:o["w3"]] = dz2.sum(axis=0)
    dz1 = (dz2 @ W2) * (z1 > 0)
    G[o["W1"]:o["b1"]] = (dz1.T @ Z).reshape(-1)
    G[o["b1"]:o["W2"]] = dz1.sum(axis=0)
    return float((d * d).mean()), G


def train(X, T, seed, epochs, lr, dtype=np.float64):
    """Full-batch training on raw feature rows X and targets T; returns the
    parameters (with the standardisation) and the last epoch's loss."""
    n, F = X.shape
    o = layout(F)
    P = init(F, seed, dtype)
    m, istd = standardise(X.astype(dtype))
    P[o["mean"]:o["istd"]] = m
    P[o["istd"]:o["end"]] = istd
    Z = (X.astype(dtype) - m) * istd
    P[o["b3"]] = T.astype(dtype).mean()
    is_w = np.zeros(o["end"], bool)
    is_w[o["W1"]:o["b1"]] = is_w[o["W2"]:o["b2"]] = is_w[o["w3"]:o["b3"]] = True
    net = slice(0, o["mean"])
    am = np.zeros(o["mean"], dtype)
    av = np.zeros(o["mean"], dtype)
    b1, b2, eps, wd = 0.9, 0.999, 1e-8, 1e-3
    b1t = b2t = 1.0
    loss = 0.0
    for _ in range(epochs):
        loss, G = loss_and_grad(P, F, Z, T.astype(dtype))
        g = G[net]
        b1t *= b1
        b2t *= b2
        am = b1 * am + (1 - b1) * g
        av = b2 * av + (1 - b2) * g * g
        upd = (am / (1 - b1t)) / (np.sqrt(av / (1 - b2t)) + eps) + np.where(is_w[net], wd * P[net], 0)
        P[net] = P[net] - lr * upd
    return P, loss
