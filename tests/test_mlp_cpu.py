"""The numpy MLP restatement (mlp_ref.py) that checks the device training
kernel: its backward pass against central finite differences of its loss, and
its initialiser's hash against known values of the splitmix-style finaliser."""
import numpy as np

import mlp_ref as R


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(3)
    F, n = 19, 40
    P = R.init(F, 777)
    o = R.layout(F)
    P[o["b1"]:o["W2"]] = rng.normal(0, 0.3, R.H)  # biases off zero, so ReLU gates vary
    P[o["b2"]:o["w3"]] = rng.normal(0, 0.3, R.H)
    Z = rng.normal(0, 1, (n, F))
    T = rng.normal(-1, 0.5, n)
    _, G = R.loss_and_grad(P, F, Z, T)
    idx = np.concatenate([rng.choice(np.arange(o["W1"], o["b1"]), 40, replace=False),
                          rng.choice(np.arange(o["W2"], o["b2"]), 40, replace=False),
                          np.arange(o["b1"], o["b1"] + 10), np.arange(o["b2"], o["b2"] + 10),
                          np.arange(o["w3"], o["b3"] + 1)])
    h = 1e-6
    for i in idx:
        Pp, Pm = P.copy(), P.copy()
        Pp[i] += h
        Pm[i] -= h
        num = (R.loss_and_grad(Pp, F, Z, T)[0] - R.loss_and_grad(Pm, F, Z, T)[0]) / (2 * h)
        assert abs(num - G[i]) <= 1e-6 + 1e-5 * abs(num), (i, num, G[i])


def test_planted_backward_mistake_is_caught():
    # a transposed W2 in the backward pass must fail the same comparison
    rng = np.random.default_rng(4)
    F, n = 19, 30
    P = R.init(F, 5)
    o = R.layout(F)
    P[o["b2"]:o["w3"]] = rng.normal(0, 0.3, R.H)
    Z = rng.normal(0, 1, (n, F))
    T = rng.normal(0, 1, n)
    W1, b1, W2, b2, w3, b3, _, _ = R.unpack(P, F)
    z1 = Z @ W1.T + b1
    h1 = np.maximum(z1, 0)
    z2 = h1 @ W2.T + b2
    dy = 2 * ((np.maximum(z2, 0) @ w3 + b3) - T) / n
    dz2 = np.outer(dy, w3) * (z2 > 0)
    wrong = ((dz2 @ W2.T) * (z1 > 0)).T @ Z  # W2 transposed
    _, G = R.loss_and_grad(P, F, Z, T)
    assert np.abs(wrong.reshape(-1) - G[o["W1"]:o["b1"]]).max() > 1e-3


def test_initialiser_hash_and_range():
    # fmix64 (murmur3's 64-bit finaliser) of 0 is 0; values land in [0, 1)
    assert R.fmix64(0) == 0
    assert R.fmix64(1) == 0xB456BCFC34C2CB2C
    u = [R.urand(12345, i) for i in range(1000)]
    assert min(u) >= 0.0 and max(u) < 1.0 and 0.4 < np.mean(u) < 0.6
