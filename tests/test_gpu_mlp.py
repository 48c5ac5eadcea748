"""The FP32 MLP bucket-size predictor on the device (k_mlp_train,
k_mlp_infer) against the numpy restatement (mlp_ref.py, float64, its backward
pinned by finite differences): inference with the device's parameters, and
training from the same initialisation and training set."""
import numpy as np
import pytest

import cfgpu as C
import mlp_ref as R
from test_gpu_parity import say

pytestmark = pytest.mark.gpu


def train_on_device(epochs):
    g = C.synth_genome(2_000_000, 0.1, 300, 3000, 0.02, seed=91)
    reads = C.synth_reads(g, 200_000, 150, 0.01, seed=92)
    inp = C.Input.from_records(reads, np.arange(200_001, dtype=np.uint64) * np.uint64(150), 31)
    with C.Handle(31, 20, 2) as h:
        h.partition(inp)
        h.set_tuning(predictor=1, mlp_epochs=epochs)  # train the MLP whatever the held-out errors say
        h.construct("learned")
        rs = h.run_stats()
        assert rs["predictor"] == 1
        m = h.sizing_model()
    say(f"[case] mlp: {m['n_train']} training buckets, {m['epochs']} epochs, loss {m['loss']:.5g}")
    return m


@pytest.fixture(scope="module")
def model():
    return train_on_device(120)


def test_first_step_follows_the_reference():
    # One full-batch Adam step from the same initialisation.  Adam's first
    # step moves every parameter by lr * g / (|g| + eps) plus the weight decay,
    # so it fixes the sign of every gradient component and the update rule
    # (bias correction, eps, decoupled decay).  Every parameter whose float64
    # gradient is at least 1e-5 (over 95 % of them) must match to 1e-6; a
    # wrong gradient term, sign or index moves hundreds of them by about
    # lr = 3e-3.  Below that an FP32 gradient can round to the other sign (a
    # ReLU gate within rounding of zero): such a parameter may take the
    # opposite step, never more.
    m = train_on_device(1)
    F = m["n_features"]
    o = R.layout(F)
    X = m["features"][m["train_buckets"]].astype(np.float64)
    T = m["train_targets"].astype(np.float64)
    P, loss = R.train(X, T, m["seed"], 1, m["lr"])
    P0 = R.init(F, m["seed"])
    mu, istd = R.standardise(X)
    P0[o["mean"]:o["istd"]], P0[o["istd"]:o["end"]], P0[o["b3"]] = mu, istd, T.mean()
    _, G = R.loss_and_grad(P0, F, (X - mu) * istd, T)
    dev = m["params"].astype(np.float64)
    net = slice(0, o["mean"])
    strong = np.abs(G[net]) >= 1e-5
    diff = np.abs(dev - P)[net]
    say(f"[case] mlp step 1: {int(strong.sum())} of {strong.size} gradients resolvable, their max diff "
        f"{diff[strong].max():.3g}; others max diff {diff[~strong].max() if (~strong).any() else 0:.3g}")
    assert strong.mean() >= 0.95
    assert diff[strong].max() <= 1e-6
    assert diff.max() <= 2.5 * m["lr"]
    np.testing.assert_allclose(dev[o["mean"]:], P[o["mean"]:], rtol=1e-5, atol=1e-6)
    assert abs(m["loss"] - loss) <= 1e-5 * abs(loss)


def test_inference_matches_reference(model):
    m = model
    want = R.forward(m["params"].astype(np.float64), m["n_features"], m["features"].astype(np.float64))
    # FP32 dot products of 19 and 64 terms: a few ulps of the activations
    np.testing.assert_allclose(m["predictions"], want, rtol=1e-4, atol=1e-5)


def test_training_matches_reference(model):
    m = model
    X = m["features"][m["train_buckets"]].astype(np.float64)
    T = m["train_targets"].astype(np.float64)
    P, loss = R.train(X, T, m["seed"], m["epochs"], m["lr"])
    dev = m["params"].astype(np.float64)
    F = m["n_features"]
    o = R.layout(F)
    dp = np.abs(dev - P)
    say(f"[case] mlp: |param diff| max {dp.max():.3g} mean {dp.mean():.3g}; loss {m['loss']:.6g} vs {loss:.6g}")
    # the standardisation is computed, not trained: FP32 rounding only
    np.testing.assert_allclose(dev[o["mean"]:], P[o["mean"]:], rtol=1e-5, atol=1e-6)
    pd = R.forward(dev, F, m["features"].astype(np.float64))
    pr = R.forward(P, F, m["features"].astype(np.float64))
    dpred = np.abs(pd - pr)
    say(f"[case] mlp: |prediction diff| max {dpred.max():.3g} mean {dpred.mean():.3g}")
    # After 120 epochs the loss is near zero and its gradients near FP32
    # rounding; Adam's normalised steps then let the FP32 and float64
    # trajectories part by a few lr-sized steps in flat directions.  The
    # trained function must still agree: loss within 1 %, every bucket's
    # predicted log(distinct / instances) within 0.02 (2 % of a table size,
    # under the 15 % sizing margin).
    assert abs(m["loss"] - loss) <= 1e-2 * abs(loss) + 1e-7
    assert dpred.max() <= 2e-2
