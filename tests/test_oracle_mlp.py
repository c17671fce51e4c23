"""Pins the FP64 oracle of the selector MLP (orc_mlp_forward / orc_mlp_grad)
before it checks the HIP forward / backward kernels:
  forward   the matrix form x @ W + b with ReLU, written with numpy on the
            parameter layout (input-major), and closed forms (zero weights
            give the output biases; a single active path gives relu(x0) * c);
  gradient  central finite differences of the oracle's own loss (the
            definition of the derivative), and the closed form of the
            output-bias gradient when every weight is zero:
            dL/db3[a] = 2 / U * sum over records with arm a of (b3[a] - t).
Planted mistakes (a transposed W2, a dropped ReLU mask, the wrong loss
target) are shown to fail these pins.
"""
import numpy as np
import pytest

import oracle

NI, NH, NO = 16, 64, 4
W1, B1 = 0, NI * NH
W2 = B1 + NH
B2 = W2 + NH * NH
W3 = B2 + NH
B3 = W3 + NH * NO
NP = B3 + NO


def unpack(w):
    return (w[W1:B1].reshape(NI, NH), w[B1:W2], w[W2:B2].reshape(NH, NH), w[B2:W3],
            w[W3:B3].reshape(NH, NO), w[B3:NP])


def numpy_forward(w, x):
    w1, b1, w2, b2, w3, b3 = unpack(w)
    h1 = np.maximum(x @ w1 + b1, 0)
    h2 = np.maximum(h1 @ w2 + b2, 0)
    return h2 @ w3 + b3


def numpy_loss(w, rec):
    x = rec[:, :NI].astype(np.float64)
    arm = rec[:, NI].astype(int)
    cost = rec[:, NI + 1].astype(np.float64)
    keep = cost > 0
    y = numpy_forward(w, x[keep])
    return float(np.mean((y[np.arange(keep.sum()), arm[keep]] - np.log2(cost[keep])) ** 2))


def rand_params(seed, scale=0.3):
    return np.random.default_rng(seed).normal(0, scale, NP)


def rand_records(seed, count=40):
    rng = np.random.default_rng(seed)
    rec = np.zeros((count, 20), np.float32)
    rec[:, :NI] = rng.uniform(0, 1, (count, NI))
    rec[:, NI] = rng.integers(0, NO, count)
    rec[:, NI + 1] = np.exp2(rng.uniform(1, 10, count))
    rec[::7, NI + 1] = -1  # iterations without a measured cost are skipped
    return rec


def test_p_count():
    assert NP == oracle.MLP_PARAMS == 5508


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_forward_matches_matrix_form(seed):
    w = rand_params(seed)
    x = np.random.default_rng(seed + 10).uniform(-1, 1, (9, NI))
    assert np.allclose(oracle.mlp_forward(w, x), numpy_forward(w, x), rtol=1e-12, atol=1e-12)


def test_forward_closed_forms():
    w = np.zeros(NP)
    w[B3:NP] = [1.5, -2.0, 0.25, 7.0]
    x = np.random.default_rng(0).uniform(0, 1, (3, NI))
    assert np.allclose(oracle.mlp_forward(w, x), np.tile(w[B3:NP], (3, 1)))
    # One path: x0 -> hidden1 unit 5 -> hidden2 unit 9 -> output 2, gain 2*3*0.5.
    w = np.zeros(NP)
    w[W1 + 0 * NH + 5] = 2.0
    w[W2 + 5 * NH + 9] = 3.0
    w[W3 + 9 * NO + 2] = 0.5
    for x0 in (-1.0, 0.0, 0.7):
        x = np.zeros(NI)
        x[0] = x0
        y = oracle.mlp_forward(w, x)[0]
        assert y[2] == pytest.approx(3.0 * max(x0, 0.0)) and y[[0, 1, 3]].tolist() == [0, 0, 0]


@pytest.mark.parametrize("seed", [4, 5])
def test_gradient_matches_finite_differences(seed):
    w = rand_params(seed)
    rec = rand_records(seed)
    g, loss, used = oracle.mlp_grad(w, rec)
    assert used == int((rec[:, NI + 1] > 0).sum())
    assert loss == pytest.approx(numpy_loss(w, rec), rel=1e-12)
    rng = np.random.default_rng(seed)
    idx = np.concatenate([rng.choice(NP, 150, replace=False), np.arange(B3, NP)])
    h = 1e-6
    for p in idx:
        wp, wm = w.copy(), w.copy()
        wp[p] += h
        wm[p] -= h
        fd = (oracle.mlp_grad(wp, rec)[1] - oracle.mlp_grad(wm, rec)[1]) / (2 * h)
        assert g[p] == pytest.approx(fd, rel=1e-5, abs=1e-7), p


def test_gradient_closed_form_output_bias():
    w = np.zeros(NP)
    w[B3:NP] = [3.0, 1.0, -2.0, 5.0]
    rec = rand_records(7)
    g, loss, used = oracle.mlp_grad(w, rec)
    keep = rec[:, NI + 1] > 0
    t = np.log2(rec[keep, NI + 1].astype(np.float64))
    arm = rec[keep, NI].astype(int)
    for a in range(NO):
        assert g[B3 + a] == pytest.approx(2.0 / used * np.sum(w[B3 + a] - t[arm == a]), rel=1e-12)
    assert np.all(g[:B3] == 0)  # zero weights: no signal reaches the hidden layers


def test_planted_mistakes_fail_the_pins():
    w = rand_params(9)
    x = np.random.default_rng(9).uniform(-1, 1, (5, NI))
    ref = oracle.mlp_forward(w, x)
    # A transposed W2 gives another function.
    wt = w.copy()
    wt[W2:B2] = w[W2:B2].reshape(NH, NH).T.ravel()
    assert not np.allclose(numpy_forward(wt, x), ref)
    # Dropping the ReLU on the first layer changes the output.
    w1, b1, w2, b2, w3, b3 = unpack(w)
    no_relu = np.maximum((x @ w1 + b1) @ w2 + b2, 0) @ w3 + b3
    assert not np.allclose(no_relu, ref)
    # A loss against the raw cost instead of log2(cost) has another value.
    rec = rand_records(9)
    keep = rec[:, NI + 1] > 0
    y = numpy_forward(w, rec[keep, :NI].astype(np.float64))
    raw = np.mean((y[np.arange(keep.sum()), rec[keep, NI].astype(int)] - rec[keep, NI + 1]) ** 2)
    assert raw != pytest.approx(oracle.mlp_grad(w, rec)[1], rel=1e-3)
