"""The FP64 MLP oracle pinned by central finite differences, a hand-built network
with a closed-form output and bfloat16 rounding known answers."""
import numpy as np
import pytest

from oracle import mlp


def _random_net(seed):
    """He-scaled weights and small random biases (outputs away from saturation)"""
    rng = np.random.default_rng(seed)
    P = {}
    for name, shape in mlp.SHAPES:
        fan_in = shape[1] if len(shape) == 2 else 64
        P[name] = rng.normal(0.0, np.sqrt(2.0 / fan_in) if len(shape) == 2 else 0.1, shape)
    return mlp.pack(P)


def test_bf16_rounding_known_answers():
    ulp = 2.0 ** -7   # spacing of bfloat16 in [1, 2)
    x = np.array([1.0, 1.0 + ulp / 2, 1.0 + 1.5 * ulp, 1.0 + ulp / 2 + 2.0 ** -20, -3.0, 0.0, 65280.0])
    got = mlp.bf16(x)
    # ties go to even: 1 + ulp/2 -> 1, 1 + 1.5 ulp -> 1 + 2 ulp; above the tie rounds up
    assert list(got) == [1.0, 1.0, 1.0 + 2 * ulp, 1.0 + ulp, -3.0, 0.0, 65280.0]


def test_closed_form_network():
    # W1 picks +x0 and -x0; layers 2 and 3 pass units 0 and 1 on; W4 = [1, -1]
    P = {name: np.zeros(shape) for name, shape in mlp.SHAPES}
    P["W1"][0, 0], P["W1"][1, 0] = 1.0, -1.0
    P["W2"][0, 0] = P["W2"][1, 1] = 1.0
    P["W3"][0, 0] = P["W3"][1, 1] = 1.0
    P["W4"][0, 0], P["W4"][0, 1] = 1.0, -1.0
    P["b4"][0] = 0.25
    w = mlp.pack(P)
    X = np.zeros((7, 16))
    X[:, 0] = np.linspace(-3, 3, 7)
    X[:, 5] = 9.0                               # ignored input
    assert np.allclose(mlp.forward(w, X), 1.0 / (1.0 + np.exp(-(X[:, 0] + 0.25))), rtol=0, atol=1e-15)
    y = (X[:, 0] > 0).astype(float)
    p = 1.0 / (1.0 + np.exp(-(X[:, 0] + 0.25)))
    assert mlp.loss(w, X, y) == pytest.approx(-np.sum(y * np.log(p) + (1 - y) * np.log(1 - p)), rel=1e-14)


@pytest.mark.parametrize("seed", [0, 1])
def test_gradient_matches_central_differences(seed):
    rng = np.random.default_rng(100 + seed)
    w = _random_net(seed)
    X = rng.normal(size=(40, 16))
    y = (rng.random(40) < 0.4).astype(float)
    rw = rng.uniform(0.5, 2.0, 40)
    lv, g = mlp.gradient(w, X, y, rw)
    assert lv == pytest.approx(mlp.loss(w, X, y, rw), rel=1e-14)
    # every bias and output weight, and a spread of hidden weights
    idx = np.unique(np.concatenate([np.arange(mlp.PARAM_COUNT - 65, mlp.PARAM_COUNT),
                                    rng.choice(mlp.PARAM_COUNT, 300, replace=False)]))
    h = 1e-5
    fd = np.empty(len(idx))
    for j, i in enumerate(idx):
        wp, wm = w.copy(), w.copy()
        wp[i] += h
        wm[i] -= h
        fd[j] = (mlp.loss(wp, X, y, rw) - mlp.loss(wm, X, y, rw)) / (2 * h)
    assert np.allclose(g[idx], fd, rtol=1e-5, atol=1e-7)
    # a wrong transpose or sign would fail this too: the directional derivative
    d = rng.normal(size=mlp.PARAM_COUNT)
    d /= np.linalg.norm(d)
    e = 1e-5
    dd = (mlp.loss(w + e * d, X, y, rw) - mlp.loss(w - e * d, X, y, rw)) / (2 * e)
    assert float(g @ d) == pytest.approx(dd, rel=1e-5)


def test_quantized_forward_close_to_exact():
    rng = np.random.default_rng(3)
    w = _random_net(3)
    X = rng.normal(size=(64, 16))
    p, q = mlp.forward(w, X), mlp.forward(w, X, quantize=True)
    assert np.max(np.abs(p - q)) < 0.02 and not np.array_equal(p, q)
