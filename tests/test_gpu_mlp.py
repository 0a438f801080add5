"""The routing MLP on the GPU: the BF16 MFMA forward and backward kernels against the FP64
MLP oracle, device training (deterministic: the same run twice gives the same weights),
and the per-leaf features against their FP64 restatement."""
import numpy as np
import pytest

import tsr_amd
from oracle import mlp
from tests.helpers import scene

pytestmark = pytest.mark.gpu


def _net(seed):
    rng = np.random.default_rng(seed)
    P = {}
    for name, shape in mlp.SHAPES:
        P[name] = rng.normal(0.0, np.sqrt(2.0 / shape[1]) if len(shape) == 2 else 0.1, shape)
    return mlp.pack(P).astype(np.float32)


@pytest.mark.parametrize("n", [1, 15, 16, 17, 1000, 40000])
def test_forward_matches_oracle(gpu_ctx, n):
    rng = np.random.default_rng(n)
    w = _net(n)
    X = rng.normal(size=(n, 16)).astype(np.float32)
    p = tsr_amd.mlp_forward(gpu_ctx, w, X)
    ref = mlp.forward(w, X, quantize=True)
    # BF16 operands as the oracle rounds them; FP32 accumulation order and the
    # BF16 rounding of hidden activations near a rounding tie are all that differ
    assert np.max(np.abs(p - ref)) < 2e-3, np.max(np.abs(p - ref))
    assert np.mean(np.abs(p - ref)) < 2e-4


@pytest.mark.parametrize("n", [16, 333, 20000])
def test_gradient_matches_oracle(gpu_ctx, n):
    rng = np.random.default_rng(7 + n)
    w = _net(n + 1)
    X = rng.normal(size=(n, 16)).astype(np.float32)
    y = (rng.random(n) < 0.3).astype(np.float32)
    rw = rng.uniform(0.5, 2.0, n).astype(np.float32)
    loss, g = tsr_amd.mlp_gradient(gpu_ctx, w, X, y, rw)
    ref_loss, ref_g = mlp.gradient(w, X, y, rw, quantize=True)
    assert loss == pytest.approx(ref_loss, rel=2e-3)
    # the backward pass rounds each layer's gradient to BF16 before the MFMA:
    # a relative error of a few 2^-9 per element, summed over rows
    err = np.linalg.norm(g - ref_g) / np.linalg.norm(ref_g)
    assert err < 1e-2, err
    for name, sl in _slices().items():
        e = np.linalg.norm(g[sl] - ref_g[sl]) / max(np.linalg.norm(ref_g[sl]), 1e-30)
        assert e < 2e-2, (name, e)


def _slices():
    out, o = {}, 0
    for name, shape in mlp.SHAPES:
        k = int(np.prod(shape))
        out[name] = slice(o, o + k)
        o += k
    return out


def test_gradient_of_empty_batch_is_zero(gpu_ctx):
    w = _net(0)
    loss, g = tsr_amd.mlp_gradient(gpu_ctx, w, np.zeros((0, 16), np.float32), np.zeros(0, np.float32))
    assert loss == 0.0 and not g.any()


def test_training_learns_a_separable_rule(gpu_ctx):
    rng = np.random.default_rng(11)
    n = 30000
    X = rng.normal(size=(n, 16)).astype(np.float32)
    y = ((X[:, 0] + 0.5 * X[:, 3] - 0.25 * X[:, 7]) > 0.2).astype(np.float32)
    w0 = tsr_amd.init_mlp_weights(5)
    w, losses = tsr_amd.mlp_train(gpu_ctx, w0, X, y, epochs=8, batch=1024, learning_rate=2e-3, seed=1)
    assert losses[-1] < 0.5 * losses[0] and losses[-1] < 0.1, losses
    Xt = rng.normal(size=(5000, 16)).astype(np.float32)
    yt = ((Xt[:, 0] + 0.5 * Xt[:, 3] - 0.25 * Xt[:, 7]) > 0.2)
    acc = np.mean((tsr_amd.mlp_forward(gpu_ctx, w, Xt) > 0.5) == yt)
    assert acc > 0.97, acc
    # the device's training step is plain Adam on the summed gradient: one step
    # from w0 with lr on one batch moves every parameter by about lr
    w1, _ = tsr_amd.mlp_train(gpu_ctx, w0, X[:256], y[:256], epochs=1, batch=256, learning_rate=1e-3, seed=0)
    _, g = mlp.gradient(w0, X[:256], y[:256], quantize=True)
    moved = np.abs(w1 - w0)
    big = np.abs(g) > 1e-3 * np.abs(g).max()
    assert np.allclose(moved[big], 1e-3, rtol=0.05)
    assert np.all(np.sign(w1 - w0)[big] == -np.sign(g[big]))
    # fixed-order gradient reduction: training is reproducible bit for bit
    w2, _ = tsr_amd.mlp_train(gpu_ctx, w0, X[:4096], y[:4096], epochs=2, batch=1024, learning_rate=1e-3, seed=3)
    w3, _ = tsr_amd.mlp_train(gpu_ctx, w0, X[:4096], y[:4096], epochs=2, batch=1024, learning_rate=1e-3, seed=3)
    assert np.array_equal(w2, w3)


@pytest.fixture(scope="module")
def small_field(gpu_ctx):
    s, c = scene(3, 0.125, seed=5)
    sc = tsr_amd.Scene(gpu_ctx, s, c)
    tree = tsr_amd.Tree(gpu_ctx, sc)
    occ = tsr_amd.Occupancy(gpu_ctx, tree, sc)
    yield sc, tree, occ
    occ.close()
    tree.close()
    sc.close()


def test_leaf_features_match_restatement(gpu_ctx, small_field):
    _, tree, occ = small_field
    t = tree.download()
    f = occ.download()
    feat = tsr_amd.leaf_features(gpu_ctx, tree, occ)
    assert feat.shape == (tree.leaf_count, 16) and np.all(np.isfinite(feat))
    ref = mlp.leaf_features(t, f)
    # FP32 features of FP64 (or compensated FP32) quantities; column 14 is a log of a
    # difference of near-equal values where the occupancy sits at EPSILON
    for col in range(16):
        tol = 1e-3 if col == 14 else 1e-5
        assert np.allclose(feat[:, col], ref[:, col], rtol=tol, atol=tol), (col, np.max(np.abs(feat[:, col] - ref[:, col])))
