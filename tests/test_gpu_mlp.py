"""Learned node culling on the GPU: the BF16 MFMA forward and backward kernels
against the FP64 MLP oracle, device training, node features against their FP64
restatement, training labels, and the culling mechanism in the occupancy pass."""
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


def test_node_features_match_restatement(gpu_ctx, small_field):
    sc, tree, occ = small_field
    t = tree.download()
    smp = tree.download_samples()
    f = occ.download()
    feat = tsr_amd.node_features(gpu_ctx, tree, sc, occ)
    assert feat.shape == (tree.node_count, 16) and np.all(np.isfinite(feat))
    Lc = tree.leaf_count
    ref, parents = mlp.node_features(t, smp, f["node_states"], f["kernel_sums"][Lc:], tree.root_min,
                                     tree.root_size)
    for col in (0, 1, 2, 4, 5, 6, 12, 13, 15):
        assert np.allclose(feat[:, col], ref[:, col], rtol=1e-4, atol=1e-5), col
    # variance and normal dispersion lose digits to FP32 cancellation
    assert np.allclose(feat[:, 3], ref[:, 3], rtol=1e-2, atol=1e-3 * np.maximum(ref[:, 2] ** 2, 1e-6).max())
    assert np.allclose(feat[:, 7], ref[:, 7], atol=1e-4)
    assert np.allclose(feat[:, 14], ref[:, 14], rtol=1e-3, atol=1e-3)
    assert np.allclose(feat[:, 11], ref[:, 11], rtol=1e-4, atol=1e-6)
    # hit-count features: leaf hits > 0 exactly where the sampleness flag is set,
    # and their subtree totals add up to the field's sampleness hit count
    leaves = t["leaf_nodes"]
    hits = np.expm1(feat[leaves, 8].astype(np.float64))
    assert np.array_equal(hits > 0.5, (f["node_states"][leaves] & 1) != 0)
    assert np.rint(hits).sum() == pytest.approx(occ.hit_counts()[1], rel=1e-6)
    assert np.expm1(feat[0, 8].astype(np.float64)) == pytest.approx(occ.hit_counts()[1], rel=1e-5)
    assert np.all(feat[:, 9] >= 0) and feat[0, 9] > 0
    # parent estimate: in [-1, 1], computed from the parent's subtree counts
    assert np.all(np.abs(feat[:, 10]) <= 1.0)


def test_node_labels_are_subtree_ors(gpu_ctx, small_field):
    _, tree, occ = small_field
    t = tree.download()
    f = occ.download()
    lab = tsr_amd.node_labels(gpu_ctx, tree, occ)
    ch = t["children"]
    leaves = t["leaf_nodes"]
    occv = f["occupancy"]
    cells = t["cells"].reshape(-1, 8)
    pos = occv[cells] > 0
    mixed = np.any(pos, 1) & ~np.all(pos, 1)
    leaf_lab = occv > 0
    leaf_lab[np.unique(cells[mixed])] = True
    assert np.array_equal(lab[leaves].astype(bool), leaf_lab)
    inner = np.nonzero(ch != 0xFFFFFFFF)[0]
    for nd in inner:
        assert lab[nd] == lab[ch[nd]:ch[nd] + 8].max()
    assert 0 < lab.mean() < 1


def test_culling_with_constant_networks(gpu_ctx, small_field):
    sc, tree, plain = small_field
    ref = plain.download()
    mref = tsr_amd.DualMarchingCells(gpu_ctx, tree, plain).download()
    keep = np.zeros(mlp.PARAM_COUNT, np.float32)
    keep[-1] = 20.0                       # p ~ 1 everywhere: nothing is culled
    with_keep = tsr_amd.Culler(gpu_ctx, keep, threshold=0.5)
    o = tsr_amd.Occupancy(gpu_ctx, tree, sc, flags=tsr_amd._lib.TSR_CULL_LEARNED, culler=with_keep)
    assert o.cull_counts() == (0, 0)
    got = o.download()
    for k in ("kernel_sums", "cone_lengths", "occupancy"):
        assert np.array_equal(got[k], ref[k], equal_nan=True), k
    m = tsr_amd.DualMarchingCells(gpu_ctx, tree, o).download()
    assert np.array_equal(m["triangles"], mref["triangles"])
    o.close()
    drop = keep.copy()
    drop[-1] = -20.0                      # p ~ 0: every node at depth >= 2 is culled
    with_drop = tsr_amd.Culler(gpu_ctx, drop, threshold=0.5)
    o = tsr_amd.Occupancy(gpu_ctx, tree, sc, flags=tsr_amd._lib.TSR_CULL_LEARNED, culler=with_drop)
    nodes, leaves = o.cull_counts()
    t = tree.download()
    assert leaves == tree.leaf_count and nodes == int(np.sum(t["node_depth"] >= 2))
    got = o.download()
    assert np.all(got["occupancy"] == -np.float64(np.float32(1e-6)))
    assert np.all(got["kernel_sums"][:tree.leaf_count] == 0)       # no emptiness evaluation
    assert np.array_equal(got["kernel_sums"][tree.leaf_count:], ref["kernel_sums"][tree.leaf_count:])
    assert got["cone_count"] == ref["cone_count"]
    o.close()
    with_keep.close()
    with_drop.close()
