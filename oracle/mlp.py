"""FP64 CPU oracle of the learned node-culling MLP (test infrastructure only:
imported by tests/ and nothing shipped).

The network has no counterpart in the reference.  It is 16-64-64-64-1 with ReLU
hidden layers and a sigmoid output; the flat weight layout is W1[64x16], b1[64],
W2[64x64], b2[64], W3[64x64], b3[64], W4[1x64], b4[1], every W row-major
[out][in].  The loss of a batch is the weighted binary cross entropy summed over
its rows, sum_r w_r * -(y_r log p_r + (1 - y_r) log(1 - p_r)), with p clamped to
[1e-7, 1 - 1e-7] inside the logarithms only.

`quantize=True` rounds what the device feeds to its BF16 matrix cores (inputs,
weights of the four layers, hidden activations) to bfloat16 with
round-to-nearest-even, so that the device's forward pass can be compared at a
tolerance set by FP32 accumulation order alone.  The gradient is the exact FP64
gradient of that (quantized) network with respect to the weights as used;
tests pin it by central finite differences of `loss`.
"""
import numpy as np

NIN, NH = 16, 64
SHAPES = [("W1", (NH, NIN)), ("b1", (NH,)), ("W2", (NH, NH)), ("b2", (NH,)),
          ("W3", (NH, NH)), ("b3", (NH,)), ("W4", (1, NH)), ("b4", (1,))]
PARAM_COUNT = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(w):
    w = np.asarray(w, np.float64).reshape(-1)
    assert w.size == PARAM_COUNT
    out, o = {}, 0
    for name, shape in SHAPES:
        n = int(np.prod(shape))
        out[name] = w[o:o + n].reshape(shape)
        o += n
    return out


def pack(parts):
    return np.concatenate([np.asarray(parts[name], np.float64).reshape(-1) for name, _ in SHAPES])


def bf16(x):
    """round to bfloat16 (nearest, ties to even) and return as float64"""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _q(x, quantize):
    return bf16(x) if quantize else np.asarray(x, np.float64)


def forward(w, X, quantize=False, keep=False):
    P = unpack(w)
    x = _q(X, quantize)
    W = [_q(P[f"W{i}"], quantize) for i in range(1, 5)]
    z1 = x @ W[0].T + P["b1"]
    h1 = _q(np.maximum(z1, 0.0), quantize)
    z2 = h1 @ W[1].T + P["b2"]
    h2 = _q(np.maximum(z2, 0.0), quantize)
    z3 = h2 @ W[2].T + P["b3"]
    h3 = _q(np.maximum(z3, 0.0), quantize)
    z4 = (h3 @ W[3].T + P["b4"])[:, 0]
    p = 1.0 / (1.0 + np.exp(-z4))
    if keep:
        return p, dict(x=x, W=W, z=(z1, z2, z3), h=(h1, h2, h3))
    return p


def loss(w, X, y, rw=None, quantize=False):
    p = forward(w, X, quantize)
    rw = np.ones(len(p)) if rw is None else np.asarray(rw, np.float64)
    y = np.asarray(y, np.float64)
    pc = np.clip(p, 1e-7, 1 - 1e-7)
    return float(np.sum(-rw * (y * np.log(pc) + (1 - y) * np.log(1 - pc))))


def gradient(w, X, y, rw=None, quantize=False):
    """(loss, d loss / d weights) by back-propagation in FP64"""
    p, c = forward(w, X, quantize, keep=True)
    rw = np.ones(len(p)) if rw is None else np.asarray(rw, np.float64)
    y = np.asarray(y, np.float64)
    z1, z2, z3 = c["z"]
    h1, h2, h3 = c["h"]
    W = c["W"]
    d4 = rw * (p - y)                      # sigmoid + cross entropy (unclamped p)
    g = {"W4": d4[None, :] @ h3, "b4": np.array([d4.sum()])}
    d3 = (d4[:, None] @ W[3]) * (z3 > 0)
    g["W3"], g["b3"] = d3.T @ h2, d3.sum(0)
    d2 = (d3 @ W[2]) * (z2 > 0)
    g["W2"], g["b2"] = d2.T @ h1, d2.sum(0)
    d1 = (d2 @ W[1]) * (z1 > 0)
    g["W1"], g["b1"] = d1.T @ c["x"], d1.sum(0)
    return loss(w, X, y, rw, quantize), pack(g)


def node_geometry(children):
    """depth and integer cell coordinates of every node from the child links
    (parents precede children in DFS block order)"""
    INVALID = 0xFFFFFFFF
    n = len(children)
    depth = np.zeros(n, np.int64)
    k = np.zeros((n, 3), np.int64)
    for nd in range(n):
        c = int(children[nd])
        if c == INVALID or c & 0x80000000:
            continue
        for s in range(8):
            depth[c + s] = depth[nd] + 1
            k[c + s] = 2 * k[nd] + [(s >> 0) & 1, (s >> 1) & 1, (s >> 2) & 1]
    return depth, k


def node_features(tree, samples, node_states, ksum_s, root_min, root_size, hits_s=None):
    """FP64 restatement of the 16 node features from a tree (DFS block order,
    `children` without a leaf flag: INVALID for leaves), its node-ordered samples,
    the node state flags (bit 0: sampleness flag) and per-leaf sampleness kernel
    sums.  Columns 9 and 10 need the subsampled emptiness hit counts and are
    returned as NaN; column 8 is NaN unless per-leaf sampleness hit counts are given.

    0 depth / 19; 1 log(1 + subtree samples); 2 mean scale / node size; 3 variance
    of scale / node size^2; 4-6 mean normal; 7 1 - |mean normal| (0 if empty);
    8 log(1 + subtree sampleness hits); 11 log(1 + subtree sampleness kernel
    sum) / 30; 12 fraction of subtree leaves with the sampleness flag; 13
    log2(subtree leaves) / 20; 14 |mean position - node centre| / node size (0 if
    empty); 15 the node's sampleness flag."""
    ch = np.asarray(tree["children"])
    n = len(ch)
    depth, k = node_geometry(ch)
    leaves = np.asarray(tree["leaf_nodes"])
    cnt_own = np.asarray(tree["sample_counts"], np.int64)
    starts = np.asarray(tree["sample_starts"], np.int64)
    owner = np.repeat(np.arange(n), cnt_own)
    sc = np.asarray(samples["scale"], np.float64)
    agg = np.zeros((n, 14))
    cols = [np.ones_like(sc), sc, sc * sc] + [np.asarray(samples[c], np.float64)
                                              for c in ("nx", "ny", "nz", "px", "py", "pz")]
    for j, v in enumerate(cols):
        agg[:, j] = np.bincount(owner, v, minlength=n)
    assert np.all(starts[cnt_own > 0] == np.concatenate([[0], np.cumsum(cnt_own)[:-1]])[cnt_own > 0])
    sflag = (np.asarray(node_states) & 1) != 0
    if hits_s is not None:
        agg[leaves, 9] = hits_s
    agg[leaves, 11] = 1.0
    agg[leaves, 12] = sflag[leaves]
    agg[leaves, 13] = ksum_s
    parents = np.full(n, -1)
    for nd in range(n - 1, -1, -1):          # children have larger ids than their parent
        c = int(ch[nd])
        if c == 0xFFFFFFFF:
            continue
        agg[nd] += agg[c:c + 8].sum(0)
        parents[c:c + 8] = nd
    h = float(root_size) * 2.0 ** -depth
    cnt = agg[:, 0]
    inv = np.where(cnt > 0, 1.0 / np.maximum(cnt, 1), 0.0)
    ms = agg[:, 1] * inv / h
    ms2 = agg[:, 2] * inv / h ** 2
    mn = agg[:, 3:6] * inv[:, None]
    ctr = np.asarray(root_min, np.float64) + (2 * k + 1) * (float(root_size) * 2.0 ** -(depth + 1))[:, None]
    dist = np.where(cnt > 0, np.linalg.norm(agg[:, 6:9] * inv[:, None] - ctr, axis=1) / h, 0.0)
    f = np.full((n, 16), np.nan)
    f[:, 0] = depth / 19.0
    f[:, 1] = np.log1p(cnt)
    f[:, 2] = ms
    f[:, 3] = np.maximum(ms2 - ms * ms, 0.0)
    f[:, 4:7] = mn
    f[:, 7] = np.where(cnt > 0, 1.0 - np.linalg.norm(mn, axis=1), 0.0)
    if hits_s is not None:
        f[:, 8] = np.log1p(agg[:, 9])
    f[:, 11] = np.log1p(agg[:, 13]) / 30.0
    f[:, 12] = np.where(agg[:, 11] > 0, agg[:, 12] / np.maximum(agg[:, 11], 1), 0.0)
    f[:, 13] = np.log2(np.maximum(agg[:, 11], 1.0)) / 20.0
    f[:, 14] = dist
    f[:, 15] = sflag
    return f, parents
