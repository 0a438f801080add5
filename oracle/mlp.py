"""FP64 CPU oracle of the routing MLP of the certified priors and of its per-leaf
features (test infrastructure only: imported by tests/ and nothing shipped).

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


def leaf_features(tree, field, eps=float(np.float32(1e-6))):
    """FP64 restatement of the 16 per-leaf routing features from a tree (download():
    children with INVALID for leaves, parents, node_depth, leaf_nodes, node_to_leaf) and
    its field (kernel_sums, cone_lengths, node_states, cone_count).

    0 depth / 19; 1, 2 log(1 + E), log(1 + S) kernel sums / 16; 3 E / (E + S) (0.5 if both
    are 0); 4, 5 log(1 + cone length) E, S / 16; for the ancestors 1..4 levels up (the root
    stays the root): 6..9 their subtree prior Esub / (Esub + Ssub) over cone lengths (0.5 if
    0) and 10..13 the occupancy margin (pS - pE) / (pS + pE) at that prior, pE = p E / M,
    pS = (1 - p) S / M (0 if both are 0); 14 sign(r) log(1 + |r|) / 20 with r = (v - EPS) / EPS,
    v = pS - pE at the prior 3 levels up; 15 log10(max((E + S) / M, 1e-30)) / 10."""
    ch = np.asarray(tree["children"])
    n = len(ch)
    leaves = np.asarray(tree["leaf_nodes"], np.int64)
    L = len(leaves)
    clen = np.asarray(field["cone_lengths"], np.float64)
    ks = np.asarray(field["kernel_sums"], np.float64)
    E, S = ks[:L], ks[L:]
    M = float(field["cone_count"]) if field["cone_count"] > 0 else 1.0
    sub = np.zeros((n, 2))
    sub[leaves, 0], sub[leaves, 1] = clen[:L], clen[L:]
    for nd in range(n - 1, -1, -1):          # children have larger ids than their parent
        c = int(ch[nd])
        if c != 0xFFFFFFFF:
            sub[nd] = sub[c:c + 8].sum(0)
    parents = np.asarray(tree["parents"], np.int64)
    depth = np.asarray(tree["node_depth"], np.float64)
    f = np.zeros((L, 16))
    f[:, 0] = depth[leaves] / 19.0
    f[:, 1], f[:, 2] = np.log1p(E) / 16, np.log1p(S) / 16
    with np.errstate(divide="ignore", invalid="ignore"):
        f[:, 3] = np.where(E + S > 0, E / (E + S), 0.5)
        f[:, 4], f[:, 5] = np.log1p(clen[:L]) / 16, np.log1p(clen[L:]) / 16
        a = leaves.copy()
        v3 = None
        for k in range(1, 5):
            pa = parents[a]
            a = np.where(pa != 0xFFFFFFFF, pa, a)
            t = sub[a].sum(1)
            p = np.where(t > 0, sub[a, 0] / t, 0.5)
            pE, pS = p * (E / M), (1 - p) * (S / M)
            f[:, 5 + k] = p
            f[:, 9 + k] = np.where(pS + pE > 0, (pS - pE) / (pS + pE), 0.0)
            if k == 3:
                v3 = pS - pE
        r = (v3 - eps) / eps
        f[:, 14] = np.sign(r) * np.log1p(np.abs(r)) / 20
        f[:, 15] = np.log10(np.maximum((E + S) / M, 1e-30)) / 10
    return f
