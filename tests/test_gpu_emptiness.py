"""GPU: the emptiness pass gathered per sampleness-flagged leaf (the default E pass) against
the per-link cone traversal (TSR_EMPTINESS_TRAVERSAL, the reference's loop order,
Occupancy.cpp:138-248) and against the FP64 oracle.

Both passes apply the same FP64 kernel to the same (link, leaf) pairs, so node flags, hit
counts, cone-count M and the mesh must be identical; the E kernel sums and cone lengths are
summed in different fixed orders and must agree within 1e-6 relative (FP32 Kahan sums).
Crafted scenes put samples close to their camera, so that cones cover many direction cells
or are listed as wide links, and give samples a second (or an invalid) camera link.
"""
import numpy as np
import pytest

from tests.helpers import PARAMS, rel_err, scene

pytestmark = pytest.mark.gpu


def _near_camera(samples, cameras, ratios, per_ratio=40, seed=3):
    """Move per_ratio samples per ratio r towards their camera until R / |e - a| = r
    (R = bandwidth 2 * scale): cones of wide angle, near the apex."""
    s = {k: v.copy() for k, v in samples.items()}
    rng = np.random.default_rng(seed)
    n = len(s["px"])
    cam = s["parent_cameras"]
    chosen = rng.choice(n, size=per_ratio * len(ratios), replace=False)
    for j, i in enumerate(chosen):
        r = ratios[j // per_ratio]
        a = np.array([cameras["x"][cam[i]], cameras["y"][cam[i]], cameras["z"][cam[i]]], np.float64)
        p = np.array([s["px"][i], s["py"][i], s["pz"][i]], np.float64)
        R = 2.0 * float(s["scale"][i])
        f = R / (r * np.linalg.norm(p - a))
        q = a + (p - a) * f
        s["px"][i], s["py"][i], s["pz"][i] = np.float32(q[0]), np.float32(q[1]), np.float32(q[2])
    return s


def _two_links(samples, cameras):
    """A second camera link per sample: the next camera, or invalid for every third sample."""
    s = {k: v.copy() for k, v in samples.items()}
    n, K = len(s["px"]), len(cameras["x"])
    cams = np.empty(2 * n, np.uint32)
    cams[0::2] = s["parent_cameras"]
    second = (s["parent_cameras"].astype(np.int64) + 1) % K
    second[np.arange(n) % 3 == 0] = 0xFFFFFFFF
    cams[1::2] = second.astype(np.uint32)
    s["parent_cameras"] = cams
    return s


def _scenes():
    """(name, samples, cameras, links per sample, compare with the oracle)"""
    out = []
    for cfg, res in ((0, 0.5), (1, 0.0625), (3, 0.125), (4, 0.03125)):
        smp, cams = scene(cfg, res)
        out.append((f"cfg{cfg}x{res}", smp, cams, 1, cfg != 4))
    # cone angles of 0.02 and 0.3 rad go to direction cells, 0.7 to the wide lists; at half
    # resolution (finer cells) 0.3 covers more than WIDE_CELLS cells and is listed as wide
    for res in (0.0625, 0.5):
        smp, cams = scene(1, res)
        out.append((f"cfg1x{res}_near_camera", _near_camera(smp, cams, (0.02, 0.3, 0.7)), cams, 1, res < 0.1))
    smp, cams = scene(3, 0.125)
    out.append(("cfg3_two_links", _two_links(smp, cams), cams, 2, True))
    return out


SCENES = _scenes()


@pytest.fixture(scope="module", params=range(len(SCENES)), ids=[s[0] for s in SCENES])
def fields(request, gpu_ctx):
    from tsr_amd import DualMarchingCells, Occupancy, Scene, Tree
    from tsr_amd._lib import TSR_EMPTINESS_TRAVERSAL, TSR_PRIORS_ALL_LEAVES
    name, samples, cameras, mc, with_oracle = SCENES[request.param]
    print(f"\n[emptiness] {name}: {len(samples['px'])} samples, {mc} link(s) per sample", flush=True)
    sc = Scene(gpu_ctx, samples, cameras, max_cams=mc)
    tree = Tree(gpu_ctx, sc)
    out = {"name": name, "samples": samples, "cameras": cameras, "mc": mc, "oracle": with_oracle}
    for key, flags in (("gather", TSR_PRIORS_ALL_LEAVES), ("traversal", TSR_PRIORS_ALL_LEAVES | TSR_EMPTINESS_TRAVERSAL),
                       ("gather2", TSR_PRIORS_ALL_LEAVES)):
        occ = Occupancy(gpu_ctx, tree, sc, flags=flags)
        mesh = DualMarchingCells(gpu_ctx, tree, occ)
        out[key] = (occ.download(), occ.hit_counts(), mesh.download())
        mesh.close()
        occ.close()
    tree.close()
    sc.close()
    return out


def test_gather_equals_traversal(fields):
    (fg, hg, mg), (ft, ht, mt) = fields["gather"], fields["traversal"]
    L = len(fg["occupancy"])
    assert np.array_equal(fg["node_states"], ft["node_states"]), "node flags differ"
    assert hg[0] == ht[0] and hg[1] == ht[1], f"hit counts {hg} vs {ht}"
    assert fg["cone_count"] == ft["cone_count"]
    for k in ("kernel_sums", "cone_lengths"):
        r = rel_err(fg[k][:L], ft[k][:L], floor=1e-30)
        assert np.max(r) <= 1e-6, f"E {k}: max rel err {np.max(r):.3e}"
        assert np.array_equal(fg[k][L:], ft[k][L:]), f"S {k} differ"
    assert np.array_equal(mg["triangles"], mt["triangles"]), "meshes differ"
    print(f"[emptiness] {fields['name']}: E hits {hg[0]}, max rel diff "
          f"{np.max(rel_err(fg['kernel_sums'][:L], ft['kernel_sums'][:L], floor=1e-30)):.2e}", flush=True)


def test_gather_deterministic(fields):
    a, b = fields["gather"][0], fields["gather2"][0]
    for k in ("kernel_sums", "cone_lengths", "priors", "occupancy", "node_states"):
        assert np.array_equal(a[k], b[k], equal_nan=True), f"{k} differs on a rerun"


def test_gather_matches_oracle(fields):
    from oracle.oracle import Oracle
    if not fields["oracle"]:
        pytest.skip("gather vs traversal only: the serial oracle at this size is slow")
    o = Oracle(fields["samples"], fields["cameras"], PARAMS, max_cams=fields["mc"])
    o.build_tree().occupancy()
    of = o.field()
    g = fields["gather"][0]
    assert np.array_equal(g["node_states"], of["node_states"]), "node flags differ from the oracle"
    for k in ("kernel_sums", "cone_lengths"):
        r = rel_err(g[k], of[k], floor=1e-30)
        assert np.max(r) <= 1e-5, f"{k}: max rel err {np.max(r):.3e} vs the oracle"
    assert abs(g["cone_count"] - of["cone_count"]) <= 1e-9 * of["cone_count"]


def test_index_rebuilt_for_other_parameters(gpu_ctx):
    """The tree keeps the direction index built from its scene with its parameters; a field
    computed with another bandwidth factor builds its own (over the tree's node-order links)
    and must still equal the per-link traversal with that bandwidth."""
    from tsr_amd import Occupancy, Scene, Tree
    from tsr_amd._lib import TSR_EMPTINESS_TRAVERSAL, TSR_PRIORS_ALL_LEAVES, default_params
    samples, cameras = scene(1, 0.0625)
    sc = Scene(gpu_ctx, samples, cameras)
    tree = Tree(gpu_ctx, sc)
    params = default_params()
    params.bandwidth_factor = 3.0   # instead of 2
    g = Occupancy(gpu_ctx, tree, sc, params=params, flags=TSR_PRIORS_ALL_LEAVES)
    t = Occupancy(gpu_ctx, tree, sc, params=params, flags=TSR_PRIORS_ALL_LEAVES | TSR_EMPTINESS_TRAVERSAL)
    fg, ft = g.download(), t.download()
    L = len(fg["occupancy"])
    assert np.array_equal(fg["node_states"], ft["node_states"])
    assert np.array_equal(g.hit_counts(), t.hit_counts())
    assert fg["cone_count"] == ft["cone_count"]
    for k in ("kernel_sums", "cone_lengths"):
        assert np.max(rel_err(fg[k][:L], ft[k][:L], floor=1e-30)) <= 1e-6, k
    for x in (t, g, tree, sc):
        x.close()
