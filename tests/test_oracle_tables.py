"""Marching-cubes tables: derivable invariants (oracle and product encodings)."""
import os
import re

import numpy as np

from oracle.oracle import mc_case, mc_edges

EDGE_ENDS = [(0, 1), (1, 2), (2, 3), (3, 0), (4, 5), (5, 6), (6, 7), (7, 4), (0, 4), (1, 5), (2, 6), (3, 7)]
# Bourke corner positions (x, y, z) for corner index 0..7
CORNER_POS = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]


def tris(cfg):
    s = mc_case(cfg)
    e = [int(c, 16) for c in s]
    return [tuple(e[i:i + 3]) for i in range(0, len(e), 3)]


def test_edge_table_known_values():
    # first entries of Bourke's edgeTable (public domain)
    assert [mc_edges(c) for c in range(4)] == [0x000, 0x109, 0x203, 0x30a]
    assert mc_edges(255) == 0 and mc_edges(0xF0) == mc_edges(0x0F)


def test_triangles_use_exactly_the_cut_edges():
    for cfg in range(256):
        used = set(e for t in tris(cfg) for e in t)
        cut = {e for e in range(12) if mc_edges(cfg) >> e & 1}
        assert used == cut, cfg
        assert len(mc_case(cfg)) % 3 == 0


def _face_of(e0, e1):
    """cube face shared by two cube edges, or None"""
    pts = [CORNER_POS[c] for c in EDGE_ENDS[e0] + EDGE_ENDS[e1]]
    for a in range(3):
        for v in (0, 1):
            if all(p[a] == v for p in pts):
                return (a, v)
    return None


def test_triangulation_is_a_manifold_patch():
    # interior triangle edges are shared by exactly two triangles, edges on a cube
    # face by exactly one: the surface is closed inside the cube
    for cfg in range(256):
        count = {}
        for t in tris(cfg):
            assert len(set(t)) == 3, cfg
            for k in range(3):
                e = tuple(sorted((t[k], t[(k + 1) % 3])))
                count[e] = count.get(e, 0) + 1
        for e, c in count.items():
            if _face_of(*e) is None:
                assert c == 2, (cfg, e, c)
            else:
                assert c == 1, (cfg, e, c)


def test_product_table_matches_oracle_table():
    src = open(os.path.join(os.path.dirname(__file__), "..", "tsr_amd", "csrc", "dmc.hip")).read()
    body = src[src.index("TRI_TABLE[256] = {"):]
    body = body[:body.index("};")]
    words = [int(w, 16) for w in re.findall(r"0x([0-9a-f]{16})ull", body)]
    assert len(words) == 256
    for cfg, w in enumerate(words):
        n = w & 0xF
        edges = "".join("%x" % ((w >> (4 * (j + 1))) & 0xF) for j in range(n))
        assert edges == mc_case(cfg), cfg
