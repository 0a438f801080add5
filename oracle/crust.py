"""CPU restatement of the crust post-processing of DualMarchingCells' constructor --
TEST INFRASTRUCTURE ONLY (the product never imports it).

Follows (paths relative to SurfaceReconstruction/SurfaceReconstruction/):
  delete_isles  FlexibleMesh::deleteIsolatedGeometry, FlexibleMesh.cpp:1411-1426, with
                IslesEraser.cpp:30-200 (components of the vertex-edge graph named by their
                lowest vertex, sized in vertices) and the order-keeping compaction of
                FlexibleMesh::computeOffsetsForFiltering / deleteGeometry (FlexibleMesh.cpp:25-110)
  fill_holes    FlexibleMesh::fillHoles, FlexibleMesh.cpp:1580-1932 (single-threaded edge order)
  normals       Mesh::computeNormalsWeightedByAngles, Mesh.cpp:145-202
Positions are FP64 here (the reference's Real is float).  Pure Python loops: small meshes only.
"""
import numpy as np

INVALID = 0xFFFFFFFF


def delete_isles(pos, scales, keys, tris, min_size):
    """returns (pos, scales, keys, tris) without the vertex components smaller than min_size"""
    V = len(scales)
    parent = list(range(V))

    def find(a):
        while parent[a] != a:
            parent[a] = parent[parent[a]]
            a = parent[a]
        return a
    t = tris.reshape(-1, 3)
    for tri in t:
        for k in range(3):
            a, b = find(int(tri[k])), find(int(tri[(k + 1) % 3]))
            if a != b:
                parent[max(a, b)] = min(a, b)
    root = np.array([find(v) for v in range(V)], np.int64)
    size = np.bincount(root, minlength=V)
    keep_v = size[root] >= min_size
    if keep_v.all():
        return pos, scales, keys, tris
    new_idx = np.cumsum(keep_v) - 1
    keep_t = keep_v[t].all(axis=1)
    p = pos.reshape(-1, 3)[keep_v].reshape(-1)
    k = keys.reshape(-1, 2)[keep_v].reshape(-1)
    return p, scales[keep_v], k, new_idx[t[keep_t]].astype(np.uint32).reshape(-1)


def _fill_small(tris, border, count, vcount):
    idx = list(border) if border is not None else [vcount - count + i for i in range(count)]
    tris += [idx[0], idx[1], idx[2]]
    if count == 3:
        return
    tris += [idx[2], idx[3], idx[0]]
    if count == 4:
        return
    tris += [idx[3], idx[4], idx[0]]


def _fill_hole(pos, scales, keys, tris, border, count, centre):
    """FlexibleMesh::fillHole(borderVertices, borderEdgeCount, center), FlexibleMesh.cpp:1766-1900"""
    if count <= 5:
        _fill_small(tris, border, count, len(scales))
        return
    old_v = len(scales)
    add = count // 2
    rings = count.bit_length() - 1
    f0, f1 = (rings - 1) / rings, 1.0 / rings
    prev = old_v - count

    def bv(i):
        return border[i] if border is not None else prev + i
    b = 0
    for v in range(old_v, old_v + add):
        v0, v1 = bv(b), bv(b + 1)
        b += 2
        v2 = bv(0) if v == old_v + add - 1 else bv(b)
        p = (pos[v0] + pos[v1] + pos[v2]) * (f0 / 3.0) + centre * f1
        pos.append(p)
        scales.append(0.0)
        keys += [INVALID, INVALID]
    odd = count & 1
    outer = 0
    for inner in range(add):
        o0, o1 = bv(outer), bv(outer + 1)
        outer += 2
        o2 = bv(0) if (not odd and inner == add - 1) else bv(outer)
        tris += [o0, o1, old_v + inner, o1, o2, old_v + inner]
    if odd:
        tris += [bv(count - 1), bv(0), len(scales) - 1]
    outer = 2
    for inner in range(add):
        i0 = old_v + inner
        if inner + 1 == add:
            i1, ov = old_v, bv(0)
        else:
            i1, ov = old_v + inner + 1, bv(outer)
        tris += [i0, ov, i1]
        outer += 2
    _fill_hole(pos, scales, keys, tris, None, add, centre)


def fill_holes(pos, scales, keys, tris):
    """returns (pos, scales, keys, tris) with the holes filled"""
    t = [int(x) for x in tris]
    T = len(t) // 3
    edge_of, ends, first, second = {}, [], [], []
    for ti in range(T):                       # FlexibleMesh::addTriangle -> addEdge order
        for k in range(3):
            a, b = t[3 * ti + k], t[3 * ti + (k + 1) % 3]
            e = (min(a, b), max(a, b))
            if e not in edge_of:
                edge_of[e] = len(ends)
                ends.append(e)
                first.append(ti)
                second.append(None)
            else:
                second[edge_of[e]] = ti
    border = []
    for ei, e in enumerate(ends):
        if second[ei] is not None:
            continue
        tri = t[3 * first[ei]:3 * first[ei] + 3]
        for d in range(2):                    # Triangle::getVerticesInWindingOrder
            o = (e[d], e[1 - d])
            if any(o[0] == tri[c] and o[1] == tri[(c + 1) % 3] for c in range(3)):
                border += [o[0], o[1]]
                break
    p = [np.array(x, np.float64) for x in pos.reshape(-1, 3)]
    s = [float(x) for x in scales]
    kk = [int(x) for x in keys]
    if not border:
        return pos, scales, keys, tris
    sets = [border, []]                       # findBorderRings
    si = 0
    while si < len(sets) - 1:
        es, rest = sets[si], sets[-1]
        start, size, n = es[0], 2, len(es)
        v0 = size
        while v0 < n:                         # findBorderRing + extendBorderRing
            e0, e1 = es[v0], es[v0 + 1]
            if e0 != es[size - 1]:
                v0 += 2
                continue
            es[v0], es[v0 + 1] = es[size], es[size + 1]
            es[size], es[size + 1] = e0, e1
            size += 2
            v0 = size
            if e1 != start:
                continue
            if n - size:
                rest.extend(es[size:])
                del es[size:]
            break
        if sets[-1]:
            sets.append([])
        si += 1
    if not sets[-1]:
        sets.pop()
    rings = [r[0::2] for r in sets]           # removeDuplicatesInRings
    for r in rings:
        centre = sum(p[v] for v in r) / float(len(r))
        _fill_hole(p, s, kk, t, r, len(r), centre)
    return (np.array(p, np.float64).reshape(-1), np.array(s, np.float64), np.array(kk, np.uint32),
            np.array(t, np.uint32))


def normals(pos, tris):
    """angle-weighted unit vertex normals (3 per vertex)"""
    P = pos.reshape(-1, 3)
    t = tris.reshape(-1, 3).astype(np.int64)
    out = np.zeros_like(P)
    for tri in t:
        v = P[tri]
        d0, d1, d2 = v[1] - v[0], v[2] - v[1], v[2] - v[0]
        nrm = np.cross(d0, d2)

        def unit(x):
            n = np.sqrt(x @ x)
            return x / n if n > 0 else x
        d0, d1, d2, nrm = unit(d0), unit(d1), unit(d2), unit(nrm)
        c = [d0 @ d2, -(d0 @ d1), d1 @ d2]
        for k in range(3):
            out[tri[k]] += nrm * np.arccos(min(1.0, max(-1.0, c[k])))
    ln = np.sqrt((out ** 2).sum(1))
    ok = ln > 0
    out[ok] /= ln[ok][:, None]
    return out.reshape(-1)


def postprocess(mesh, min_isle_size):
    """DualMarchingCells.cpp:26-39 after extractMesh, on a mesh dict (positions, scales,
    edge_keys, triangles); returns the post-processed dict with normals"""
    p, s, k, t = delete_isles(mesh["positions"], mesh["scales"], mesh["edge_keys"], mesh["triangles"],
                              min_isle_size)
    if len(s) == 0 or len(t) == 0:
        return {"positions": p, "scales": s, "edge_keys": k, "triangles": t, "normals": np.zeros(3 * len(s))}
    p, s, k, t = fill_holes(p, s, k, t)
    return {"positions": p, "scales": s, "edge_keys": k, "triangles": t, "normals": normals(p, t)}
