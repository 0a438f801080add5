// Crust post-processing after DMC (DualMarchingCells::DualMarchingCells,
// SR/SurfaceExtraction/DualMarchingCells.cpp:26-39), on a mesh handle in place:
//
//   isles    FlexibleMesh::deleteIsolatedGeometry (SR/Geometry/FlexibleMesh.cpp:1411-1426) with
//            IslesEraser (SR/Geometry/IslesEraser.cpp): components of the vertex-edge graph, each
//            named by its lowest vertex index (min-label propagation with pointer jumping on the
//            GPU), sized by vertex count; components below the minimum size lose their vertices
//            and triangles, the rest keep their order (FlexibleMesh.cpp:25-110 offsets).
//   holes    FlexibleMesh::fillHoles (FlexibleMesh.cpp:1580-1932) on the host: border edges in
//            the mesh's edge creation order, chained into rings, each ring closed by rings of
//            new vertices shrinking towards the ring's centre.  Serial per hole, as in the
//            reference; the holes are few and small.
//   normals  Mesh::computeNormalsWeightedByAngles (SR/Geometry/Mesh.cpp:145-202): angle-weighted
//            sum of unit triangle normals per vertex, normalised; on the GPU, summed per vertex
//            in triangle order (FP64) instead of the reference's atomics.
#include <algorithm>
#include <cmath>
#include <unordered_map>
#include <vector>

#include "handles.h"

namespace tsr {

namespace {

constexpr u32 BLOCK = 256;

__global__ void k_iotaLabels(u32 *label, u32 n) {
    const u32 i = blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) label[i] = i;
}

// every triangle edge pulls both ends' labels to the smaller one
__global__ void k_hookEdges(const u32 *__restrict__ tris, u32 T, u32 *label, u32 *changed) {
    const u32 t = blockIdx.x * BLOCK + threadIdx.x;
    if (t >= T) return;
    for (int k = 0; k < 3; ++k) {
        const u32 a = tris[3 * (size_t)t + k], b = tris[3 * (size_t)t + (k + 1) % 3];
        const u32 la = label[a], lb = label[b];
        if (la == lb) continue;
        const u32 m = min(la, lb);
        const u32 big = max(la, lb);
        // link the larger root to the smaller label and the vertices themselves
        if (atomicMin(&label[big], m) > m) atomicOr(changed, 1u);
        if (atomicMin(&label[a], m) > m) atomicOr(changed, 1u);
        if (atomicMin(&label[b], m) > m) atomicOr(changed, 1u);
    }
}

__global__ void k_jump(u32 *label, u32 n, u32 *changed) {
    const u32 v = blockIdx.x * BLOCK + threadIdx.x;
    if (v >= n) return;
    const u32 l = label[v], ll = label[l];
    if (ll < l) {
        label[v] = ll;
        atomicOr(changed, 1u);
    }
}

__global__ void k_isleSizes(const u32 *__restrict__ label, u32 n, u32 *size) {
    const u32 v = blockIdx.x * BLOCK + threadIdx.x;
    if (v < n) atomicAdd(&size[label[v]], 1u);
}

__global__ void k_keepVertex(const u32 *__restrict__ label, const u32 *__restrict__ size, u32 n, u32 minSize,
                             u32 *keep) {
    const u32 v = blockIdx.x * BLOCK + threadIdx.x;
    if (v < n) keep[v] = size[label[v]] >= minSize ? 1u : 0u;
}

__global__ void k_keepTriangle(const u32 *__restrict__ tris, u32 T, const u32 *__restrict__ keepV, u32 *keepT) {
    const u32 t = blockIdx.x * BLOCK + threadIdx.x;
    if (t < T)
        keepT[t] = keepV[tris[3 * (size_t)t]] & keepV[tris[3 * (size_t)t + 1]] & keepV[tris[3 * (size_t)t + 2]];
}

__global__ void k_compactVertices(const u32 *__restrict__ keep, const u32 *__restrict__ newIdx, u32 n,
                                  const double *__restrict__ pos, const double *__restrict__ scale,
                                  const u32 *__restrict__ keys, double *pos2, double *scale2, u32 *keys2) {
    const u32 v = blockIdx.x * BLOCK + threadIdx.x;
    if (v >= n || !keep[v]) return;
    const u32 j = newIdx[v];
    for (int a = 0; a < 3; ++a) pos2[3 * (size_t)j + a] = pos[3 * (size_t)v + a];
    scale2[j] = scale[v];
    keys2[2 * (size_t)j] = keys[2 * (size_t)v];
    keys2[2 * (size_t)j + 1] = keys[2 * (size_t)v + 1];
}

__global__ void k_compactTriangles(const u32 *__restrict__ keep, const u32 *__restrict__ newT, u32 T,
                                   const u32 *__restrict__ tris, const u32 *__restrict__ newV, u32 *tris2) {
    const u32 t = blockIdx.x * BLOCK + threadIdx.x;
    if (t >= T || !keep[t]) return;
    for (int k = 0; k < 3; ++k) tris2[3 * (size_t)newT[t] + k] = newV[tris[3 * (size_t)t + k]];
}

// per triangle corner: (vertex, weighted unit normal) records, Mesh.cpp:145-199
__global__ void k_cornerNormals(const u32 *__restrict__ tris, u32 T, const double *__restrict__ pos,
                                u32 *__restrict__ vtx, double *__restrict__ wn) {
    const u32 t = blockIdx.x * BLOCK + threadIdx.x;
    if (t >= T) return;
    double p[3][3];
    u32 v[3];
    for (int k = 0; k < 3; ++k) {
        v[k] = tris[3 * (size_t)t + k];
        for (int a = 0; a < 3; ++a) p[k][a] = pos[3 * (size_t)v[k] + a];
    }
    double d0[3], d1[3], d2[3];
    for (int a = 0; a < 3; ++a) {
        d0[a] = p[1][a] - p[0][a];
        d1[a] = p[2][a] - p[1][a];
        d2[a] = p[2][a] - p[0][a];
    }
    double nrm[3] = {d0[1] * d2[2] - d0[2] * d2[1], d0[2] * d2[0] - d0[0] * d2[2], d0[0] * d2[1] - d0[1] * d2[0]};
    auto unit = [](double *x) {
        const double l = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
        if (l > 0.0)
            for (int a = 0; a < 3; ++a) x[a] /= l;
    };
    unit(d0);
    unit(d1);
    unit(d2);
    unit(nrm);
    const double c[3] = {d0[0] * d2[0] + d0[1] * d2[1] + d0[2] * d2[2],
                         -(d0[0] * d1[0] + d0[1] * d1[1] + d0[2] * d1[2]),
                         d1[0] * d2[0] + d1[1] * d2[1] + d1[2] * d2[2]};
    for (int k = 0; k < 3; ++k) {
        const double ang = acos(fmin(1.0, fmax(-1.0, c[k])));
        const size_t r = 3 * (size_t)t + k;
        vtx[r] = v[k];
        for (int a = 0; a < 3; ++a) wn[3 * r + a] = nrm[a] * ang;
    }
}

// per vertex: sum of its corner records in triangle order (the sort by vertex is stable)
__global__ void k_sumNormals(const u32 *__restrict__ vSorted, const u32 *__restrict__ rSorted, u32 m,
                             const double *__restrict__ wn, const u32 *__restrict__ start, u32 V,
                             double *__restrict__ normals) {
    const u32 v = blockIdx.x * BLOCK + threadIdx.x;
    if (v >= V) return;
    double s[3] = {0.0, 0.0, 0.0};
    for (u32 j = start[v]; j < m && vSorted[j] == v; ++j)
        for (int a = 0; a < 3; ++a) s[a] += wn[3 * (size_t)rSorted[j] + a];
    const double l = sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
    for (int a = 0; a < 3; ++a) normals[3 * (size_t)v + a] = l > 0.0 ? s[a] / l : 0.0;
}

__global__ void k_firstOfVertex(const u32 *__restrict__ vSorted, u32 m, u32 *start) {
    const u32 j = blockIdx.x * BLOCK + threadIdx.x;
    if (j < m && (j == 0 || vSorted[j] != vSorted[j - 1])) start[vSorted[j]] = j;
}

__global__ void k_fillU32c(u32 *p, u32 n, u32 v) {
    const u32 i = blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) p[i] = v;
}

// ---- hole filling (host) ----------------------------------------------------------------
struct HostMesh {
    std::vector<double> pos, scale;
    std::vector<u32> keys, tris;
    u32 vertexCount() const { return (u32)scale.size(); }
    void addVertex(const double p[3]) {
        pos.insert(pos.end(), p, p + 3);
        scale.push_back(0.0);   // FlexibleMesh::resize: new vertices have default attributes
        keys.push_back(INVALID);
        keys.push_back(INVALID);
    }
    void addTriangle(u32 a, u32 b, u32 c) {
        tris.push_back(a);
        tris.push_back(b);
        tris.push_back(c);
    }
};

// FlexibleMesh.cpp:1734-1932: rings of ringSize / 2 new vertices towards the centre
void fillHole(HostMesh &m, const u32 *border, u32 count, const double centre[3]);

void fillSmallHole(HostMesh &m, const u32 *border, u32 count) {   // FlexibleMesh.cpp:1902-1931
    u32 idx[5];
    const u32 start = m.vertexCount() - count;
    for (u32 i = 0; i < count; ++i) idx[i] = border ? border[i] : start + i;
    m.addTriangle(idx[0], idx[1], idx[2]);
    if (count == 3) return;
    m.addTriangle(idx[2], idx[3], idx[0]);
    if (count == 4) return;
    m.addTriangle(idx[3], idx[4], idx[0]);
}

void fillHole(HostMesh &m, const u32 *border, u32 count, const double centre[3]) {
    if (count < 3) throw tsr_error(TSR_ERR_INVALID_ARGUMENT, "hole border with fewer than 3 edges");
    if (count <= 5) return fillSmallHole(m, border, count);
    const u32 oldV = m.vertexCount(), add = count / 2, newV = oldV + add;
    // createHoleFillingRingVertices (FlexibleMesh.cpp:1785-1822)
    u32 rings = 0;
    for (u32 t = count; t; t >>= 1) ++rings;
    rings -= 1;
    const double f0 = (rings - 1) / (double)rings, f1 = 1.0 / rings;
    const u32 prevStart = oldV - count, last = newV - 1;
    auto bv = [&](u32 i) { return border ? border[i] : prevStart + i; };
    u32 b = 0;
    for (u32 v = oldV; v < newV; ++v) {
        const u32 v0 = bv(b++), v1 = bv(b++);
        const u32 v2 = v == last ? bv(0) : bv(b);
        double p[3];
        for (int a = 0; a < 3; ++a)
            p[a] = (m.pos[3 * (size_t)v0 + a] + m.pos[3 * (size_t)v1 + a] + m.pos[3 * (size_t)v2 + a]) * (f0 / 3.0) +
                   centre[a] * f1;
        m.addVertex(p);
    }
    // connectHoleFillingRings (FlexibleMesh.cpp:1824-1900)
    const bool odd = count & 1;
    for (u32 outer = 0, inner = 0; inner < add; ++inner) {
        const u32 o0 = bv(outer++), o1 = bv(outer++);
        const u32 o2 = (!odd && inner == add - 1) ? bv(0) : bv(outer);
        m.addTriangle(o0, o1, oldV + inner);
        m.addTriangle(o1, o2, oldV + inner);
    }
    if (odd) m.addTriangle(bv(count - 1), bv(0), m.vertexCount() - 1);
    for (u32 outer = 2, inner = 0; inner < add; outer += 2) {
        const u32 i0 = oldV + inner;
        ++inner;
        u32 i1, ov;
        if (inner == add) {
            i1 = oldV;
            ov = bv(0);
        } else {
            i1 = oldV + inner;
            ov = bv(outer);
        }
        m.addTriangle(i0, ov, i1);
    }
    fillHole(m, nullptr, add, centre);
}

// FlexibleMesh.cpp:1580-1732: border edges in edge creation order, their rings, the fill
void fillHoles(HostMesh &m) {
    const u32 T = (u32)(m.tris.size() / 3);
    // edges in creation order (addTriangle -> addEdge for local edges 0, 1, 2)
    std::unordered_map<u64, u32> edgeId;
    edgeId.reserve((size_t)T * 2);
    std::vector<u32> edgeTri0, edgeTri1;
    std::vector<u64> edgeKey;
    for (u32 t = 0; t < T; ++t)
        for (int k = 0; k < 3; ++k) {
            const u32 a = m.tris[3 * (size_t)t + k], b = m.tris[3 * (size_t)t + (k + 1) % 3];
            const u64 key = ((u64)std::min(a, b) << 32) | std::max(a, b);
            auto it = edgeId.find(key);
            if (it == edgeId.end()) {
                edgeId.emplace(key, (u32)edgeKey.size());
                edgeKey.push_back(key);
                edgeTri0.push_back(t);
                edgeTri1.push_back(INVALID);
            } else if (edgeTri1[it->second] == INVALID) {
                edgeTri1[it->second] = t;
            } else {
                throw tsr_error(TSR_ERR_NON_MANIFOLD_EDGE, "BOOM: a crust edge is shared by more than two triangles");
            }
        }
    std::vector<u32> border;
    for (size_t e = 0; e < edgeKey.size(); ++e) {
        if (edgeTri1[e] != INVALID) continue;
        const u32 *tri = &m.tris[3 * (size_t)edgeTri0[e]];
        const u32 u[2] = {(u32)(edgeKey[e] >> 32), (u32)edgeKey[e]};
        // Triangle::getVerticesInWindingOrder (Triangle.cpp:104-124)
        bool done = false;
        for (int dir = 0; dir < 2 && !done; ++dir) {
            const u32 o0 = u[dir], o1 = u[!dir];
            for (int c = 0; c < 3; ++c)
                if (o0 == tri[c] && o1 == tri[(c + 1) % 3]) {
                    border.push_back(o0);
                    border.push_back(o1);
                    done = true;
                    break;
                }
        }
    }
    if (border.empty()) return;
    // findBorderRings / findBorderRing / extendBorderRing (FlexibleMesh.cpp:1616-1710)
    std::vector<std::vector<u32>> sets(1);
    sets[0].swap(border);
    sets.resize(2);
    for (size_t s = 0; s + 1 < sets.size(); ++s) {
        std::vector<u32> &set = sets[s];
        std::vector<u32> &rest = sets.back();
        const u32 ringStart = set[0];
        u32 ringSize = 2;
        const u32 n = (u32)set.size();
        for (u32 v0 = ringSize; v0 < n;) {
            const u32 v1 = v0 + 1;
            const u32 e0 = set[v0], e1 = set[v1];
            if (e0 != set[ringSize - 1]) {
                v0 += 2;
                continue;
            }
            set[v0] = set[ringSize];
            set[v1] = set[ringSize + 1];
            set[ringSize] = e0;
            set[ringSize + 1] = e1;
            ringSize += 2;
            v0 = ringSize;
            if (e1 != ringStart) continue;
            if (n - ringSize == 0) break;
            rest.assign(set.begin() + ringSize, set.end());
            set.resize(ringSize);
            break;
        }
        if (!sets.back().empty()) sets.resize(sets.size() + 1);
    }
    if (sets.back().empty()) sets.pop_back();
    // removeDuplicatesInRings: every second entry is a ring vertex
    for (std::vector<u32> &r : sets) {
        u32 w = 0;
        for (size_t i = 0; i < r.size(); i += 2) r[w++] = r[i];
        r.resize(w);
    }
    for (const std::vector<u32> &r : sets) {
        double c[3] = {0.0, 0.0, 0.0};
        for (u32 v : r)
            for (int a = 0; a < 3; ++a) c[a] += m.pos[3 * (size_t)v + a];
        for (int a = 0; a < 3; ++a) c[a] /= (double)r.size();
        fillHole(m, r.data(), (u32)r.size(), c);
    }
}

}  // namespace

void postprocessCrust(tsr_context *ctx, tsr_mesh *mesh, u32 minIsleSize, bool fill, bool normals) {
    hipStream_t st = ctx->stream;
    Scratch &sc = ctx->scratch;
    u32 V = mesh->vertexCount, T = mesh->triangleCount;
    // ---- isles (FlexibleMesh::deleteIsolatedGeometry)
    if (V && minIsleSize > 1) {
        DBuf<u32> label(V), changed(1), size(V), keepV(V), newV(V), keepT(std::max(T, 1u)), newT(std::max(T, 1u));
        k_iotaLabels<<<gridFor(V, BLOCK), BLOCK, 0, st>>>(label.p, V);
        for (int iter = 0;; ++iter) {
            if (iter > 4096) throw tsr_error(TSR_ERR_HIP, "isle labelling did not converge");
            HIPCHECK(hipMemsetAsync(changed.p, 0, 4, st));
            if (T) k_hookEdges<<<gridFor(T, BLOCK), BLOCK, 0, st>>>(mesh->tris.p, T, label.p, changed.p);
            for (int j = 0; j < 4; ++j) k_jump<<<gridFor(V, BLOCK), BLOCK, 0, st>>>(label.p, V, changed.p);
            if (!readScalar(changed.p, st)) break;
        }
        HIPCHECK(hipMemsetAsync(size.p, 0, (size_t)V * 4, st));
        k_isleSizes<<<gridFor(V, BLOCK), BLOCK, 0, st>>>(label.p, V, size.p);
        k_keepVertex<<<gridFor(V, BLOCK), BLOCK, 0, st>>>(label.p, size.p, V, minIsleSize, keepV.p);
        exclusiveScan(sc, keepV.p, newV.p, V, st);
        const u32 V2 = readScalar(newV.p + V - 1, st) + readScalar(keepV.p + V - 1, st);
        if (V2 < V) {
            u32 T2 = 0;
            if (T) {
                k_keepTriangle<<<gridFor(T, BLOCK), BLOCK, 0, st>>>(mesh->tris.p, T, keepV.p, keepT.p);
                exclusiveScan(sc, keepT.p, newT.p, T, st);
                T2 = readScalar(newT.p + T - 1, st) + readScalar(keepT.p + T - 1, st);
            }
            DBuf<double> pos2(3 * (size_t)std::max(V2, 1u)), scale2(std::max(V2, 1u));
            DBuf<u32> keys2(2 * (size_t)std::max(V2, 1u)), tris2(3 * (size_t)std::max(T2, 1u));
            k_compactVertices<<<gridFor(V, BLOCK), BLOCK, 0, st>>>(keepV.p, newV.p, V, mesh->pos.p, mesh->scale.p,
                                                                   mesh->keys.p, pos2.p, scale2.p, keys2.p);
            if (T)
                k_compactTriangles<<<gridFor(T, BLOCK), BLOCK, 0, st>>>(keepT.p, newT.p, T, mesh->tris.p, newV.p,
                                                                        tris2.p);
            HIPCHECK(hipStreamSynchronize(st));
            mesh->pos = std::move(pos2);
            mesh->scale = std::move(scale2);
            mesh->keys = std::move(keys2);
            mesh->tris = std::move(tris2);
            V = mesh->vertexCount = V2;
            T = mesh->triangleCount = T2;
        }
    }
    // DualMarchingCells.cpp:31-32: an empty crust is left as it is
    if (V == 0 || T == 0) {
        mesh->normals.release();
        return;
    }
    // ---- holes (host)
    if (fill) {
        HostMesh h;
        h.pos.resize(3 * (size_t)V);
        h.scale.resize(V);
        h.keys.resize(2 * (size_t)V);
        h.tris.resize(3 * (size_t)T);
        HIPCHECK(hipMemcpyAsync(h.pos.data(), mesh->pos.p, h.pos.size() * 8, hipMemcpyDeviceToHost, st));
        HIPCHECK(hipMemcpyAsync(h.scale.data(), mesh->scale.p, h.scale.size() * 8, hipMemcpyDeviceToHost, st));
        HIPCHECK(hipMemcpyAsync(h.keys.data(), mesh->keys.p, h.keys.size() * 4, hipMemcpyDeviceToHost, st));
        HIPCHECK(hipMemcpyAsync(h.tris.data(), mesh->tris.p, h.tris.size() * 4, hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
        fillHoles(h);
        if (h.vertexCount() != V || h.tris.size() != 3 * (size_t)T) {
            V = mesh->vertexCount = h.vertexCount();
            T = mesh->triangleCount = (u32)(h.tris.size() / 3);
            mesh->pos.alloc(h.pos.size());
            mesh->scale.alloc(h.scale.size());
            mesh->keys.alloc(h.keys.size());
            mesh->tris.alloc(h.tris.size());
            HIPCHECK(hipMemcpyAsync(mesh->pos.p, h.pos.data(), h.pos.size() * 8, hipMemcpyHostToDevice, st));
            HIPCHECK(hipMemcpyAsync(mesh->scale.p, h.scale.data(), h.scale.size() * 8, hipMemcpyHostToDevice, st));
            HIPCHECK(hipMemcpyAsync(mesh->keys.p, h.keys.data(), h.keys.size() * 4, hipMemcpyHostToDevice, st));
            HIPCHECK(hipMemcpyAsync(mesh->tris.p, h.tris.data(), h.tris.size() * 4, hipMemcpyHostToDevice, st));
            HIPCHECK(hipStreamSynchronize(st));
        }
    }
    // ---- normals (Mesh::computeNormalsWeightedByAngles)
    if (normals) {
        const u32 m3 = 3 * T;
        DBuf<u32> vtx(m3), rid(m3), vS(m3), rS(m3), start(V);
        DBuf<double> wn(3 * (size_t)m3);
        k_cornerNormals<<<gridFor(T, BLOCK), BLOCK, 0, st>>>(mesh->tris.p, T, mesh->pos.p, vtx.p, wn.p);
        k_iotaLabels<<<gridFor(m3, BLOCK), BLOCK, 0, st>>>(rid.p, m3);
        int bits = 1;
        while (bits < 32 && (1ull << bits) < V) ++bits;
        sortPairs(sc, vtx.p, vS.p, rid.p, rS.p, m3, 0, bits, st);
        k_fillU32c<<<gridFor(V, BLOCK), BLOCK, 0, st>>>(start.p, V, m3);
        k_firstOfVertex<<<gridFor(m3, BLOCK), BLOCK, 0, st>>>(vS.p, m3, start.p);
        mesh->normals.alloc(3 * (size_t)V);
        k_sumNormals<<<gridFor(V, BLOCK), BLOCK, 0, st>>>(vS.p, rS.p, m3, wn.p, start.p, V, mesh->normals.p);
        HIPCHECK(hipStreamSynchronize(st));
    }
}

}  // namespace tsr
