// CPU oracle for the TSR occupancy-field + dual-marching-cells path.
//
// TEST INFRASTRUCTURE ONLY.  This file restates, in plain serial C++, the
// reference algorithm of SamirAroudj/TSR for the scene octree, the occupancy
// scalar field and DMC crust extraction.  It is the checker that the HIP
// product is compared against; the product never links, loads or calls it.
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it.
//
// Every function cites the reference file:line it follows (paths relative to
// SurfaceReconstruction/SurfaceReconstruction/).  Where the reference depends
// on the un-vendored BaseProject library (ObliqueCircularCone, EPSILON,
// intersection helpers, Plane) the definitions used here are fixed in the
// block "BaseProject stand-ins" below; the product uses the same definitions.
//
// Geometry frame.  Node boundaries are B(k,d) = rootMin + k * (rootSize*2^-d)
// evaluated in FP64 (one rounding), node centres are B(2k+1,d+1).  The
// reference builds coordinates by repeated float additions (Nodes.cpp:46-58);
// the closed form makes every comparison reproducible bit for bit.
//
// Two numeric modes:
//   * Real = double, serial:   the parity oracle (deterministic).
//   * Real = float, OpenMP:    the CPU baseline that mirrors the reference's
//     own parallel structure (Occupancy.cpp:171 dynamic,128 + omp atomic;
//     Occupancy.cpp:540 dynamic,512).
//
// Parity pins (tests/test_oracle_*.py): closed-form kernel integrals, cone
// invariants, brute-force traversal on tiny trees, tree invariants (container
// size rule, 2:1 balance, DFS numbering, contiguous sample ranges), Schaefer-
// Warren cell enumeration vs the leaf-corner characterisation, MC tables vs
// their derivable invariants, marching-cubes known answer on a uniform tree.
// parity unpinned: ObliqueCircularCone::{getTCutCone,getLength,isDegenerated,
// haveSeparatingAxis}, Math::EPSILON's value, intersectAABBWithSphere /
// intersectAABBWithLineSegment boundary conventions (BaseProject absent).

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef uint32_t u32;
typedef uint64_t u64;

namespace {

// ---------------------------------------------------------------------------
// BaseProject stand-ins (Math::EPSILON, REAL_MAX, ObliqueCircularCone, ...)
// ---------------------------------------------------------------------------
const float EPS_F = 1e-6f;                 // Math::EPSILON (value unpinned)
const double EPS_D = (double)EPS_F;
const u32 INVALID = 0xffffffffu;           // Nodes.cpp:30
const u32 MAX_TREE_DEPTH = 19;             // location codes + depth fit 64-bit sort keys
const u32 MAX_DEPTH_DIFFERENCE = 3;        // Occupancy.cpp:72
const u32 FLAG_S = 1u, FLAG_E = 2u;        // Occupancy.h:36-42
const double PI_D = 3.14159265358979323846;

struct Params {
    double bandwidthFactor;           // App.cfg:63-73
    double confidenceThreshold;
    double relativeRadiusForPriors;
    double relativeSampleConeLength;
    double maxRelativeSamplingDistance;
};

// ---------------------------------------------------------------------------
// Input samples (float32 SoA as handed over by the caller)
// ---------------------------------------------------------------------------
struct Samples {
    u32 n = 0, maxCams = 1, camCount = 0;
    std::vector<float> px, py, pz, nx, ny, nz, scale, conf;
    std::vector<u32> cams;       // n * maxCams parent camera links
    std::vector<float> cx, cy, cz;
    float maxRel = 2.0f;
    // Samples.h:409-412 getMaxSamplingDistance (Real = float arithmetic)
    float msd(u32 i) const { return scale[i] * maxRel; }
};

// ---------------------------------------------------------------------------
// Octree (Nodes.h:143-147 arrays + integer coordinates for the FP64 frame)
// ---------------------------------------------------------------------------
struct Tree {
    float rootMinF[3] = {0, 0, 0}, rootSizeF = 0;
    double rootMin[3] = {0, 0, 0}, rootSize = 0;
    std::vector<u32> children, parents, counts, starts;
    std::vector<u32> kx, ky, kz;
    std::vector<uint8_t> depth;
    // leaves (Leaves.h:86-99)
    std::vector<u32> leafNode, nodeToLeaf;
    // dual cells (DualCells.h:67)
    std::vector<u32> cells;
    // sample permutation produced by reorderSamplesAndFillNodes
    std::vector<u32> sampleTarget;
    std::vector<u32> sampleDepth;   // container depth per ORIGINAL sample idx
    // statistics of the sampling pass
    u64 samplingCreatingUnits = 0, samplingNodesCreated = 0;

    double h(u32 d) const { return std::ldexp(rootSize, -(int)d); }
    double B(int a, u64 k, u32 d) const { return rootMin[a] + (double)k * h(d); }
    bool isLeaf(u32 n) const { return n != INVALID && children[n] == INVALID; }
    u32 count() const { return (u32)children.size(); }
    u64 kc(int a, u32 n) const { return a == 0 ? kx[n] : a == 1 ? ky[n] : kz[n]; }
    double centre(int a, u32 n) const { return B(a, 2 * kc(a, n) + 1, depth[n] + 1u); }
    double size(u32 n) const { return h(depth[n]); }

    u32 newNode(u32 parent, u64 x, u64 y, u64 z, u32 d) {
        children.push_back(INVALID); parents.push_back(parent);
        counts.push_back(0); starts.push_back(INVALID);
        kx.push_back((u32)x); ky.push_back((u32)y); kz.push_back((u32)z);
        depth.push_back((uint8_t)d);
        return (u32)children.size() - 1;
    }
    // Nodes.cpp:397-415 createChildren; child slot bits x=1, y=2, z=4 (Nodes.cpp:46-58)
    void createChildren(u32 n) {
        if ((u64)count() + 8 >= (u64)INVALID)
            throw std::runtime_error("Required scene tree node count is larger than supported.");
        if (depth[n] + 1u > MAX_TREE_DEPTH)
            throw std::runtime_error("octree deeper than 19 levels");
        u32 first = count();
        for (u32 c = 0; c < 8; ++c)
            newNode(n, 2ull * kx[n] + (c & 1), 2ull * ky[n] + ((c >> 1) & 1), 2ull * kz[n] + ((c >> 2) & 1),
                    depth[n] + 1u);
        children[n] = first;
    }
};

// Samples.cpp:451-455 getDistanceToPlane: signed distance to the plane through
// the sample with its normal (Plane convention fixed here: dot(x - p, n)).
inline double planeDistance(const Samples &s, u32 i, double x, double y, double z) {
    return (x - (double)s.px[i]) * (double)s.nx[i] + (y - (double)s.py[i]) * (double)s.ny[i] +
           (z - (double)s.pz[i]) * (double)s.nz[i];
}

// ---------------------------------------------------------------------------
// Root scope: Tree.cpp:70-101 + Samples.cpp:290-298,442-449 (Real = float)
// ---------------------------------------------------------------------------
void computeRoot(const Samples &s, Tree &t) {
    float mn[3] = {3.4028235e38f, 3.4028235e38f, 3.4028235e38f};
    float mx[3] = {-3.4028235e38f, -3.4028235e38f, -3.4028235e38f};
    for (u32 i = 0; i < s.n; ++i) {
        const float r = s.msd(i) + EPS_F;
        const float p[3] = {s.px[i], s.py[i], s.pz[i]};
        for (int a = 0; a < 3; ++a) {
            mn[a] = std::min(mn[a], p[a] - r);
            mx[a] = std::max(mx[a], p[a] + r);
        }
    }
    float c[3], e[3];
    for (int a = 0; a < 3; ++a) { c[a] = (mn[a] + mx[a]) * 0.5f; e[a] = mx[a] - mn[a]; }
    float size = std::max(std::max(e[0], e[1]), e[2]);
    float half = size * 0.5f;
    t.rootSizeF = size;
    t.rootSize = (double)size;
    for (int a = 0; a < 3; ++a) {
        t.rootMinF[a] = c[a] - half;
        t.rootMin[a] = (double)t.rootMinF[a];
    }
}

// Tree.cpp:86-91 getSampleDepth, restated exactly: the depth at which
// Nodes::findContainer (Nodes.cpp:250) stops, i.e. the smallest d with
// rootSize * 2^-d <= maxSamplingDistance.
u32 containerDepth(const Tree &t, double msd) {
    for (u32 d = 0; d <= MAX_TREE_DEPTH; ++d)
        if (t.h(d) <= msd) return d;
    throw std::runtime_error("sample needs an octree deeper than 19 levels");
}

// ---------------------------------------------------------------------------
// Node build: Nodes.cpp:196-337
// ---------------------------------------------------------------------------
struct Builder {
    const Samples &s;
    Tree &t;
    Builder(const Samples &s_, Tree &t_) : s(s_), t(t_) {}

    // Nodes.cpp:32-44 getChildIndexForOffset relative to the node centre
    u32 childFor(u32 n, double x, double y, double z) const {
        u32 c = 0;
        if (x - t.centre(0, n) >= 0.0) c |= 1;
        if (y - t.centre(1, n) >= 0.0) c |= 2;
        if (z - t.centre(2, n) >= 0.0) c |= 4;
        return c;
    }

    // Nodes.cpp:231-269 findContainer; returns true and the container in
    // `node`, or false and the leaf where the descent ended.
    bool findContainer(u32 &node, u32 i) const {
        const double x = s.px[i], y = s.py[i], z = s.pz[i];
        const double msd = (double)s.msd(i);
        for (;;) {
            if (t.size(node) <= msd) return true;
            if (t.isLeaf(node)) return false;
            node = t.children[node] + childFor(node, x, y, z);
        }
    }

    // Nodes.cpp:222-229 createNodesForContainment
    void containment(u32 i) {
        u32 node = 0;
        const double x = s.px[i], y = s.py[i], z = s.pz[i];
        for (int a = 0; a < 3; ++a) {
            const double p = a == 0 ? x : a == 1 ? y : z;
            if (!(p >= t.B(a, 0, 0) - EPS_D && p < t.B(a, 1, 0) + EPS_D))
                throw std::runtime_error("Cannot find a container node for a position which is not contained by the start node for the search.");
        }
        while (!findContainer(node, i)) t.createChildren(node);
        ++t.counts[node];
    }

    // Nodes.cpp:471-488 fittingSamplingNodeCenter for the node with integer
    // coordinates (kx,ky,kz) at depth d (its centre is B(2k+1,d+1)).
    bool fitting(u64 x, u64 y, u64 z, u32 d, bool positive, u32 i) const {
        const double cx = t.B(0, 2 * x + 1, d + 1), cy = t.B(1, 2 * y + 1, d + 1), cz = t.B(2, 2 * z + 1, d + 1);
        const double dx = cx - (double)s.px[i], dy = cy - (double)s.py[i], dz = cz - (double)s.pz[i];
        const double distSq = dx * dx + dy * dy + dz * dz;
        const double msd = (double)s.msd(i);
        if (distSq > msd * msd) return false;
        const bool onPositive = planeDistance(s, i, cx, cy, cz) > 0.0;
        return onPositive == positive;
    }

    // Nodes.cpp:490-517 intersect: some corner on the wanted side and the
    // node box overlaps the sample's sphere of radius maxSamplingDistance
    // (intersectAABBWithSphere convention: squared distance <= r^2).
    bool intersect(u32 n, bool positive, u32 i) const {
        const u32 d = t.depth[n];
        bool allWrong = true;
        for (u32 c = 0; c < 8 && allWrong; ++c) {
            const double x = t.B(0, t.kx[n] + (c & 1), d), y = t.B(1, t.ky[n] + ((c >> 1) & 1), d),
                         z = t.B(2, t.kz[n] + ((c >> 2) & 1), d);
            const bool inPositive = planeDistance(s, i, x, y, z) > 0.0;
            if (inPositive == positive) allWrong = false;
        }
        if (allWrong) return false;
        const double p[3] = {s.px[i], s.py[i], s.pz[i]};
        double d2 = 0.0;
        for (int a = 0; a < 3; ++a) {
            const double lo = t.B(a, t.kc(a, n), d), hi = t.B(a, t.kc(a, n) + 1, d);
            if (p[a] < lo) d2 += (lo - p[a]) * (lo - p[a]);
            else if (p[a] > hi) d2 += (p[a] - hi) * (p[a] - hi);
        }
        const double r = (double)s.msd(i);
        return d2 <= r * r;
    }

    // Nodes.cpp:433-469 findReliableSamplingNode
    u32 findReliable(u32 n, bool positive, u32 i, u64 &visits) const {
        ++visits;
        if (t.isLeaf(n)) {
            const u32 d = t.depth[n];
            if (fitting(t.kx[n], t.ky[n], t.kz[n], d, positive, i))
                for (u32 c = 0; c < 8; ++c)
                    if (fitting(2ull * t.kx[n] + (c & 1), 2ull * t.ky[n] + ((c >> 1) & 1),
                                2ull * t.kz[n] + ((c >> 2) & 1), d + 1, positive, i))
                        return n;
            return INVALID;
        }
        const u32 c0 = t.children[n];
        for (u32 c = 0; c < 8; ++c) {
            if (!intersect(c0 + c, positive, i)) continue;
            const u32 r = findReliable(c0 + c, positive, i, visits);
            if (r != INVALID) return r;
        }
        return INVALID;
    }

    // Slab test standing in for CollisionDetection::intersectAABBWithLineSegment
    // (Nodes.cpp:319): parameters clipped to [0,1] along start->end.
    static bool segmentBox(double &tIn, double &tOut, const double st[3], const double en[3],
                           const double lo[3], const double hi[3]) {
        tIn = 0.0; tOut = 1.0;
        for (int a = 0; a < 3; ++a) {
            const double dir = en[a] - st[a];
            if (dir == 0.0) {
                if (st[a] < lo[a] || st[a] > hi[a]) return false;
                continue;
            }
            double t1 = (lo[a] - st[a]) / dir, t2 = (hi[a] - st[a]) / dir;
            if (t1 > t2) std::swap(t1, t2);
            if (t1 > tIn) tIn = t1;
            if (t2 < tOut) tOut = t2;
        }
        return tIn <= tOut;
    }

    // Nodes.cpp:271-337 createNodesForSampling; returns whether nodes were created
    bool sampling(bool positive, u32 i, u64 &visits) {
        u32 n = 0;
        if (findReliable(0, positive, i, visits) != INVALID) return false;
        const double msd = (double)s.msd(i);
        const double st[3] = {s.px[i], s.py[i], s.pz[i]};
        const double sgn = positive ? msd : -msd;
        const double en[3] = {st[0] + (double)s.nx[i] * sgn, st[1] + (double)s.ny[i] * sgn,
                              st[2] + (double)s.nz[i] * sgn};
        for (;;) {
            if (t.isLeaf(n)) {
                t.createChildren(n);
                t.samplingNodesCreated += 8;
                if (findReliable(n, positive, i, visits) != INVALID) return true;
            }
            const u32 d = t.depth[n] + 1;
            if (t.h(d) == 0.0) throw std::runtime_error("Nodes::createNodesForSampling: Reached a tree node of size 0");
            double score = -1.0e300;
            u32 best = INVALID;
            for (u32 c = 0; c < 8; ++c) {
                const u32 ch = t.children[n] + c;
                double lo[3], hi[3];
                for (int a = 0; a < 3; ++a) { lo[a] = t.B(a, t.kc(a, ch), d); hi[a] = t.B(a, t.kc(a, ch) + 1, d); }
                double tIn, tOut;
                if (!segmentBox(tIn, tOut, st, en, lo, hi)) continue;
                const double overlap = tOut - tIn;
                if (overlap <= score) continue;
                score = overlap;
                best = c;
                if (overlap >= 1.0 - EPS_D) break;
            }
            if (best == INVALID) throw std::runtime_error("createNodesForSampling: segment leaves the node");
            n = t.children[n] + best;
        }
    }

    // Nodes.cpp:133-143 / 145-184 getNode restated on integer coordinates:
    // deepest existing node (depth <= maxDepth) containing cell (x,y,z,dq).
    u32 getNode(u64 x, u64 y, u64 z, u32 dq, u32 maxDepth, u32 &outDepth) const {
        u32 n = 0, d = 0;
        for (;;) {
            if (t.isLeaf(n) || d == maxDepth) { outDepth = d; return n; }
            const u32 sh = dq - d - 1;
            const u32 c = (u32)(((x >> sh) & 1) | (((y >> sh) & 1) << 1) | (((z >> sh) & 1) << 2));
            n = t.children[n] + c;
            ++d;
        }
    }

    // Nodes.cpp:347-395 createNodesForBalancing (one sweep, recursive DFS)
    bool balanceSweep(u32 n) {
        if (!t.isLeaf(n)) {
            bool created = false;
            const u32 c0 = t.children[n];
            for (u32 c = 0; c < 8; ++c) created |= balanceSweep(c0 + c);
            return created;
        }
        const u32 d = t.depth[n];
        const u64 lim = 1ull << d;
        bool created = false;
        for (u32 side = 0; side < 6; ++side) {
            // Nodes.cpp:60-84 getSideCenterPositionWS with stepSize = size/2
            // -> centre of the same-size face neighbour.
            int64_t k[3] = {(int64_t)t.kx[n], (int64_t)t.ky[n], (int64_t)t.kz[n]};
            k[side >> 1] += (side & 1) ? 1 : -1;
            if (k[side >> 1] < 0 || (u64)k[side >> 1] >= lim) continue;  // outside the root
            u32 nd;
            u32 nb = getNode((u64)k[0], (u64)k[1], (u64)k[2], d, d, nd);
            if (!t.isLeaf(nb)) continue;
            while (d > nd + 1) {
                t.createChildren(nb);
                nb = getNode((u64)k[0], (u64)k[1], (u64)k[2], d, d, nd);
                created = true;
            }
        }
        return created;
    }

    // Nodes.cpp:536-554 computeTargetIndices
    u32 targets(std::vector<u32> &tgt, u32 n, u32 next) const {
        if (t.isLeaf(n)) return next;
        const u32 c0 = t.children[n];
        u32 newNext = next + 8;
        for (u32 c = 0; c < 8; ++c) {
            tgt[c0 + c] = next + c;
            newNext = targets(tgt, c0 + c, newNext);
        }
        return newNext;
    }

    // Nodes.cpp:519-592 reorder
    void reorder() {
        const u32 cnt = t.count();
        std::vector<u32> tgt(cnt, INVALID);
        tgt[0] = 0;
        targets(tgt, 0, 1);
        auto remapLinks = [&](std::vector<u32> &links) {
            std::vector<u32> tmp(cnt);
            for (u32 i = 0; i < cnt; ++i) tmp[tgt[i]] = links[i] == INVALID ? INVALID : tgt[links[i]];
            links.swap(tmp);
        };
        remapLinks(t.children);
        remapLinks(t.parents);
        auto perm = [&](auto &v) {
            auto tmp = v;
            for (u32 i = 0; i < cnt; ++i) tmp[tgt[i]] = v[i];
            v.swap(tmp);
        };
        perm(t.counts); perm(t.kx); perm(t.ky); perm(t.kz); perm(t.depth);
        u32 acc = 0;
        for (u32 i = 0; i < cnt; ++i) { t.starts[i] = acc; acc += t.counts[i]; }
        std::fill(t.counts.begin(), t.counts.end(), 0u);
    }

    // Nodes.cpp:594-636 reorderSamplesAndFillNodes; within a node the slots
    // follow the original sample order (the reference's omp critical order
    // is nondeterministic).
    void reorderSamples() {
        t.sampleTarget.assign(s.n, INVALID);
        for (u32 i = 0; i < s.n; ++i) {
            u32 node = 0;
            if (!findContainer(node, i)) throw std::runtime_error("There is a sample which was not inserted into the Tree object!");
            const u32 off = t.counts[node]++;
            t.sampleTarget[i] = t.starts[node] + off;
        }
    }

    // Leaves.cpp:44-80 createScopes / gatherScopes
    void gather(u32 n) {
        if (t.isLeaf(n)) {
            t.nodeToLeaf[n] = (u32)t.leafNode.size();
            t.leafNode.push_back(n);
            return;
        }
        for (u32 c = 0; c < 8; ++c) gather(t.children[n] + c);
    }

    // DualCells.cpp:49-357 Schaefer-Warren recursion
    u32 ch(u32 n, u32 slot) const { return t.isLeaf(n) ? n : t.children[n] + slot; }
    void byNode(u32 n) {
        if (t.isLeaf(n)) return;
        u32 c[8];
        for (u32 i = 0; i < 8; ++i) c[i] = t.children[n] + i;
        for (u32 i = 0; i < 8; ++i) byNode(c[i]);
        facesYZ(c); facesXZ(c); facesXY(c);
        edgesZ(c); edgesY(c); edgesX(c);
        byVertex(c);
    }
    void by2X(u32 a, u32 b) {
        if (t.isLeaf(a) && t.isLeaf(b)) return;
        u32 c[8] = {ch(a, 1), ch(b, 0), ch(a, 3), ch(b, 2), ch(a, 5), ch(b, 4), ch(a, 7), ch(b, 6)};
        facesYZ(c); edgesZ(c); edgesY(c); byVertex(c);
    }
    void by2Y(u32 a, u32 b) {  // a = upper (y=1) node, b = lower node
        if (t.isLeaf(a) && t.isLeaf(b)) return;
        u32 c[8] = {ch(b, 2), ch(b, 3), ch(a, 0), ch(a, 1), ch(b, 6), ch(b, 7), ch(a, 4), ch(a, 5)};
        facesXZ(c); edgesZ(c); edgesX(c); byVertex(c);
    }
    void by2Z(u32 a, u32 b) {
        if (t.isLeaf(a) && t.isLeaf(b)) return;
        u32 c[8] = {ch(a, 4), ch(a, 5), ch(a, 6), ch(a, 7), ch(b, 0), ch(b, 1), ch(b, 2), ch(b, 3)};
        facesXY(c); edgesY(c); edgesX(c); byVertex(c);
    }
    void by4XY(u32 a, u32 b, u32 c_, u32 d) {
        if (t.isLeaf(a) && t.isLeaf(b) && t.isLeaf(c_) && t.isLeaf(d)) return;
        u32 c[8] = {ch(a, 3), ch(b, 2), ch(c_, 1), ch(d, 0), ch(a, 7), ch(b, 6), ch(c_, 5), ch(d, 4)};
        edgesZ(c); byVertex(c);
    }
    void by4XZ(u32 a, u32 b, u32 c_, u32 d) {
        if (t.isLeaf(a) && t.isLeaf(b) && t.isLeaf(c_) && t.isLeaf(d)) return;
        u32 c[8] = {ch(a, 5), ch(b, 4), ch(a, 7), ch(b, 6), ch(c_, 1), ch(d, 0), ch(c_, 3), ch(d, 2)};
        edgesY(c); byVertex(c);
    }
    void by4YZ(u32 a, u32 b, u32 c_, u32 d) {
        if (t.isLeaf(a) && t.isLeaf(b) && t.isLeaf(c_) && t.isLeaf(d)) return;
        u32 c[8] = {ch(a, 6), ch(a, 7), ch(b, 4), ch(b, 5), ch(c_, 2), ch(c_, 3), ch(d, 0), ch(d, 1)};
        edgesX(c); byVertex(c);
    }
    void edgesZ(const u32 c[8]) { by4XY(c[0], c[1], c[2], c[3]); by4XY(c[4], c[5], c[6], c[7]); }
    void edgesY(const u32 c[8]) { by4XZ(c[0], c[1], c[4], c[5]); by4XZ(c[2], c[3], c[6], c[7]); }
    void edgesX(const u32 c[8]) { by4YZ(c[0], c[2], c[4], c[6]); by4YZ(c[1], c[3], c[5], c[7]); }
    void facesYZ(const u32 c[8]) { by2X(c[0], c[1]); by2X(c[4], c[5]); by2X(c[2], c[3]); by2X(c[6], c[7]); }
    void facesXZ(const u32 c[8]) { by2Y(c[2], c[0]); by2Y(c[3], c[1]); by2Y(c[6], c[4]); by2Y(c[7], c[5]); }
    void facesXY(const u32 c[8]) { by2Z(c[0], c[4]); by2Z(c[1], c[5]); by2Z(c[2], c[6]); by2Z(c[3], c[7]); }
    void byVertex(const u32 n[8]) {
        bool all = true;
        for (u32 i = 0; i < 8; ++i) all &= t.isLeaf(n[i]);
        if (all) {
            for (u32 i = 0; i < 8; ++i) t.cells.push_back(t.nodeToLeaf[n[i]]);
            return;
        }
        u32 c[8];
        for (u32 i = 0; i < 8; ++i) c[i] = ch(n[i], 7 - i);
        byVertex(c);
    }

    // Tree.cpp:26-37 -> Nodes.cpp:186-194, Leaves.cpp:22-28, DualCells.cpp:21-31
    void build() {
        if (s.n == 0) throw std::runtime_error("no samples");
        computeRoot(s, t);
        t.newNode(INVALID, 0, 0, 0, 0);
        t.sampleDepth.resize(s.n);
        for (u32 i = 0; i < s.n; ++i) t.sampleDepth[i] = containerDepth(t, (double)s.msd(i));
        for (u32 i = 0; i < s.n; ++i) containment(i);
        u64 visits = 0;
        for (u32 i = 0; i < s.n; ++i) {
            bool a = sampling(true, i, visits);
            bool b = sampling(false, i, visits);
            t.samplingCreatingUnits += (u64)a + (u64)b;
        }
        while (balanceSweep(0)) {}
        reorder();
        reorderSamples();
        t.nodeToLeaf.assign(t.count(), INVALID);
        gather(0);
        byNode(0);
    }
};

// ---------------------------------------------------------------------------
// Occupancy field: Occupancy.cpp:105-597
// ---------------------------------------------------------------------------
template <typename Real>
struct Cone {
    // ObliqueCircularCone stand-in (BaseProject absent; definitions fixed here):
    //   apex a, end centre e, end-disc outside normal n, end radius R, cut length L.
    //   tComplete(x) = dot(x - a, n) / dot(e - a, n);  disc at t: centre a + (e-a) t, radius R t
    //   cut cone = t in [tStart, 1], tStart = max(0, 1 - L/|e-a|); getLength() = min(L, |e-a|)
    //   tCut = (tComplete - tStart) / (1 - tStart); degenerate iff dot(e-a, n) <= 0.
    Real a[3], v[3], n[3], R, invDenom, tStart, invCut, length, A;
    bool degenerate;
    // swept-ball culling bound, haveSeparatingAxis stand-in (conservative)
    Real vhat[3], slope;
    void init(const Real apex[3], const Real endNormal[3], const Real end[3], Real endRadius, Real coneLength) {
        Real denom = 0, len2 = 0;
        for (int i = 0; i < 3; ++i) {
            a[i] = apex[i]; n[i] = endNormal[i]; v[i] = end[i] - apex[i];
        }
        denom = v[0] * n[0] + v[1] * n[1] + v[2] * n[2];
        len2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
        A = std::sqrt(len2);
        R = endRadius;
        degenerate = !(denom > Real(0)) || !(A > Real(0)) || !(R > Real(0));
        if (degenerate) return;
        invDenom = Real(1) / denom;
        if (coneLength >= A) { tStart = 0; length = A; invCut = 1; }
        else { tStart = Real(1) - coneLength / A; length = coneLength; invCut = Real(1) / (Real(1) - tStart); }
        for (int i = 0; i < 3; ++i) vhat[i] = v[i] / A;
        slope = R < A ? R / std::sqrt(A * A - R * R) : Real(-1);
    }
    // conservative: true if the node box [lo, lo+size]^3 cannot meet the cut cone
    bool separated(const Real lo[3], Real size) const {
        const Real half = size * Real(0.5);
        Real b[3], u = 0, w2 = 0;
        for (int i = 0; i < 3; ++i) { b[i] = lo[i] + half - a[i]; u += b[i] * vhat[i]; }
        for (int i = 0; i < 3; ++i) { Real p = b[i] - u * vhat[i]; w2 += p * p; }
        const Real w = std::sqrt(w2);
        Real ts = slope < 0 ? Real(1) : (u + slope * w) / A;
        ts = std::min(Real(1), std::max(tStart, ts));
        const Real du = u - ts * A;
        const Real f = std::sqrt(du * du + w2) - R * ts;
        const Real rho = half * Real(1.7320508075688772) * Real(1.0001) + size * Real(1e-6);
        return f > rho;
    }
};

template <typename Real>
struct OccupancyT {
    const Samples &s;
    const Tree &t;
    Params prm;
    std::vector<Real> kernelSum[2], coneLength[2], priorE;   // [0] = EMPTINESS, [1] = SAMPLENESS
    std::vector<u32> states;
    Real coneCount = 0;
    std::vector<double> occupancy;
    std::vector<Real> lcx, lcy, lcz;   // leaf centres
    std::vector<Real> ncx, ncy, ncz, nsz;  // node min corner + size
    u64 hits[2] = {0, 0}, visits[2] = {0, 0};
    bool bruteForce = false;
    int threads = 1;

    OccupancyT(const Samples &s_, const Tree &t_, const Params &p) : s(s_), t(t_), prm(p) {
        const u32 L = (u32)t.leafNode.size(), N = t.count();
        for (int k = 0; k < 2; ++k) { kernelSum[k].assign(L, 0); coneLength[k].assign(L, 0); }
        priorE.assign(L, 0);
        states.assign(N, 0);
        ncx.resize(N); ncy.resize(N); ncz.resize(N); nsz.resize(N);
        for (u32 i = 0; i < N; ++i) {
            ncx[i] = (Real)t.B(0, t.kx[i], t.depth[i]); ncy[i] = (Real)t.B(1, t.ky[i], t.depth[i]);
            ncz[i] = (Real)t.B(2, t.kz[i], t.depth[i]); nsz[i] = (Real)t.size(i);
        }
        lcx.resize(L); lcy.resize(L); lcz.resize(L);
        for (u32 l = 0; l < L; ++l) {
            const u32 n = t.leafNode[l];
            lcx[l] = (Real)t.centre(0, n); lcy[l] = (Real)t.centre(1, n); lcz[l] = (Real)t.centre(2, n);
        }
    }

    // Occupancy.cpp:429-439 computeRayKernel (CONE_RAY_KERNEL_NEG_PARABLE)
    static Real rayKernel(Real z, Real Z) {
        if (z < Real(0) || z >= Z) return 0;
        const Real z2 = z * z, Z2 = Z * Z;
        return (Real(3) / (Real(2) * Z)) * (-z2 / Z2 + Real(1));
    }
    // Occupancy.cpp:392-400 computeCircularDiscKernelFromSquared (CONE_PCD_KERNEL_NEG_PARABLE)
    static Real discKernel(Real r2, Real R) {
        const Real R2 = R * R;
        if (r2 >= R2) return 0;
        const Real notNormed = Real(1) - r2 / R2;
        return (Real(2) / (Real(PI_D) * R2)) * notNormed;
    }
    // Occupancy.cpp:279-346 addKernel; returns the scaled kernel or 0 if rejected
    static Real kernelAt(const Cone<Real> &c, const Real x[3], bool samplenessType, Real confidence) {
        const Real d0 = x[0] - c.a[0], d1 = x[1] - c.a[1], d2 = x[2] - c.a[2];
        const Real tc = (d0 * c.n[0] + d1 * c.n[1] + d2 * c.n[2]) * c.invDenom;
        const Real tcut = (tc - c.tStart) * c.invCut;
        if (tcut < Real(0) || tcut >= Real(1)) return 0;
        const Real q0 = x[0] - (c.a[0] + c.v[0] * tc), q1 = x[1] - (c.a[1] + c.v[1] * tc),
                   q2 = x[2] - (c.a[2] + c.v[2] * tc);
        const Real r2 = q0 * q0 + q1 * q1 + q2 * q2;
        const Real pcdR = c.R * tc;
        if (r2 >= pcdR * pcdR) return 0;
        Real Z, z, norm;
        if (!samplenessType) { Z = c.length; z = tcut * Z; norm = 1; }
        else { const Real tc2 = std::fabs(tcut - Real(0.5)) * Real(2); Z = Real(0.5) * c.length; z = tc2 * Z; norm = Real(0.5); }
        const Real k = rayKernel(z, Z) * norm * discKernel(r2, pcdR);
        if (Real(EPS_D) > k) return 0;
        const Real sk = confidence * k;
        if (Real(EPS_D) > sk) return 0;
        return sk;
    }

    // Occupancy.cpp:491-516 forwardAnyChildStateFlag
    u32 forward(u32 n, u32 flag) {
        if (t.isLeaf(n)) return states[n];
        states[n] &= ~flag;
        for (u32 c = 0; c < 8; ++c)
            if (forward(t.children[n] + c, flag) & flag) states[n] |= flag;
        return states[n];
    }

    // Occupancy.cpp:138-248 updateKernels; with `chosen`, over the links of those samples only
    // (getCameraToSampleLinkData, Occupancy.cpp:250-277), kernels and lengths negated and no
    // flags set when `removal` (addKernel's `negative`, Occupancy.cpp:279-346)
    void pass(bool sampleness, const u32 *chosen = nullptr, u32 chosenCount = 0, bool removal = false) {
        const int type = sampleness ? 1 : 0;
        if (!sampleness) forward(0, FLAG_S);
        const u32 leafFlag = sampleness ? FLAG_S : FLAG_E;
        const u32 required = sampleness ? 0u : FLAG_S;
        const Real sign = removal ? Real(-1) : Real(1);
        const long long linkCount = (long long)(chosen ? chosenCount : s.n) * s.maxCams;
        double mLocal = 0;
        u64 hitsLocal = 0, visitsLocal = 0;
#pragma omp parallel for schedule(dynamic, 128) reduction(+ : mLocal, hitsLocal, visitsLocal) num_threads(threads) if (threads > 1)
        for (long long pos = 0; pos < linkCount; ++pos) {
            const u32 i = chosen ? chosen[pos / s.maxCams] : (u32)(pos / s.maxCams);   // Samples.h:419-437
            const long long link = (long long)i * s.maxCams + pos % s.maxCams;
            const u32 cam = s.cams[link];
            if (cam >= s.camCount) continue;          // Cameras.h:171-176 isValid
            const Real camWS[3] = {(Real)s.cx[cam], (Real)s.cy[cam], (Real)s.cz[cam]};
            const Real p[3] = {(Real)s.px[i], (Real)s.py[i], (Real)s.pz[i]};
            const Real nrm[3] = {-(Real)s.nx[i], -(Real)s.ny[i], -(Real)s.nz[i]};
            const Real R = (Real)prm.bandwidthFactor * (Real)s.scale[i];
            Real L = std::numeric_limits<float>::max();   // REAL_MAX
            Real end[3] = {p[0], p[1], p[2]};
            if (sampleness) {
                const Real cs[3] = {p[0] - camWS[0], p[1] - camWS[1], p[2] - camWS[2]};
                const Real total = std::sqrt(cs[0] * cs[0] + cs[1] * cs[1] + cs[2] * cs[2]);
                L = R * (Real)prm.relativeSampleConeLength;
                for (int a = 0; a < 3; ++a) end[a] += (cs[a] / total) * (L * Real(0.5));
            }
            Cone<Real> cone;
            cone.init(camWS, nrm, end, R, L);
            if (cone.degenerate) continue;
            const Real conf = (Real)s.conf[i];
            const Real weightedLength = sign * conf * cone.length;
            if (!sampleness) mLocal += (double)(sign * conf);
            const u32 maxDepth = t.sampleDepth[i] + MAX_DEPTH_DIFFERENCE;   // Occupancy.cpp:518-521
            // LeavesIterator over ConeNodesChecker (NodesIterator.cpp:38-86, ConeNodesChecker.cpp:50-66)
            u32 stack[8 * 24];
            int sp = 0;
            stack[sp++] = 0;
            while (sp > 0) {
                const u32 nd = stack[--sp];
                ++visitsLocal;
                const u32 d = t.depth[nd];
                if (d > maxDepth) continue;
                if ((required & states[nd]) != required) continue;
                const bool leaf = t.isLeaf(nd);
                if (!bruteForce && !leaf) {
                    const Real lo[3] = {ncx[nd], ncy[nd], ncz[nd]};
                    if (cone.separated(lo, nsz[nd])) continue;
                }
                if (!leaf) {
                    const u32 c0 = t.children[nd];
                    for (int c = 7; c >= 0; --c) stack[sp++] = c0 + (u32)c;
                    continue;
                }
                const u32 l = t.nodeToLeaf[nd];
                const Real x[3] = {lcx[l], lcy[l], lcz[l]};
                const Real k = kernelAt(cone, x, sampleness, conf);
                if (k == Real(0)) continue;
                ++hitsLocal;
                if (threads > 1) {
#pragma omp atomic
                    kernelSum[type][l] += sign * k;
#pragma omp atomic
                    coneLength[type][l] += weightedLength;
                    if (!removal) states[nd] |= leafFlag;   // benign race as in the reference (Occupancy.cpp:233)
                } else {
                    kernelSum[type][l] += sign * k;
                    coneLength[type][l] += weightedLength;
                    if (!removal) states[nd] |= leafFlag;
                }
            }
        }
        if (!sampleness) coneCount += (Real)mLocal;
        hits[type] += hitsLocal;
        visits[type] += visitsLocal;
    }

    // Occupancy.cpp:530-568 computePriors
    void priors() {
        const u32 L = (u32)t.leafNode.size();
        u64 vis = 0;
#pragma omp parallel for schedule(dynamic, 512) reduction(+ : vis) num_threads(threads) if (threads > 1)
        for (long long li = 0; li < (long long)L; ++li) {
            const u32 l = (u32)li;
            const u32 ln = t.leafNode[l];
            const Real c[3] = {lcx[l], lcy[l], lcz[l]};
            const Real radius = nsz[ln] * (Real)prm.relativeRadiusForPriors;
            const u32 maxDepth = t.depth[ln] + MAX_DEPTH_DIFFERENCE;
            Real sumE = 0, sumS = 0;
            u32 stack[8 * 24];
            int sp = 0;
            stack[sp++] = 0;
            while (sp > 0) {
                const u32 nd = stack[--sp];
                ++vis;
                if ((states[nd] & FLAG_S) != FLAG_S) continue;    // SphereNodeStatesChecker.cpp:28-36
                if (t.depth[nd] > maxDepth) continue;             // SphereNodesChecker.cpp:29-40
                Real d2 = 0;
                const Real lo[3] = {ncx[nd], ncy[nd], ncz[nd]};
                for (int a = 0; a < 3; ++a) {
                    const Real hi = lo[a] + nsz[nd];
                    if (c[a] < lo[a]) d2 += (lo[a] - c[a]) * (lo[a] - c[a]);
                    else if (c[a] > hi) d2 += (c[a] - hi) * (c[a] - hi);
                }
                if (d2 > radius * radius) continue;
                if (!t.isLeaf(nd)) {
                    const u32 c0 = t.children[nd];
                    for (int k = 7; k >= 0; --k) stack[sp++] = c0 + (u32)k;
                    continue;
                }
                const u32 j = t.nodeToLeaf[nd];
                sumE += coneLength[0][j];
                sumS += coneLength[1][j];
            }
            priorE[l] = sumE / (sumE + sumS);
        }
        visits[0] += 0;
        priorVisits = vis;
    }
    u64 priorVisits = 0;

    // Depth-capped bound of leaf l's prior (no reference counterpart: the certified priors
    // of the GPU path).  The traversal of priors() in FP64, with two additions: an inner
    // node entirely inside the sphere (farthest corner distance^2 <= r^2 (1 - 1e-9)) whose
    // deepest leaf is within the depth limit adds its subtree sums at once (the same leaves
    // priors() would sum one by one), and a straddling inner node other than the root at
    // depth >= leaf depth - capLevels (and below the depth limit) adds its subtree sums to
    // upper bounds instead of being opened.  Returns [E / (E + S + Su), (E + Eu) / (E + Eu + S)].
    void priorBound(u32 l, int capLevels, double &lo, double &hi) const {
        const u32 ln = t.leafNode[l];
        const double c[3] = {(double)lcx[l], (double)lcy[l], (double)lcz[l]};
        const double radius = (double)nsz[ln] * (double)prm.relativeRadiusForPriors;
        const double r2 = radius * radius, ri2 = r2 * (1.0 - 1e-9);
        const int md = (int)t.depth[ln] + (int)MAX_DEPTH_DIFFERENCE;
        const int capD = (int)t.depth[ln] - capLevels;
        double E = 0, S = 0, Eu = 0, Su = 0;
        std::vector<u32> stack{0u};
        while (!stack.empty()) {
            const u32 nd = stack.back();
            stack.pop_back();
            if ((states[nd] & FLAG_S) != FLAG_S) continue;
            const int d = (int)t.depth[nd];
            if (d > md) continue;
            double dmin = 0, dmax = 0;
            for (int a = 0; a < 3; ++a) {
                const double b0 = t.B(a, t.kc(a, nd), t.depth[nd]), b1 = t.B(a, t.kc(a, nd) + 1, t.depth[nd]);
                if (c[a] < b0) dmin += (b0 - c[a]) * (b0 - c[a]);
                else if (c[a] > b1) dmin += (c[a] - b1) * (c[a] - b1);
                const double far = std::max(std::fabs(c[a] - b0), std::fabs(b1 - c[a]));
                dmax += far * far;
            }
            if (!(dmin <= r2)) continue;
            if (t.isLeaf(nd) || (nd != 0 && deepest[nd] <= md && dmax <= ri2)) {
                E += subE[nd];
                S += subS[nd];
                continue;
            }
            if (d >= md) continue;
            if (nd != 0 && d >= capD) {
                Eu += subE[nd];
                Su += subS[nd];
                continue;
            }
            for (u32 k = 0; k < 8; ++k) stack.push_back(t.children[nd] + k);
        }
        lo = E / (E + S + Su);
        hi = (E + Eu) / (E + Eu + S);
    }
    // FP64 subtree sums of the cone lengths and deepest leaf depth per node (priorBound);
    // clen: per-leaf cone lengths [E | S] to use instead of this field's
    std::vector<double> subE, subS;
    std::vector<int> deepest;
    void subtreeSums(const float *clen = nullptr) {
        const u32 L = (u32)t.leafNode.size();
        const u32 N = (u32)t.children.size();
        subE.assign(N, 0.0);
        subS.assign(N, 0.0);
        deepest.assign(N, 0);
        std::vector<u32> order;   // children after parents: walk it backwards
        order.reserve(N);
        order.push_back(0);
        for (size_t i = 0; i < order.size(); ++i)
            if (!t.isLeaf(order[i]))
                for (u32 k = 0; k < 8; ++k) order.push_back(t.children[order[i]] + k);
        for (size_t i = order.size(); i-- > 0;) {
            const u32 nd = order[i];
            if (t.isLeaf(nd)) {
                const u32 j = t.nodeToLeaf[nd];
                subE[nd] = clen ? (double)clen[j] : (double)coneLength[0][j];
                subS[nd] = clen ? (double)clen[L + j] : (double)coneLength[1][j];
                deepest[nd] = (int)t.depth[nd];
            } else {
                for (u32 k = 0; k < 8; ++k) {
                    const u32 cn = t.children[nd] + k;
                    subE[nd] += subE[cn];
                    subS[nd] += subS[cn];
                    deepest[nd] = std::max(deepest[nd], deepest[cn]);
                }
            }
        }
    }

    // Occupancy.cpp:570-587 getOccupancy, evaluated in FP64
    double getOccupancy(u32 l) const {
        const double e = (double)kernelSum[0][l], sm = (double)kernelSum[1][l];
        if (e + sm < prm.confidenceThreshold) return -EPS_D;
        const double prior = (double)priorE[l];
        const double M = (double)coneCount;
        const double pE = prior * (e / M);
        const double pS = (1.0 - prior) * (sm / M);
        if (pE >= pS) return -EPS_D;
        return pS - pE;
    }

    void compute() {
        pass(true);
        pass(false);
        priors();
        const u32 L = (u32)t.leafNode.size();
        occupancy.resize(L);
        for (u32 l = 0; l < L; ++l) occupancy[l] = getOccupancy(l);
    }
};

// ---------------------------------------------------------------------------
// DMC: DualMarchingCells.cpp:41-169, DualMarchingCell.cpp:34-101,
// DualMarchingCellTables.cpp:17-314 (Paul Bourke's public-domain tables)
// ---------------------------------------------------------------------------
const u32 CORNER_REORDER[8] = {0, 1, 5, 4, 2, 3, 7, 6};
const u32 EDGE_ENDS[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {4, 5}, {5, 6}, {6, 7}, {7, 4}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};
// triangulation per cube configuration, one hex digit per edge, 3 digits per triangle
const char *const TRI[256] = {
#include "mc_cases.inc"
};

struct Mesh {
    std::vector<double> vx, vy, vz, vs;
    std::vector<u32> vkey0, vkey1;   // (min leaf, max leaf) edge-vertex key
    std::vector<u32> tri;
    u64 conflicts = 0;
};

u32 edgeTableEntry(u32 cfg) {
    u32 m = 0;
    for (u32 e = 0; e < 12; ++e)
        if (((cfg >> EDGE_ENDS[e][0]) & 1) != ((cfg >> EDGE_ENDS[e][1]) & 1)) m |= 1u << e;
    return m;
}

// DualMarchingCell.cpp:51-71 getInterpolationFactor
double interpolationFactor(double d0, double d1) {
    if (EPS_D >= std::fabs(d1 - d0)) return 0.5;
    if (EPS_D >= std::fabs(d0)) return 0.0;
    if (EPS_D >= std::fabs(d1)) return 1.0;
    return (0.0 - d0) / (d1 - d0);
}

void extractMesh(const Tree &t, const std::vector<double> &occ, Mesh &m) {
    std::map<std::pair<u32, u32>, u32> keyToVertex;
    std::map<std::pair<u32, u32>, u32> edgeUse;    // FlexibleMesh edge multiplicity
    const u32 C = (u32)(t.cells.size() / 8);
    for (u32 ci = 0; ci < C; ++ci) {
        const u32 *cell = &t.cells[8 * ci];
        double pos[8][3], dist[8], scl[8];
        u32 leaf[8];
        for (u32 i = 0; i < 8; ++i) {
            const u32 l = cell[CORNER_REORDER[i]];
            const u32 n = t.leafNode[l];
            for (int a = 0; a < 3; ++a) pos[i][a] = t.centre(a, n);
            dist[i] = occ[l];
            scl[i] = t.size(n);
            leaf[i] = l;
        }
        u32 cfg = 0;
        for (u32 i = 0; i < 8; ++i)
            if (dist[i] > 0.0) cfg |= 1u << i;
        const u32 cut = edgeTableEntry(cfg);
        if (!cut) continue;
        double ip[12][3], is[12];
        for (u32 e = 0; e < 12; ++e) {
            if (!(cut & (1u << e))) continue;
            const u32 a = EDGE_ENDS[e][0], b = EDGE_ENDS[e][1];
            const double f = interpolationFactor(dist[a], dist[b]);
            for (int k = 0; k < 3; ++k) ip[e][k] = pos[a][k] + (pos[b][k] - pos[a][k]) * f;
            is[e] = scl[a] + (scl[b] - scl[a]) * f;
        }
        const char *tri = TRI[cfg];
        const size_t ne = std::strlen(tri);
        for (size_t q = 0; q + 2 < ne; q += 3) {
            u32 te[3];
            for (int k = 0; k < 3; ++k) {
                const char ch = tri[q + k];
                te[k] = (u32)(ch <= '9' ? ch - '0' : ch - 'a' + 10);
            }
            bool degen = false;   // DualMarchingCell.cpp:73-101
            for (int k = 0; k < 3; ++k)
                if (leaf[EDGE_ENDS[te[k]][0]] == leaf[EDGE_ENDS[te[k]][1]]) degen = true;
            for (int k = 0; k < 3 && !degen; ++k) {
                const u32 *e0 = EDGE_ENDS[te[k]], *e1 = EDGE_ENDS[te[(k + 1) % 3]];
                if (leaf[e0[0]] == leaf[e1[0]] && leaf[e0[1]] == leaf[e1[1]]) degen = true;
                if (leaf[e0[0]] == leaf[e1[1]] && leaf[e0[1]] == leaf[e1[0]]) degen = true;
            }
            if (degen) continue;
            u32 vid[3];
            for (int k = 0; k < 3; ++k) {
                const u32 l0 = leaf[EDGE_ENDS[te[k]][0]], l1 = leaf[EDGE_ENDS[te[k]][1]];
                const std::pair<u32, u32> key(std::min(l0, l1), std::max(l0, l1));
                auto it = keyToVertex.find(key);
                if (it != keyToVertex.end()) { vid[k] = it->second; continue; }
                const u32 id = (u32)m.vx.size();
                m.vx.push_back(ip[te[k]][0]); m.vy.push_back(ip[te[k]][1]); m.vz.push_back(ip[te[k]][2]);
                m.vs.push_back(is[te[k]]);
                m.vkey0.push_back(key.first); m.vkey1.push_back(key.second);
                keyToVertex.emplace(key, id);
                vid[k] = id;
            }
            for (int k = 0; k < 3; ++k) m.tri.push_back(vid[k]);
            for (int k = 0; k < 3; ++k) {   // FlexibleMesh.cpp:1184-1211 addEdge
                const u32 a = vid[k], b = vid[(k + 1) % 3];
                const std::pair<u32, u32> ek(std::min(a, b), std::max(a, b));
                if (++edgeUse[ek] > 2) ++m.conflicts;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Handle + C ABI for the test harness
// ---------------------------------------------------------------------------
// Leaves.cpp:82-180 createNeighborsEdges: for every leaf (in leaf order) and side 0..5
// (Nodes.cpp:60-84: -x, +x, -y, +y, -z, +z) the deepest node, at most as deep as the leaf,
// holding the same-size neighbour cell (getUnresponsibleEdgeNeighbor).  An inner node (finer
// leaves there) or a same-depth leaf with a smaller node index is skipped; otherwise the pair
// is an edge, appended to both leaves' lists in this loop order (insertNeighborsEdge).
void leafNeighbours(const Tree &t, std::vector<u32> &offsets, std::vector<u32> &nbrs) {
    const u32 L = (u32)t.leafNode.size();
    struct E { u32 a, b; };
    std::vector<E> edges;
    for (u32 l = 0; l < L; ++l) {
        const u32 nd = t.leafNode[l];
        const u32 d = t.depth[nd];
        for (u32 side = 0; side < 6; ++side) {
            const int axis = (int)(side / 2);
            const long long step = (side & 1) ? 1 : -1;
            long long q[3] = {(long long)t.kx[nd], (long long)t.ky[nd], (long long)t.kz[nd]};
            q[axis] += step;
            if (q[axis] < 0 || q[axis] >= (1ll << d)) continue;   // outside the root: invalid node
            u32 n = 0, nd2 = 0;
            for (;;) {   // Nodes.cpp:145-184 getNode with maxDepth = d
                if (t.isLeaf(n) || nd2 == d) break;
                const u32 sh = d - nd2 - 1;
                const u32 c = (u32)(((q[0] >> sh) & 1) | (((q[1] >> sh) & 1) << 1) | (((q[2] >> sh) & 1) << 2));
                n = t.children[n] + c;
                ++nd2;
            }
            if (!t.isLeaf(n)) continue;
            if (nd2 == d && n < nd) continue;
            edges.push_back({l, t.nodeToLeaf[n]});
        }
    }
    offsets.assign(L + 1, 0);
    for (const E &e : edges) { ++offsets[e.a]; ++offsets[e.b]; }
    u32 acc = 0;
    for (u32 i = 0; i <= L; ++i) { const u32 c = offsets[i]; offsets[i] = acc; acc += c; }
    nbrs.assign(acc, INVALID);
    std::vector<u32> fill(offsets.begin(), offsets.end() - 1);
    for (const E &e : edges) { nbrs[fill[e.a]++] = e.b; nbrs[fill[e.b]++] = e.a; }
}

struct Handle {
    Samples s;
    Params prm{2, 100, 25, 1, 2};
    Tree t;
    OccupancyT<double> *occ = nullptr;
    OccupancyT<float> *occf = nullptr;
    Mesh mesh;
    std::vector<double> occValues;
    std::string err;
    std::vector<double> timings;   // seconds: tree, S, E, priors, occupancy, dmc
    std::vector<float> leafMin;    // leaf scopes as float (min xyz, size)
    std::vector<u32> leafDepth;
    std::vector<double> ksum, clen, prior;   // exported per-leaf field arrays (double)
    std::vector<u32> nbrOffsets, nbrs;       // leaf neighbour CSR (on request)
    ~Handle() { delete occ; delete occf; }
};

double now() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

extern "C" {

void *oracle_create(u32 n, const float *px, const float *py, const float *pz, const float *nx, const float *ny,
                    const float *nz, const float *scale, const float *conf, const u32 *cams, u32 maxCams,
                    u32 camCount, const float *cx, const float *cy, const float *cz, const double *params5) {
    Handle *h = new Handle();
    Samples &s = h->s;
    s.n = n; s.maxCams = maxCams; s.camCount = camCount;
    s.px.assign(px, px + n); s.py.assign(py, py + n); s.pz.assign(pz, pz + n);
    s.nx.assign(nx, nx + n); s.ny.assign(ny, ny + n); s.nz.assign(nz, nz + n);
    s.scale.assign(scale, scale + n); s.conf.assign(conf, conf + n);
    s.cams.assign(cams, cams + (size_t)n * maxCams);
    s.cx.assign(cx, cx + camCount); s.cy.assign(cy, cy + camCount); s.cz.assign(cz, cz + camCount);
    if (params5) {
        h->prm.bandwidthFactor = params5[0];
        h->prm.confidenceThreshold = params5[1];
        h->prm.relativeRadiusForPriors = params5[2];
        h->prm.relativeSampleConeLength = params5[3];
        h->prm.maxRelativeSamplingDistance = params5[4];
    }
    s.maxRel = (float)h->prm.maxRelativeSamplingDistance;
    return h;
}

void oracle_destroy(void *hp) { delete (Handle *)hp; }

const char *oracle_error(void *hp) { return ((Handle *)hp)->err.c_str(); }

int oracle_build_tree(void *hp) {
    Handle *h = (Handle *)hp;
    try {
        double t0 = now();
        Builder b(h->s, h->t);
        b.build();
        h->timings.assign(6, 0.0);
        h->timings[0] = now() - t0;
        const Tree &t = h->t;
        const u32 L = (u32)t.leafNode.size();
        h->leafMin.resize(4 * (size_t)L);
        h->leafDepth.resize(L);
        for (u32 l = 0; l < L; ++l) {
            const u32 nd = t.leafNode[l];
            for (int a = 0; a < 3; ++a) h->leafMin[4 * l + a] = (float)t.B(a, t.kc(a, nd), t.depth[nd]);
            h->leafMin[4 * l + 3] = (float)t.size(nd);
            h->leafDepth[l] = t.depth[nd];
        }
        return 0;
    } catch (const std::exception &e) {
        h->err = e.what();
        return 1;
    }
}

// mode bit0: brute force traversal (no cone culling); bit1: FP64 even with
// threads; bits 8..15: thread count (0/1 = serial FP64 parity mode; >1 =
// OpenMP float baseline mode, or OpenMP FP64 with bit1 -- the FP64 sums then
// accumulate in a thread-dependent order, ~1e-15 relative)
int oracle_occupancy(void *hp, int mode) {
    Handle *h = (Handle *)hp;
    try {
        const int threads = (mode >> 8) & 0xff;
        const bool fp64 = (mode & 2) != 0;
        const u32 L = (u32)h->t.leafNode.size();
        h->ksum.assign(2 * (size_t)L, 0); h->clen.assign(2 * (size_t)L, 0); h->prior.assign(L, 0);
        if (threads > 1 && !fp64) {
            delete h->occf;
            h->occf = new OccupancyT<float>(h->s, h->t, h->prm);
            auto &o = *h->occf;
            o.threads = threads;
            o.bruteForce = mode & 1;
            double t0 = now(); o.pass(true);
            double t1 = now(); o.pass(false);
            double t2 = now(); o.priors();
            double t3 = now();
            o.occupancy.resize(L);
            for (u32 l = 0; l < L; ++l) o.occupancy[l] = o.getOccupancy(l);
            double t4 = now();
            h->timings[1] = t1 - t0; h->timings[2] = t2 - t1; h->timings[3] = t3 - t2; h->timings[4] = t4 - t3;
            h->occValues = o.occupancy;
            for (u32 l = 0; l < L; ++l) {
                for (int k = 0; k < 2; ++k) { h->ksum[k * L + l] = o.kernelSum[k][l]; h->clen[k * L + l] = o.coneLength[k][l]; }
                h->prior[l] = o.priorE[l];
            }
            return 0;
        }
        delete h->occ;
        h->occ = new OccupancyT<double>(h->s, h->t, h->prm);
        auto &o = *h->occ;
        o.threads = threads > 1 ? threads : 1;
        o.bruteForce = mode & 1;
        double t0 = now(); o.pass(true);
        double t1 = now(); o.pass(false);
        double t2 = now(); o.priors();
        double t3 = now();
        o.occupancy.resize(L);
        for (u32 l = 0; l < L; ++l) o.occupancy[l] = o.getOccupancy(l);
        double t4 = now();
        h->timings[1] = t1 - t0; h->timings[2] = t2 - t1; h->timings[3] = t3 - t2; h->timings[4] = t4 - t3;
        h->occValues = o.occupancy;
        for (u32 l = 0; l < L; ++l) {
            for (int k = 0; k < 2; ++k) { h->ksum[k * L + l] = o.kernelSum[k][l]; h->clen[k * L + l] = o.coneLength[k][l]; }
            h->prior[l] = o.priorE[l];
        }
        return 0;
    } catch (const std::exception &e) {
        h->err = e.what();
        return 1;
    }
}

// Occupancy::addSamples / eraseSamples (Occupancy.h:167-175, 222-227) after an FP64
// oracle_occupancy: the S and E passes over the chosen samples (ORIGINAL sample indices),
// then priors and the occupancy ratio of every leaf
int oracle_update(void *hp, int removal, const u32 *chosen, u32 n) {
    Handle *h = (Handle *)hp;
    try {
        if (!h->occ) throw std::runtime_error("needs an FP64 occupancy");
        auto &o = *h->occ;
        o.pass(true, chosen, n, removal != 0);
        o.pass(false, chosen, n, removal != 0);
        o.priors();
        const u32 L = (u32)h->t.leafNode.size();
        for (u32 l = 0; l < L; ++l) o.occupancy[l] = o.getOccupancy(l);
        h->occValues = o.occupancy;
        for (u32 l = 0; l < L; ++l) {
            for (int k = 0; k < 2; ++k) { h->ksum[k * L + l] = o.kernelSum[k][l]; h->clen[k * L + l] = o.coneLength[k][l]; }
            h->prior[l] = o.priorE[l];
        }
        return 0;
    } catch (const std::exception &e) {
        h->err = e.what();
        return 1;
    }
}

// Depth-capped prior bounds (priorBound) of n leaves after an FP64 oracle_occupancy;
// clen (2 * leaf count floats, may be NULL) replaces the field's cone lengths
int oracle_prior_bounds(void *hp, const u32 *leaves, u32 n, int capLevels, const float *clen, double *lo,
                        double *hi) {
    Handle *h = (Handle *)hp;
    try {
        if (!h->occ) throw std::runtime_error("needs an FP64 occupancy");
        auto &o = *h->occ;
        o.subtreeSums(clen);
        for (u32 i = 0; i < n; ++i) o.priorBound(leaves[i], capLevels, lo[i], hi[i]);
        return 0;
    } catch (const std::exception &e) {
        h->err = e.what();
        return 1;
    }
}

// DMC on an occupancy vector (NULL = the one computed by oracle_occupancy)
int oracle_dmc(void *hp, const double *occ) {
    Handle *h = (Handle *)hp;
    try {
        double t0 = now();
        h->mesh = Mesh();
        std::vector<double> o = occ ? std::vector<double>(occ, occ + h->t.leafNode.size()) : h->occValues;
        extractMesh(h->t, o, h->mesh);
        h->timings[5] = now() - t0;
        return 0;
    } catch (const std::exception &e) {
        h->err = e.what();
        return 1;
    }
}

// generic accessors for the harness
u64 oracle_count(void *hp, const char *what) {
    Handle *h = (Handle *)hp;
    const std::string w(what);
    const Tree &t = h->t;
    if (w == "nodes") return t.count();
    if (w == "leaves") return t.leafNode.size();
    if (w == "cells") return t.cells.size() / 8;
    if (w == "samples") return h->s.n;
    if (w == "vertices") return h->mesh.vx.size();
    if (w == "triangles") return h->mesh.tri.size() / 3;
    if (w == "conflicts") return h->mesh.conflicts;
    if (w == "sampling_creating_units") return t.samplingCreatingUnits;
    if (w == "sampling_nodes_created") return t.samplingNodesCreated;
    if (w == "hits_s") return h->occ ? h->occ->hits[1] : h->occf ? h->occf->hits[1] : 0;
    if (w == "hits_e") return h->occ ? h->occ->hits[0] : h->occf ? h->occf->hits[0] : 0;
    if (w == "visits_s") return h->occ ? h->occ->visits[1] : h->occf ? h->occf->visits[1] : 0;
    if (w == "visits_e") return h->occ ? h->occ->visits[0] : h->occf ? h->occf->visits[0] : 0;
    if (w == "visits_prior") return h->occ ? h->occ->priorVisits : h->occf ? h->occf->priorVisits : 0;
    return 0;
}

const void *oracle_data(void *hp, const char *what) {
    Handle *h = (Handle *)hp;
    const std::string w(what);
    Tree &t = h->t;
    if (w == "children") return t.children.data();
    if (w == "parents") return t.parents.data();
    if (w == "sample_counts") return t.counts.data();
    if (w == "sample_starts") return t.starts.data();
    if (w == "node_to_leaf") return t.nodeToLeaf.data();
    if (w == "leaf_node") return t.leafNode.data();
    if (w == "leaf_scope") return h->leafMin.data();      // float4 per leaf
    if (w == "leaf_depth") return h->leafDepth.data();
    if (w == "cells") return t.cells.data();
    if (w == "sample_target") return t.sampleTarget.data();
    if (w == "sample_depth") return t.sampleDepth.data();
    if (w == "root_min") return t.rootMinF;                // float3
    if (w == "root_size") return &t.rootSizeF;
    if (w == "kernel_sums") return h->ksum.data();          // double [2][L], [0]=E [1]=S
    if (w == "cone_lengths") return h->clen.data();
    if (w == "priors") return h->prior.data();
    if (w == "occupancy") return h->occValues.data();
    if (w == "node_states") return h->occ ? (const void *)h->occ->states.data() : h->occf ? (const void *)h->occf->states.data() : nullptr;
    if (w == "cone_count") return h->occ ? (const void *)&h->occ->coneCount : nullptr;
    if (w == "timings") return h->timings.data();
    if (w == "vpos_x") return h->mesh.vx.data();
    if (w == "vpos_y") return h->mesh.vy.data();
    if (w == "vpos_z") return h->mesh.vz.data();
    if (w == "vscale") return h->mesh.vs.data();
    if (w == "vkey0") return h->mesh.vkey0.data();
    if (w == "vkey1") return h->mesh.vkey1.data();
    if (w == "triangles") return h->mesh.tri.data();
    if (w == "neighbour_offsets") return h->nbrOffsets.data();
    if (w == "neighbours") return h->nbrs.data();
    return nullptr;
}

// Leaves::createNeighborsEdges on the current tree; returns the edge list length
u64 oracle_leaf_neighbours(void *hp) {
    Handle *h = (Handle *)hp;
    leafNeighbours(h->t, h->nbrOffsets, h->nbrs);
    return h->nbrs.size();
}

// MC tables for the table-invariant tests
const char *oracle_mc_case(u32 cfg) { return cfg < 256 ? TRI[cfg] : ""; }
u32 oracle_mc_edges(u32 cfg) { return edgeTableEntry(cfg); }

}  // extern "C"

// ---------------------------------------------------------------------------
// Test hooks: closed-form kernels, one cone evaluation, uniform trees
// ---------------------------------------------------------------------------
extern "C" {

double oracle_ray_kernel(double z, double Z) { return OccupancyT<double>::rayKernel(z, Z); }
double oracle_disc_kernel(double r2, double R) { return OccupancyT<double>::discKernel(r2, R); }

// Cone(apex, endNormal, end, R, L) evaluated at x; out[0] = tComplete, out[1] = tCut,
// out[2] = getLength(), out[3] = degenerate flag.  Returns the scaled kernel (0 = rejected).
double oracle_cone_kernel(const double *apex, const double *endNormal, const double *end, double R, double L,
                          const double *x, int sampleness, double conf, double *out) {
    Cone<double> c;
    c.init(apex, endNormal, end, R, L);
    if (out) {
        out[3] = c.degenerate ? 1.0 : 0.0;
        if (!c.degenerate) {
            const double tc = ((x[0] - c.a[0]) * c.n[0] + (x[1] - c.a[1]) * c.n[1] + (x[2] - c.a[2]) * c.n[2]) * c.invDenom;
            out[0] = tc;
            out[1] = (tc - c.tStart) * c.invCut;
            out[2] = c.length;
        }
    }
    if (c.degenerate) return 0.0;
    return OccupancyT<double>::kernelAt(c, x, sampleness != 0, conf);
}

// conservative culling bound of the cone for the cube [lo, lo + size]^3 (1 = culled)
int oracle_cone_separated(const double *apex, const double *endNormal, const double *end, double R, double L,
                          const double *lo, double size) {
    Cone<double> c;
    c.init(apex, endNormal, end, R, L);
    if (c.degenerate) return 1;
    return c.separated(lo, size) ? 1 : 0;
}

// Replace the handle's tree by a uniform tree of the given depth over the unit
// cube [0,1)^3 (leaves in Morton order) and build leaves + dual cells; used by
// the closed-form DMC tests.
int oracle_uniform_tree(void *hp, u32 depth) {
    Handle *h = (Handle *)hp;
    try {
        h->t = Tree();
        Tree &t = h->t;
        t.rootSizeF = 1.0f; t.rootSize = 1.0;
        t.newNode(INVALID, 0, 0, 0, 0);
        std::vector<u32> frontier{0};
        for (u32 d = 0; d < depth; ++d) {
            std::vector<u32> next;
            for (u32 nd : frontier) {
                t.createChildren(nd);
                for (u32 c = 0; c < 8; ++c) next.push_back(t.children[nd] + c);
            }
            frontier.swap(next);
        }
        Builder b(h->s, t);
        b.reorder();
        t.nodeToLeaf.assign(t.count(), INVALID);
        b.gather(0);
        b.byNode(0);
        h->timings.assign(6, 0.0);
        return 0;
    } catch (const std::exception &e) {
        h->err = e.what();
        return 1;
    }
}

// Replace the handle's tree by the tree over the unit cube whose inner nodes are the given
// location codes ((1 << 3d) | Morton code, parents listed before children; the root, code
// 1, is implied) and build leaves + dual cells; used by the dual-cell order known answers.
int oracle_split_tree(void *hp, const u64 *codes, u32 n) {
    Handle *h = (Handle *)hp;
    try {
        h->t = Tree();
        Tree &t = h->t;
        t.rootSizeF = 1.0f; t.rootSize = 1.0;
        t.newNode(INVALID, 0, 0, 0, 0);
        t.createChildren(0);
        for (u32 i = 0; i < n; ++i) {
            const u64 code = codes[i];
            u32 d = 0;
            while ((code >> (3 * (d + 1))) != 0) ++d;
            u32 nd = 0;
            for (u32 lvl = 0; lvl < d; ++lvl) {
                if (t.isLeaf(nd)) throw std::runtime_error("split code below a leaf");
                nd = t.children[nd] + (u32)((code >> (3 * (d - 1 - lvl))) & 7);
            }
            if (t.isLeaf(nd)) t.createChildren(nd);
        }
        Builder b(h->s, t);
        b.reorder();
        t.nodeToLeaf.assign(t.count(), INVALID);
        b.gather(0);
        b.byNode(0);
        h->timings.assign(6, 0.0);
        return 0;
    } catch (const std::exception &e) {
        h->err = e.what();
        return 1;
    }
}

// leaf centre of leaf l (3 doubles)
void oracle_leaf_centre(void *hp, u32 l, double *out) {
    Handle *h = (Handle *)hp;
    const u32 n = h->t.leafNode[l];
    for (int a = 0; a < 3; ++a) out[a] = h->t.centre(a, n);
}

}  // extern "C"
