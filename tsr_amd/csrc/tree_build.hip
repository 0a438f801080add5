// GPU scene-octree build (Tree::Tree, SR/Scene/Tree/Tree.cpp:26-37).
//
// Stages (reference -> here):
//   root scope        Tree.cpp:70-101, Samples.cpp:290-298   block min/max reduction (float, exact)
//   containment       Nodes.cpp:207-211,222-269              container location codes + ancestor
//                                                             closure into a GPU hash set
//   sampling refine   Nodes.cpp:214-219,271-337,433-517      order-dependent: speculative GPU
//                                                             evaluation of every (sample, side) unit
//                                                             on the containment tree; exact in-order
//                                                             replay of the few creating / disturbed
//                                                             units on the host (see replaySampling)
//   balancing         Nodes.cpp:339-395                       level-free fixpoint of forced splits
//   reorder           Nodes.cpp:519-592                       radix sort by pre-order DFS key
//   sample reorder    Nodes.cpp:594-636, Samples.cpp:825-835 stable radix sort by container node id
//   leaves            Leaves.cpp:44-80                        leaves sorted by DFS key
//   dual cells        DualCells.cpp:49-357                    one cell per interior leaf corner, ordered
//                                                             by the Schaefer-Warren recursion path key
#include <algorithm>
#include <chrono>
#include <functional>
#include <atomic>
#include <future>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <queue>
#include <thread>
#include <condition_variable>
#include <mutex>
#include <vector>

#include "handles.h"

namespace tsr {

namespace {

constexpr u32 BLOCK = 256;
constexpr u32 VISIT_CAP = 48;   // visited-leaf records per sampling unit before it counts as an event
// transferred part of the visit index, in max sampling distances: a refinement subdivides
// only nodes that meet its unit's segment [p, p +- n msd], so the T0 leaves holding them lie
// within msd of p (5% more covers float rounding of p, n and msd); leaves outside are
// looked up exactly on the device when a cascade reaches them.  TSR_HOT_RADIUS overrides
// the radius (read at every build; a smaller one sends more lookups to the device path,
// with the same tree).
double hotRadius() {
    const char *e = std::getenv("TSR_HOT_RADIUS");
    return e ? std::atof(e) : 1.05;
}

// error bits written by kernels
enum : u32 { ERR_OUTSIDE = 1, ERR_TOO_DEEP = 2, ERR_HASH_FULL = 4, ERR_NOT_INSERTED = 8 };

// ---------------------------------------------------------------------------
// root scope
// ---------------------------------------------------------------------------
__global__ void k_aabb(const float *__restrict__ px, const float *__restrict__ py, const float *__restrict__ pz,
                       const float *__restrict__ scale, float maxRel, u32 n, float *__restrict__ partial) {
    __shared__ float s[6][BLOCK];
    float mn[3] = {3.4028235e38f, 3.4028235e38f, 3.4028235e38f};
    float mx[3] = {-3.4028235e38f, -3.4028235e38f, -3.4028235e38f};
    for (u32 i = blockIdx.x * BLOCK + threadIdx.x; i < n; i += gridDim.x * BLOCK) {
        const float msd = scale[i] * maxRel;   // Samples.h:409-412
        const float r = msd + EPS_F;           // Samples.cpp:445
        const float p[3] = {px[i], py[i], pz[i]};
        for (int a = 0; a < 3; ++a) {
            mn[a] = fminf(mn[a], p[a] - r);
            mx[a] = fmaxf(mx[a], p[a] + r);
        }
    }
    for (int a = 0; a < 3; ++a) { s[a][threadIdx.x] = mn[a]; s[3 + a][threadIdx.x] = mx[a]; }
    __syncthreads();
    for (u32 w = BLOCK / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w)
            for (int a = 0; a < 3; ++a) {
                s[a][threadIdx.x] = fminf(s[a][threadIdx.x], s[a][threadIdx.x + w]);
                s[3 + a][threadIdx.x] = fmaxf(s[3 + a][threadIdx.x], s[3 + a][threadIdx.x + w]);
            }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int a = 0; a < 6; ++a) partial[6 * blockIdx.x + a] = s[a][0];
}

__global__ void k_checkScales(const float *__restrict__ scale, u32 n, u32 *err) {
    const u32 i = blockIdx.x * BLOCK + threadIdx.x;
    if (i < n && !(scale[i] > 0.0f)) atomicOr(err, 16u);
}

// ---------------------------------------------------------------------------
// containment (Nodes.cpp:222-269): container = node at the sample depth
// containing the sample; its ancestors become inner nodes.
// ---------------------------------------------------------------------------
__global__ void k_containers(Frame f, const float *__restrict__ px, const float *__restrict__ py,
                             const float *__restrict__ pz, const float *__restrict__ scale, float maxRel, u32 n,
                             u64 *__restrict__ containerLoc, uint8_t *__restrict__ depthOut, u32 *err) {
    const u32 i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    const double p[3] = {px[i], py[i], pz[i]};
    for (int a = 0; a < 3; ++a)   // Nodes.cpp:237-238 (tolerance EPSILON)
        if (!(p[a] >= f.B(a, 0, 0) - EPS_D && p[a] < f.B(a, 1, 0) + EPS_D)) atomicOr(err, (u32)ERR_OUTSIDE);
    const double msd = (double)(scale[i] * maxRel);
    const u32 d = containerDepth(f, msd);
    if (d > MAX_DEPTH) {
        atomicOr(err, (u32)ERR_TOO_DEEP);
        containerLoc[i] = 1;
        depthOut[i] = 0;
        return;
    }
    containerLoc[i] = makeLoc(cellCoord(f, 0, p[0], d), cellCoord(f, 1, p[1], d), cellCoord(f, 2, p[2], d), d);
    depthOut[i] = (uint8_t)d;
}

// Inner-node set of the containment tree with, as each node's value, the mask of its
// inner children (k_zeroVals, then k_childMasks over the set's keys).  A DFS that tests the
// 8 children of an open node then looks the set up once per node it opens, through a
// one-entry cache of the last parent's mask, instead of once per child.
__global__ void k_zeroVals(HashView h, const u64 *__restrict__ keys, u32 n) {
    const u32 i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    if (u32 *v = h.slot(keys[i])) *v = 0u;
}
__global__ void k_childMasks(HashView h, const u64 *__restrict__ keys, u32 n) {
    const u32 i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n || keys[i] <= 1ull) return;
    if (u32 *v = h.slot(keys[i] >> 3)) atomicOr(v, 1u << (keys[i] & 7));   // the parent is inner too
}
struct InnerMasks {
    HashView h;
    u64 cacheKey = 0;
    u32 cacheMask = 0;
    __device__ bool operator()(u64 loc) {
        if (loc == 1ull) return h.contains(1ull);
        const u64 parent = loc >> 3;
        if (parent != cacheKey) {
            const u32 *v = h.slot(parent);
            cacheKey = parent;
            cacheMask = v ? *v : 0u;
        }
        return (cacheMask >> (loc & 7)) & 1u;
    }
};

__global__ void k_insertAncestors(HashView h, const u64 *__restrict__ containerLoc, u32 n, u32 *err) {
    const u32 i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    u64 loc = containerLoc[i];
    while (loc > 1) {
        loc >>= 3;
        if (!h.insert(loc, 0, err)) break;
    }
}

// ---------------------------------------------------------------------------
// sampling refinement, speculative evaluation on a frozen tree
// Nodes.cpp:433-469 findReliableSamplingNode as a stackless DFS over location
// codes (parent = loc >> 3, next sibling = loc + 1).
// ---------------------------------------------------------------------------
// A leaf is reliable when its centre fits and so does one of its 8 children's
// (fittingCentre at depth d + 1, with the two centre offsets per axis, their squares and
// plane terms formed once: the same operations in the same order).
TSR_HD bool reliableLeafAtK(const Frame &f, const u32 k[3], u32 d, bool positive, const SamplingUnit &u) {
    if (!fittingCentre(f, k[0], k[1], k[2], d, positive, u)) return false;
    double sq[3][2], pl[3][2];
    for (int a = 0; a < 3; ++a)
        for (u32 b = 0; b < 2; ++b) {
            const double dx = f.B(a, 2 * (2 * k[a] + b) + 1, d + 2) - u.p[a];
            sq[a][b] = dx * dx;
            pl[a][b] = dx * u.n[a];
        }
    const double msd2 = u.msd * u.msd;
    for (u32 c = 0; c < 8; ++c) {
        const u32 bx = c & 1, by = (c >> 1) & 1, bz = c >> 2;
        if ((sq[0][bx] + sq[1][by]) + sq[2][bz] > msd2) continue;
        if (((pl[0][bx] + pl[1][by]) + pl[2][bz] > 0.0) == positive) return true;
    }
    return false;
}

// Nodes.cpp:490-517 intersect, per child box: some corner on the wanted side of the sample
// plane (its plane distance ((x - px) nx + (y - py) ny) + (z - pz) nz, with x, y, z the
// box's planes, > 0 for the positive side) and box/ball overlap (the squared distances to
// the box's slabs, summed over the axes in order, <= msd^2).
// The terms an opened node's 8 children share in that test: per axis and child bit b, over
// the child's two boundary planes, the extreme of (plane - p) n (the largest for the
// positive side, the smallest for the negative one), and the squared distance from p to
// the child's slab.  FP64 rounding is monotone, so the extremes added in the corner sums'
// order give exactly the extreme of the eight computed corner sums: some corner is on the
// wanted side iff that sum is.  The distance sum adds the same terms in the same order.
struct OpenTerms {
    double ext[3][2], dist[3][2];
};
TSR_HD void openTerms(const Frame &f, const u32 k[3], u32 d, bool positive, const SamplingUnit &u, OpenTerms &o) {
    for (int a = 0; a < 3; ++a) {
        double pl[3], t[3];
        for (u32 j = 0; j < 3; ++j) {
            pl[j] = f.B(a, 2 * k[a] + j, d + 1);
            t[j] = (pl[j] - u.p[a]) * u.n[a];
        }
        for (u32 b = 0; b < 2; ++b) {
            o.ext[a][b] = positive ? fmax(t[b], t[b + 1]) : fmin(t[b], t[b + 1]);
            const double lo = pl[b], hi = pl[b + 1];
            o.dist[a][b] = u.p[a] < lo ? (lo - u.p[a]) * (lo - u.p[a])
                                       : (u.p[a] > hi ? (u.p[a] - hi) * (u.p[a] - hi) : 0.0);
        }
    }
}
// the test for child c of the opened node
// (selects, not indexing by the runtime slot: that would put the terms in scratch memory)
TSR_HD bool intersectChild(const OpenTerms &o, u32 c, bool positive, double msd2) {
    const bool bx = c & 1u, by = (c >> 1) & 1u, bz = (c >> 2) & 1u;
    const double v = ((bx ? o.ext[0][1] : o.ext[0][0]) + (by ? o.ext[1][1] : o.ext[1][0])) +
                     (bz ? o.ext[2][1] : o.ext[2][0]);
    if ((v > 0.0) != positive) return false;
    const double d2 = ((bx ? o.dist[0][1] : o.dist[0][0]) + (by ? o.dist[1][1] : o.dist[1][0])) +
                      (bz ? o.dist[2][1] : o.dist[2][0]);
    return d2 <= msd2;
}

// Depth-first search of findReliableSamplingNode from `start` (Nodes.cpp:433-469): children
// in slot order, inner ones opened as met, each leaf met reported to onLeaf; true at the
// first reliable leaf.  The opened node is kept as grid coordinates and depth, with its
// children's shared terms formed on opening (and again on returning to it).
template <typename IsInner, typename OnLeaf>
TSR_HD bool findReliable(const Frame &f, u64 start, bool positive, const SamplingUnit &u, IsInner isInner,
                         OnLeaf onLeaf) {
    u32 k[3];
    locCoords(start, k[0], k[1], k[2]);
    u32 d = locDepth(start);
    if (!isInner(start)) {
        onLeaf(start);
        return reliableLeafAtK(f, k, d, positive, u);
    }
    const double msd2 = u.msd * u.msd;
    u64 loc = start;
    OpenTerms o;
    openTerms(f, k, d, positive, u, o);
    u32 next = 0;
    for (;;) {
        if (next == 8) {
            if (loc == start) return false;
            next = (u32)(loc & 7) + 1;
            loc >>= 3;
            for (int a = 0; a < 3; ++a) k[a] >>= 1;
            --d;
            openTerms(f, k, d, positive, u, o);
            continue;
        }
        const u32 c = next;
        if (intersectChild(o, c, positive, msd2)) {
            const u64 ch = (loc << 3) | c;
            const u32 kc[3] = {2 * k[0] + (c & 1u), 2 * k[1] + ((c >> 1) & 1u), 2 * k[2] + (c >> 2)};
            if (isInner(ch)) {
                loc = ch;
                for (int a = 0; a < 3; ++a) k[a] = kc[a];
                ++d;
                openTerms(f, k, d, positive, u, o);
                next = 0;
                continue;
            }
            onLeaf(ch);
            if (reliableLeafAtK(f, kc, d + 1, positive, u)) return true;
        }
        ++next;
    }
}

// Host form of findReliable for the in-order replay: the same depth-first search and the
// same predicates, with each open node's shared child terms kept on a stack of open nodes
// (no recomputation on returning to a node).
template <typename IsInner, typename OnLeaf>
bool findReliableHost(const Frame &f, u64 start, bool positive, const SamplingUnit &u, IsInner isInner,
                      OnLeaf onLeaf) {
    struct Open {
        u64 loc;
        u32 k[3], d;
        OpenTerms o;
    };
    Open stack[MAX_DEPTH + 2];
    int top = 0;
    {
        Open &s0 = stack[0];
        s0.loc = start;
        locCoords(start, s0.k[0], s0.k[1], s0.k[2]);
        s0.d = locDepth(start);
        if (!isInner(start)) {
            onLeaf(start);
            return reliableLeafAtK(f, s0.k, s0.d, positive, u);
        }
        openTerms(f, s0.k, s0.d, positive, u, s0.o);
        top = 1;
    }
    const double msd2 = u.msd * u.msd;
    u32 next = 0;
    for (;;) {
        const Open &o = stack[top - 1];
        if (next == 8) {
            if (o.loc == start) return false;
            next = (u32)(o.loc & 7) + 1;
            --top;
            continue;
        }
        const u32 c = next;
        if (intersectChild(o.o, c, positive, msd2)) {
            const u64 ch = (o.loc << 3) | c;
            const u32 kc[3] = {2 * o.k[0] + (c & 1u), 2 * o.k[1] + ((c >> 1) & 1u), 2 * o.k[2] + (c >> 2)};
            if (isInner(ch)) {
                Open &n = stack[top++];
                n.loc = ch;
                n.k[0] = kc[0]; n.k[1] = kc[1]; n.k[2] = kc[2];
                n.d = o.d + 1;
                openTerms(f, n.k, n.d, positive, u, n.o);
                next = 0;
                continue;
            }
            onLeaf(ch);
            if (reliableLeafAtK(f, kc, o.d + 1, positive, u)) return true;
        }
        ++next;
    }
}

// Where the DFS of findReliable can start without changing its result: the
// deepest existing node whose box strictly contains the unit's search ball
// (centre p, radius msd) -- every node that the intersect test accepts lies inside
// it, all its ancestors pass that test (the sample plane crosses their
// interior), and their other children are rejected, so a DFS from the root
// reaches this node first and yields the same leaves in the same order.  The
// ball is inflated far above FP64 rounding before the strict-containment test.
// Returns the root when the ball is not strictly inside the root box.
// Deepest node on the ancestor chain of the box [a0, a1] that strictly contains it:
// the deepest inner one, or its child on the chain when that child is a leaf
// (innerOnly = false).  Returns the root when the box is not strictly inside the root.
template <typename IsInner>
TSR_HD u64 deepestContaining(const Frame &f, const double a0[3], const double a1[3], bool innerOnly, IsInner isInner) {
    u32 lo[3], diff = 0;
    for (int a = 0; a < 3; ++a) {
        if (!(a0[a] > f.rootMin[a]) || !(a1[a] < f.B(a, 1ull << MAX_DEPTH, MAX_DEPTH))) return 1ull;
        // cellCoord: B(k) <= x < B(k + 1), so a0 > B(lo) needs a0 != B(lo)
        lo[a] = cellCoord(f, a, a0[a], MAX_DEPTH);
        if (f.B(a, lo[a], MAX_DEPTH) == a0[a]) return 1ull;
        diff |= lo[a] ^ cellCoord(f, a, a1[a], MAX_DEPTH);
    }
    const u32 dc = diff ? MAX_DEPTH - (32u - (u32)__builtin_clz(diff)) : MAX_DEPTH;
    auto anc = [&](u32 d) { return makeLoc(lo[0] >> (MAX_DEPTH - d), lo[1] >> (MAX_DEPTH - d), lo[2] >> (MAX_DEPTH - d), d); };
    if (dc == 0 || !isInner(1ull)) return 1ull;
    // inner-ness is prefix-closed along the ancestor chain: binary search the deepest inner one
    u32 good = 0, bad = dc + 1;
    while (bad - good > 1) {
        const u32 mid = (good + bad) / 2;
        if (isInner(anc(mid))) good = mid;
        else bad = mid;
    }
    return (good == dc || innerOnly) ? anc(good) : anc(good + 1);
}

// Where the DFS of findReliable can start without changing its result: the
// deepest existing node whose box strictly contains the unit's search ball
// (centre p, radius msd) -- every node that the intersect test accepts lies inside
// it, all its ancestors pass that test (the sample plane crosses their
// interior), and their other children are rejected, so a DFS from the root
// reaches this node first and yields the same leaves in the same order.  The
// ball is inflated far above FP64 rounding before the strict-containment test.
template <typename IsInner>
TSR_HD u64 searchStart(const Frame &f, const SamplingUnit &u, IsInner isInner) {
    const double r = u.msd * (1.0 + 1e-9) + 1e-12 * f.rootSize;
    double a0[3], a1[3];
    for (int a = 0; a < 3; ++a) {
        a0[a] = u.p[a] - r;
        a1[a] = u.p[a] + r;
    }
    return deepestContaining(f, a0, a1, false, isInner);
}

__device__ SamplingUnit loadUnit(const float *px, const float *py, const float *pz, const float *nx, const float *ny,
                                 const float *nz, const float *scale, float maxRel, u32 i) {
    SamplingUnit u;
    u.p[0] = px[i]; u.p[1] = py[i]; u.p[2] = pz[i];
    u.n[0] = nx[i]; u.n[1] = ny[i]; u.n[2] = nz[i];
    u.msd = (double)(scale[i] * maxRel);
    return u;
}

// mode 0: count visited leaves and flag creating units; mode 1: write (leaf, unit) pairs;
// mode 2: as mode 0, and stage the first VISIT_CAP visited leaves in the unit's slots
// (visitSlot: groups of 256 units own contiguous rows, row j = j-th visit)
__device__ __forceinline__ size_t visitSlot(u32 unit, u32 j) {
    return ((size_t)(unit / 256) * VISIT_CAP + j) * 256 + (unit % 256);
}
__global__ void k_samplingEval(Frame f, HashView inner, const float *__restrict__ px, const float *__restrict__ py,
                               const float *__restrict__ pz, const float *__restrict__ nx,
                               const float *__restrict__ ny, const float *__restrict__ nz,
                               const float *__restrict__ scale, float maxRel, u32 units, int mode,
                               u32 *__restrict__ visitCount, uint8_t *__restrict__ isEvent,
                               const u32 *__restrict__ offsets, u64 *__restrict__ pairLeaf,
                               u32 *__restrict__ pairUnit) {
    const u32 unit = blockIdx.x * BLOCK + threadIdx.x;
    if (unit >= units) return;
    if (mode == 1 && isEvent[unit]) return;
    const u32 i = unit >> 1;
    const bool positive = (unit & 1) == 0;   // Nodes.cpp:217-218: positive side first
    const SamplingUnit u = loadUnit(px, py, pz, nx, ny, nz, scale, maxRel, i);
    u32 count = 0;
    const u32 base = mode == 1 ? offsets[unit] : 0;
    InnerMasks masks{inner};
    auto isInner = [&](u64 loc) { return masks(loc); };
    auto onLeaf = [&](u64 loc) {
        if (mode == 1 && count < VISIT_CAP) {
            pairLeaf[base + count] = loc;
            pairUnit[base + count] = unit;
        } else if (mode == 2 && count < VISIT_CAP) {
            pairLeaf[visitSlot(unit, count)] = loc;
        }
        ++count;
    };
    const bool found = findReliable(f, searchStart(f, u, isInner), positive, u, isInner, onLeaf);
    if (mode == 0 || mode == 2) {
        const bool event = !found || count > VISIT_CAP;
        isEvent[unit] = event ? 1 : 0;
        visitCount[unit] = event ? 0 : count;
    }
}

// the occupied keys of a hash table, in slot order (a stream compaction: no counter
// atomics, deterministic order); their number goes to *count
struct NonZeroKey {
    __host__ __device__ bool operator()(const u64 &k) const { return k != 0ull; }
};
void compactKeys(Scratch &sc, const DHash &h, u64 *out, u32 *count, hipStream_t st) {
    selectIf(sc, h.keys.p, out, count, h.capacity, NonZeroKey(), st);
}

__global__ void k_insertList(HashView h, const u64 *__restrict__ locs, u32 n, u32 *err) {
    const u32 i = blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) h.insert(locs[i], 0, err);
}

// ---------------------------------------------------------------------------
// 2:1 face balancing (Nodes.cpp:339-395): for every leaf (child of an inner
// node that is not inner) and each face, the neighbour leaf must be at most one
// level coarser; otherwise the neighbour's ancestors down to depth d-2 are split.
// ---------------------------------------------------------------------------
// One pass over a list of inner nodes: their leaf children are checked, and the nodes
// made inner go to `added` (the next pass needs only those: a split can make new
// leaves too deep for a neighbour region, never an existing leaf).
__global__ void k_balance(HashView inner, const u64 *__restrict__ parents, u32 n, u64 *__restrict__ added,
                          u32 *addedCount, u32 addedCap, u32 *err) {
    const u32 s = blockIdx.x * BLOCK + threadIdx.x;
    if (s >= n) return;
    const u64 parent = parents[s];
    const u32 d = locDepth(parent) + 1;
    if (d < 2) return;
    for (u32 c = 0; c < 8; ++c) {
        const u64 leaf = (parent << 3) | c;
        if (inner.contains(leaf)) continue;
        u32 k[3];
        locCoords(leaf, k[0], k[1], k[2]);
        const u32 lim = 1u << d;
        for (u32 side = 0; side < 6; ++side) {
            u32 q[3] = {k[0], k[1], k[2]};
            const u32 a = side >> 1;
            if (side & 1) {
                if (q[a] + 1 >= lim) continue;
                q[a] += 1;
            } else {
                if (q[a] == 0) continue;
                q[a] -= 1;
            }
            const u64 nb = makeLoc(q[0], q[1], q[2], d);
            if (inner.contains(nb >> 3)) continue;   // neighbour exists at depth d
            // deepest inner ancestor of nb, depth m < d - 1
            u32 m = d - 2;
            while (m > 0 && !inner.contains(nb >> (3 * (d - m)))) --m;
            if (m + 1 >= d - 1) continue;            // neighbour leaf at depth m+1 >= d-1
            for (u32 j = m + 1; j <= d - 2; ++j) {
                const u64 split = nb >> (3 * (d - j));
                if (inner.insert(split, 0, err)) {
                    const u32 at = atomicAdd(addedCount, 1u);
                    if (at < addedCap) added[at] = split;
                    else atomicOr(err, (u32)ERR_HASH_FULL);
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// finalisation: node ids in DFS block order, leaves, samples, dual cells
// ---------------------------------------------------------------------------
__global__ void k_dfsKeys(const u64 *__restrict__ locs, u32 n, u64 *__restrict__ keys) {
    const u32 i = blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) keys[i] = dfsKey(locs[i]);
}

// children of the inner node with pre-order rank r get ids 1 + 8r + c (Nodes.cpp:536-554)
// each inner node's value: its rank in DFS block order, so that the id of a node is
// 1 + 8 rank(parent) + slot (the root: 0) with one lookup in the (smaller) inner set
__global__ void k_innerRanks(HashView inner, const u64 *__restrict__ innerSorted, u32 innerCount) {
    const u32 r = blockIdx.x * BLOCK + threadIdx.x;
    if (r >= innerCount) return;
    if (u32 *v = inner.slot(innerSorted[r])) *v = r;
}
__device__ __forceinline__ u32 nodeIdByRank(HashView innerRank, u64 loc) {
    if (loc == 1ull) return 0u;
    const u32 r = innerRank.find(loc >> 3);
    return r == INVALID ? INVALID : 1u + 8u * r + (u32)(loc & 7);
}

__global__ void k_nodeArrays(HashView inner, const u64 *__restrict__ innerSorted, u32 innerCount,
                             u32 *__restrict__ child, u32 *__restrict__ parents, uint8_t *__restrict__ depth,
                             u64 *__restrict__ nodeLoc, u64 *__restrict__ leafKeys, u32 *__restrict__ leafIds,
                             u32 *leafCounter) {
    const u32 r = blockIdx.x * BLOCK + threadIdx.x;
    if (r == 0) {
        parents[0] = INVALID;
        depth[0] = 0;
        nodeLoc[0] = 1ull;
        if (innerCount == 0) {   // single-leaf tree
            child[0] = INVALID;
            const u32 slot = atomicAdd(leafCounter, 1u);
            leafKeys[slot] = dfsKey(1ull);
            leafIds[slot] = 0;
        }
    }
    if (r >= innerCount) return;
    const u64 loc = innerSorted[r];
    const u32 id = nodeIdByRank(inner, loc);
    child[id] = 1u + 8u * r;
    const u32 d = locDepth(loc) + 1;
    for (u32 c = 0; c < 8; ++c) {
        const u64 cl = (loc << 3) | c;
        const u32 cid = 1u + 8u * r + c;
        parents[cid] = id;
        depth[cid] = (uint8_t)d;
        nodeLoc[cid] = cl;
        if (!inner.contains(cl)) {
            const u32 slot = atomicAdd(leafCounter, 1u);
            leafKeys[slot] = dfsKey(cl);
            leafIds[slot] = cid;
        }
    }
}

__global__ void k_leafArrays(const u32 *__restrict__ leafIdsSorted, u32 leafCount, const u64 *__restrict__ nodeLoc,
                             u32 *__restrict__ child, u32 *__restrict__ nodeToLeaf, u64 *__restrict__ leafLoc,
                             u32 *__restrict__ leafNode) {
    const u32 l = blockIdx.x * BLOCK + threadIdx.x;
    if (l >= leafCount) return;
    const u32 id = leafIdsSorted[l];
    child[id] = LEAF_BIT | l;
    nodeToLeaf[id] = l;
    leafLoc[l] = nodeLoc[id];
    leafNode[l] = id;
}

__global__ void k_fillU32(u32 *p, size_t n, u32 v) {
    const size_t i = (size_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_sampleNodes(HashView innerRank, const u64 *__restrict__ containerLoc, u32 n,
                              u32 *__restrict__ nodeOf, u32 *__restrict__ counts, u32 *err) {
    const u32 i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    const u32 id = nodeIdByRank(innerRank, containerLoc[i]);
    if (id == INVALID) {
        atomicOr(err, (u32)ERR_NOT_INSERTED);
        nodeOf[i] = 0;
        return;
    }
    nodeOf[i] = id;
    atomicAdd(&counts[id], 1u);
}

__global__ void k_iota(u32 *p, u32 n) {
    const u32 i = blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) p[i] = i;
}

__global__ void k_scatterTarget(const u32 *__restrict__ sortedOrig, u32 n, u32 *__restrict__ target) {
    const u32 j = blockIdx.x * BLOCK + threadIdx.x;
    if (j < n) target[sortedOrig[j]] = j;
}

template <typename T>
__global__ void k_gather(const T *__restrict__ src, const u32 *__restrict__ order, u32 n, u32 stride,
                         T *__restrict__ dst) {
    const u32 j = blockIdx.x * BLOCK + threadIdx.x;
    if (j >= n) return;
    const u32 o = order[j];
    for (u32 s = 0; s < stride; ++s) dst[(size_t)j * stride + s] = src[(size_t)o * stride + s];
}

// dual cells: every interior corner of every leaf, packed (x | y << 21 | z << 42) at MAX_DEPTH
// leaf corners as vertex codes on the grid of the tree's deepest level D:
// x | y << D | z << 2D (3D bits), boundary corners ~0 (bit 3D set: they sort last)
__global__ void k_cornerCodes(const u64 *__restrict__ leafLoc, u32 leafCount, u32 D, u64 *__restrict__ codes) {
    const u32 l = blockIdx.x * BLOCK + threadIdx.x;
    if (l >= leafCount) return;
    const u64 loc = leafLoc[l];
    u32 k[3];
    locCoords(loc, k[0], k[1], k[2]);
    const u32 sh = D - locDepth(loc);
    const u32 lim = 1u << D;
    for (u32 c = 0; c < 8; ++c) {
        const u32 x = (k[0] + (c & 1)) << sh, y = (k[1] + ((c >> 1) & 1)) << sh, z = (k[2] + ((c >> 2) & 1)) << sh;
        const bool interior = x > 0 && x < lim && y > 0 && y < lim && z > 0 && z < lim;
        codes[8 * (size_t)l + c] = interior ? ((u64)x | ((u64)y << D) | ((u64)z << (2 * D))) : ~0ull;
    }
}

// Schaefer-Warren recursion path of the cell at vertex v (DualCells.cpp:49-357):
// call types 0 node, 1-3 face normal x/y/z, 4-6 edge along x/y/z; 5 bits per level.
TSR_HD void swKey(u32 vx, u32 vy, u32 vz, u64 &hi, u64 &lo) {
    const u32 v[3] = {vx, vy, vz};
    u32 k[3] = {0, 0, 0};
    int type = 0;
    hi = 0;
    lo = 0;
    for (u32 lvl = 0; lvl < MAX_DEPTH; ++lvl) {
        const u32 half = 1u << (MAX_DEPTH - lvl - 1);
        u32 eq[3], bit[3];
        for (int a = 0; a < 3; ++a) {
            const u32 m = (2 * k[a] + 1) * half;
            eq[a] = v[a] == m;
            bit[a] = v[a] > m;
        }
        u32 idx = 0;
        int next = -1;   // -1 = terminal vertex call
        switch (type) {
        case 0: {
            const u32 ne = eq[0] + eq[1] + eq[2];
            if (ne == 0) { idx = bit[0] | (bit[1] << 1) | (bit[2] << 2); next = 0; }
            else if (ne == 1) {
                if (eq[0]) { idx = 8 + 2 * bit[1] + bit[2]; next = 1; }
                else if (eq[1]) { idx = 12 + bit[0] + 2 * bit[2]; next = 2; }
                else { idx = 16 + bit[0] + 2 * bit[1]; next = 3; }
            } else if (ne == 2) {
                if (!eq[2]) { idx = 20 + bit[2]; next = 6; }
                else if (!eq[1]) { idx = 22 + bit[1]; next = 5; }
                else { idx = 24 + bit[0]; next = 4; }
            } else { idx = 26; }
            break;
        }
        case 1:   // face with normal x: y, z free
            if (!eq[1] && !eq[2]) { idx = 2 * bit[1] + bit[2]; next = 1; }
            else if (eq[1] && !eq[2]) { idx = 4 + bit[2]; next = 6; }
            else if (!eq[1] && eq[2]) { idx = 6 + bit[1]; next = 5; }
            else idx = 8;
            break;
        case 2:   // face with normal y: x, z free
            if (!eq[0] && !eq[2]) { idx = bit[0] + 2 * bit[2]; next = 2; }
            else if (eq[0] && !eq[2]) { idx = 4 + bit[2]; next = 6; }
            else if (!eq[0] && eq[2]) { idx = 6 + bit[0]; next = 4; }
            else idx = 8;
            break;
        case 3:   // face with normal z: x, y free
            if (!eq[0] && !eq[1]) { idx = bit[0] + 2 * bit[1]; next = 3; }
            else if (eq[0] && !eq[1]) { idx = 4 + bit[1]; next = 5; }
            else if (!eq[0] && eq[1]) { idx = 6 + bit[0]; next = 4; }
            else idx = 8;
            break;
        default: {   // edges: 4 along x, 5 along y, 6 along z
            const int a = type - 4;
            if (!eq[a]) { idx = bit[a]; next = type; }
            else idx = 2;
            break;
        }
        }
        if (lvl < 12) hi |= (u64)idx << (5 * (11 - lvl) + 4);
        else lo |= (u64)idx << (5 * (18 - lvl) + 4);
        if (next < 0) return;
        // descend: free axes refine, fixed axes keep their plane
        const bool fixedAxis[3] = {next == 1 || next == 5 || next == 6, next == 2 || next == 4 || next == 6,
                                   next == 3 || next == 4 || next == 5};
        for (int a = 0; a < 3; ++a) k[a] = fixedAxis[a] ? k[a] * 2 : 2 * k[a] + bit[a];
        type = next;
    }
}

// The leaf holding finest-grid cell (cx, cy, cz).  `guess` is the inner depth found for a
// neighbouring cell (-1: none): the depth of the deepest inner ancestor is first tested at
// guess and guess +- 1 (2:1 balanced neighbours), then found by bisection; the result is
// the same either way and is returned in `guess`.
__device__ u32 leafOfCell(HashView inner, const u32 *__restrict__ nodeToLeaf, u32 cx, u32 cy, u32 cz,
                          int &guess) {
    const u64 finest = makeLoc(cx, cy, cz, MAX_DEPTH);
    if (guess >= 0 && guess < (int)MAX_DEPTH) {
        // the node at the guessed leaf depth: when it exists and is a leaf, it is the one
        // leaf holding the point
        const u32 id = nodeIdByRank(inner, finest >> (3 * (MAX_DEPTH - (u32)(guess + 1))));
        if (id != INVALID && nodeToLeaf[id] != INVALID) return nodeToLeaf[id];
    }
    auto innerAt = [&](int m) { return m < (int)MAX_DEPTH && inner.contains(finest >> (3 * (MAX_DEPTH - m))); };
    // largest m with the depth-m ancestor inner (monotone in m): bracketed by steps of 1, 2,
    // 4, ... away from the guess (neighbouring leaves differ by a level or two), then bisected
    int lo = 0, hi = (int)MAX_DEPTH;   // inner(lo) true (root), inner(hi) false
    if (guess < 0) {
        if (!inner.contains(1ull)) return nodeToLeaf[0];
    } else if (innerAt(guess)) {
        lo = guess;
        for (int step = 1;; step *= 2) {
            const int m = lo + step;
            if (m >= (int)MAX_DEPTH) break;
            if (!innerAt(m)) {
                hi = m;
                break;
            }
            lo = m;
        }
    } else {
        hi = guess;
        for (int step = 1;; step *= 2) {
            const int m = hi - step;
            if (m <= 0) break;   // the root is inner
            if (innerAt(m)) {
                lo = m;
                break;
            }
            hi = m;
        }
    }
    while (hi - lo > 1) {
        const int mid = (lo + hi) / 2;
        if (inner.contains(finest >> (3 * (MAX_DEPTH - mid)))) lo = mid;
        else hi = mid;
    }
    guess = lo;
    const u64 leafLoc = finest >> (3 * (MAX_DEPTH - (lo + 1)));
    const u32 id = nodeIdByRank(inner, leafLoc);
    return id == INVALID ? INVALID : nodeToLeaf[id];
}

__global__ void k_checkCells(const u32 *__restrict__ cells, size_t n, u32 *err) {
    const size_t j = (size_t)blockIdx.x * BLOCK + threadIdx.x;
    if (j < n && cells[j] == INVALID) atomicOr(err, (u32)ERR_NOT_INSERTED);
}

// keyHi / keyLo: the path key halves shifted right by hiShift / loShift (bits below
// that no level of this tree sets), so the sorts start at bit 0
__global__ void k_cells(HashView inner, const u32 *__restrict__ nodeToLeaf,
                        const u64 *__restrict__ vcodes, u32 count, u32 D, u32 hiShift, u32 loShift,
                        u32 *__restrict__ cellsUnsorted, u64 *__restrict__ keyHi, u64 *__restrict__ keyLo,
                        u32 *__restrict__ idx) {
    const u32 c = blockIdx.x * BLOCK + threadIdx.x;
    if (c >= count) return;
    const u64 code = vcodes[c];
    const u64 m = (1ull << D) - 1ull;
    const u32 up = MAX_DEPTH - D;   // vertex coordinates on the MAX_DEPTH grid
    const u32 v[3] = {(u32)(code & m) << up, (u32)((code >> D) & m) << up, (u32)((code >> (2 * D)) & m) << up};
    int guess = -1;
    for (u32 s = 0; s < 8; ++s) {
        const u32 x = v[0] - 1 + (s & 1), y = v[1] - 1 + ((s >> 1) & 1), z = v[2] - 1 + ((s >> 2) & 1);
        cellsUnsorted[8 * (size_t)c + s] = leafOfCell(inner, nodeToLeaf, x, y, z, guess);
    }
    u64 h, l;
    swKey(v[0], v[1], v[2], h, l);
    keyHi[c] = h >> hiShift;
    keyLo[c] = l >> loShift;
    idx[c] = c;
}

// ---------------------------------------------------------------------------
// Host replay of the order-dependent sampling refinement (Nodes.cpp:214-219,
// 271-337).  Units are (sample, side) in reference order.  Every unit that the
// GPU found reliable on the containment tree T0 stays a no-op unless a leaf it
// visited on T0 gets subdivided by an earlier unit; creating units and such
// disturbed units are re-run here in order, on the evolving tree.
// ---------------------------------------------------------------------------
// open-addressing set of location codes (host), key 0 = empty
struct FlatSet {
    std::vector<u64> keys;
    u64 mask = 0;
    void init(size_t expected) {
        size_t cap = 1024;
        while (cap < 2 * expected + 16) cap <<= 1;
        keys.assign(cap, 0ull);
        mask = cap - 1;
    }
    bool contains(u64 k) const {
        for (u64 s = hashMix(k) & mask;; s = (s + 1) & mask) {
            const u64 v = keys[s];
            if (v == k) return true;
            if (v == 0ull) return false;
        }
    }
    bool insert(u64 k) {
        for (u64 s = hashMix(k) & mask;; s = (s + 1) & mask) {
            const u64 v = keys[s];
            if (v == k) return false;
            if (v == 0ull) { keys[s] = k; return true; }
        }
    }
    // insert from several threads at once (no concurrent readers)
    void insertConcurrent(u64 k) {
        for (u64 s = hashMix(k) & mask;; s = (s + 1) & mask) {
            u64 expected = 0ull;
            if (__atomic_compare_exchange_n(&keys[s], &expected, k, false, __ATOMIC_RELAXED, __ATOMIC_RELAXED) ||
                expected == k)
                return;
        }
    }
};

// open-addressing u64 -> u32 map (key 0 = empty slot), host side.  Key and value share a
// 16-byte slot, so a lookup that hits costs one cache line.
struct FlatMap {
    struct Slot {
        u64 key;
        u32 val;
        u32 pad;
    };
    std::vector<Slot> slots;
    u64 mask = 0;
    void init(size_t expected) {
        size_t cap = 1024;
        while (cap < 2 * expected + 16) cap <<= 1;
        slots.assign(cap, Slot{0ull, 0u, 0u});
        mask = cap - 1;
    }
    size_t capacity() const { return slots.size(); }
    const u32 *find(u64 k) const {
        for (u64 s = hashMix(k) & mask;; s = (s + 1) & mask) {
            const u64 v = slots[s].key;
            if (v == k) return &slots[s].val;
            if (v == 0ull) return nullptr;
        }
    }
    void put(u64 k, u32 val) {
        for (u64 s = hashMix(k) & mask;; s = (s + 1) & mask) {
            const u64 v = slots[s].key;
            if (v == k || v == 0ull) {
                slots[s].key = k;
                slots[s].val = val;
                return;
            }
        }
    }
    // from several threads at once (no concurrent readers): insert k -> val if k is
    // absent; returns the slot of k
    u64 putIfAbsentConcurrent(u64 k, u32 val) {
        for (u64 s = hashMix(k) & mask;; s = (s + 1) & mask) {
            u64 expected = 0ull;
            if (__atomic_compare_exchange_n(&slots[s].key, &expected, k, false, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {
                slots[s].val = val;
                return s;
            }
            if (expected == k) return s;
        }
    }
};

// host view of the (T0 leaf -> units that visited it) index for the "hot" leaves
// (those near units the speculation could not settle); other leaves are looked
// up on the device on demand
// One slot of the hot-leaf table, laid out like the device HashView (same hash, linear
// probing, key 0 = empty) with the leaf's run of visiting units in the same 16 bytes, so a
// lookup costs one cache line.
struct RunSlot {
    u64 key;
    u32 first, count;   // units [first, first + count) of the unit list; count 0: no visits
};
struct RunSlotView {
    const RunSlot *slots = nullptr;
    u64 mask = 0;
    const RunSlot *find(u64 k) const {
        for (u64 s = hashMix(k) & mask;; s = (s + 1) & mask) {
            const u64 v = slots[s].key;
            if (v == k) return &slots[s];
            if (v == 0ull) return nullptr;
        }
    }
};

struct VisitIndex {
    RunSlotView run;                 // hot leaf -> its run of visiting units
    const u32 *unit = nullptr;       // visiting units, ascending within a run
    std::function<void(u64, std::vector<u32> &)> fetch;   // exact lookup on the device
    u64 misses = 0;
    // [first, last) of the units that visited T0 leaf `loc`, ascending; `spill`
    // holds the list of a leaf outside the hot part
    void units(u64 loc, const u32 *&first, const u32 *&last, std::vector<u32> &spill) {
        first = last = nullptr;
        if (const RunSlot *r = run.find(loc)) {
            if (r->count) {
                first = unit + r->first;
                last = first + r->count;
            }
            return;
        }
        ++misses;
        fetch(loc, spill);
        first = spill.data();
        last = spill.data() + spill.size();
    }
};

// Host set of inner nodes for the in-order replay, laid out for the DFS: one map
// entry per inner node holding the bit mask of its inner children, so that the
// eight siblings a DFS step tests cost one map lookup (cached) instead of eight
// scattered ones.  Inner-ness is prefix-closed (parents of inner nodes are inner).
struct InnerSet {
    FlatMap map;              // inner node -> mask of its inner children
    size_t count = 0;
    mutable u64 cacheKey = 0;
    mutable u32 cacheMask = 0;
    void init(size_t expected) { map.init(expected); }
    bool contains(u64 loc) const {
        if (loc == 1ull) return map.find(1ull) != nullptr;
        const u64 parent = loc >> 3;
        if (parent != cacheKey) {
            const u32 *m = map.find(parent);
            if (!m) return false;
            cacheKey = parent;
            cacheMask = *m;
        }
        return (cacheMask >> (loc & 7)) & 1u;
    }
    // loc's parent must be inner already (or loc is the root)
    void insert(u64 loc) {
        if (map.find(loc)) return;
        map.put(loc, 0u);
        ++count;
        if (loc != 1ull) {
            const u64 parent = loc >> 3;
            const u32 *m = map.find(parent);
            map.put(parent, (m ? *m : 0u) | (1u << (loc & 7)));
            if (cacheKey == parent) cacheKey = 0;
        }
    }
    size_t capacity() const { return map.capacity(); }
    // all keys of a prefix-closed set at once (several threads)
    void build(const u64 *keys, size_t n) {
        parallelFor(n, [&](size_t b, size_t e) {
            for (size_t i = b; i < e; ++i) map.putIfAbsentConcurrent(keys[i], 0u);
        });
        parallelFor(n, [&](size_t b, size_t e) {
            for (size_t i = b; i < e; ++i) {
                const u64 k = keys[i];
                if (k == 1ull) continue;
                const u64 slot = map.putIfAbsentConcurrent(k >> 3, 0u);   // the parent is present
                __atomic_fetch_or(&map.slots[slot].val, 1u << (k & 7), __ATOMIC_RELAXED);
            }
        });
        count = n;
    }
};

// read-only view of an InnerSet with its own one-entry cache (one per thread)
struct InnerView {
    const FlatMap *map;
    u64 cacheKey = 0;
    u32 cacheMask = 0;
    bool contains(u64 loc) {
        if (loc == 1ull) return map->find(1ull) != nullptr;
        const u64 parent = loc >> 3;
        if (parent != cacheKey) {
            const u32 *m = map->find(parent);
            if (!m) return false;
            cacheKey = parent;
            cacheMask = *m;
        }
        return (cacheMask >> (loc & 7)) & 1u;
    }
};

// host copies of the containment tree's inner set for the replay: T0 (frozen) and the
// growing inner set (built on host threads while the GPU runs the speculation)
struct ReplaySets {
    FlatSet t0;
    InnerSet inner;
};
ReplaySets buildReplaySets(const u64 *t0Inner, size_t n) {
    ReplaySets r;
    r.t0.init(n);
    r.inner.init(n + 1024);   // grown by the replay when needed
    parallelFor(n, [&](size_t b, size_t e) {
        for (size_t i = b; i < e; ++i) r.t0.insertConcurrent(t0Inner[i]);
    });
    r.inner.build(t0Inner, n);
    return r;
}

unsigned hostThreads();   // worker threads of parallelFor (defined with its pool below)

// Outcome of the refinement of one unit whose reliable-leaf search failed.
enum class Refine { DONE, TOO_DEEP, ZERO_SIZE, LEFT_NODE };

// Nodes.cpp:284-336: subdivide along the child with the largest overlap of the segment
// [p, p +- n msd] until a reliable leaf exists.  `create(n)` turns leaf n into an inner node
// (isInner(n) answers true afterwards).  The descent from the root passes, without creating
// anything, through the inner nodes that strictly contain the segment (in each, the one
// child holding the whole segment has overlap 1 and every other child misses it), so it
// starts at the deepest of them.  The result is a function of the isInner answers alone.
template <typename IsInner, typename Create>
Refine refineUnit(const Frame &f, const SamplingUnit &u, bool positive, IsInner isInner, Create create) {
    const double sgn = positive ? u.msd : -u.msd;
    const double en[3] = {u.p[0] + u.n[0] * sgn, u.p[1] + u.n[1] * sgn, u.p[2] + u.n[2] * sgn};
    double s0[3], s1[3];
    for (int a = 0; a < 3; ++a) {
        const double m = 1e-9 * std::fabs(en[a] - u.p[a]) + 1e-12 * f.rootSize;
        s0[a] = std::min(u.p[a], en[a]) - m;
        s1[a] = std::max(u.p[a], en[a]) + m;
    }
    u64 n = deepestContaining(f, s0, s1, true, isInner);
    auto noLeaf = [](u64) {};
    for (;;) {
        if (!isInner(n)) {
            if (locDepth(n) + 1 > MAX_DEPTH) return Refine::TOO_DEEP;
            create(n);
            if (findReliableHost(f, n, positive, u, isInner, noLeaf)) return Refine::DONE;
        }
        const u32 d = locDepth(n) + 1;
        if (f.h(d) == 0.0) return Refine::ZERO_SIZE;
        double score = -1.0e300;
        u32 best = INVALID;
        // segmentBox for the 8 children: per axis the children's bounds are the three
        // planes B(2k), B(2k+1), B(2k+2) at depth d, so their slab parameters
        // (plane - p) / dir are formed once (the same quotients segmentBox forms per child)
        u32 kp[3];
        locCoords(n, kp[0], kp[1], kp[2]);
        double plane[3][3], q[3][3], dir[3];
        for (int a = 0; a < 3; ++a) {
            dir[a] = en[a] - u.p[a];
            for (u32 j = 0; j < 3; ++j) {
                plane[a][j] = f.B(a, 2 * kp[a] + j, d);
                if (dir[a] != 0.0) q[a][j] = (plane[a][j] - u.p[a]) / dir[a];
            }
        }
        auto childSegment = [&](u32 c, double &tIn, double &tOut) {
            tIn = 0.0;
            tOut = 1.0;
            for (int a = 0; a < 3; ++a) {
                const u32 b = (c >> a) & 1u;
                if (dir[a] == 0.0) {
                    if (u.p[a] < plane[a][b] || u.p[a] > plane[a][b + 1]) return false;
                    continue;
                }
                double t1 = q[a][b], t2 = q[a][b + 1];
                if (t1 > t2) { const double tmp = t1; t1 = t2; t2 = tmp; }
                if (t1 > tIn) tIn = t1;
                if (t2 < tOut) tOut = t2;
            }
            return tIn <= tOut;
        };
        for (u32 c = 0; c < 8; ++c) {
            double tIn, tOut;
            if (!childSegment(c, tIn, tOut)) continue;
            const double overlap = tOut - tIn;
            if (overlap <= score) continue;
            score = overlap;
            best = c;
            if (overlap >= 1.0 - EPS_D) break;
        }
        if (best == INVALID) return Refine::LEFT_NODE;
        n = (n << 3) | best;
    }
}

void throwRefine(Refine r) {
    switch (r) {
    case Refine::TOO_DEEP: throw tsr_error(TSR_ERR_TREE_TOO_DEEP, "octree deeper than 19 levels");
    case Refine::ZERO_SIZE: throw tsr_error(TSR_ERR_ZERO_SIZE_NODE, "Reached a tree node of size 0");
    case Refine::LEFT_NODE: throw tsr_error(TSR_ERR_INVALID_ARGUMENT, "sampling refinement: segment leaves the node");
    case Refine::DONE: break;
    }
}

void replaySampling(const Frame &f, const tsr_scene *scene, float maxRel, ReplaySets &sets,
                    u64 *queued, u32 units, VisitIndex &index, std::vector<u64> &created,
                    u64 &eventCount) {
    const auto tInit = std::chrono::steady_clock::now();
    FlatSet &t0 = sets.t0;
    InnerSet &inner = sets.inner;
    // The units still to run, in ascending order, are the set bits of `queued` (one bit per
    // unit, starting as the event flags): the replay only ever asks for the smallest queued
    // unit in a range.  `pendingBits` marks the units that were ever queued (events, then
    // tainted units); one bit per unit keeps it cache-resident for the taints' random probes.
    const size_t queuedWords = ((size_t)units + 63) / 64;
    std::vector<u64> pendingBits(queued, queued + queuedWords);
    size_t initialEvents = 0;
    for (size_t w = 0; w < queuedWords; ++w) initialEvents += (size_t)__builtin_popcountll(queued[w]);
    auto enqueue = [&](u32 v) { queued[v >> 6] |= 1ull << (v & 63); };
    // the smallest queued unit in [from, end) (removed from the queue), or INVALID
    auto takeFrom = [&](u32 from, u32 end) -> u32 {
        if (from >= end) return INVALID;
        const size_t last = std::min(queuedWords - 1, ((size_t)end - 1) >> 6);
        size_t w = from >> 6;
        u64 bits = queued[w] & (~0ull << (from & 63));
        while (!bits) {
            if (++w > last) return INVALID;
            bits = queued[w];
        }
        const u32 v = (u32)(w * 64 + (size_t)__builtin_ctzll(bits));
        if (v >= end) return INVALID;
        queued[w] &= ~(1ull << (v & 63));
        return v;
    };
    u64 tainted = 0, creating = 0;
    auto isInner = [&](u64 loc) { return inner.contains(loc); };
    auto noLeaf = [](u64) {};
    std::vector<u32> tmp;
    auto loadUnitHost = [&](u32 w, SamplingUnit &su) {
        const u32 i = w >> 1;
        su.p[0] = scene->hpx[i]; su.p[1] = scene->hpy[i]; su.p[2] = scene->hpz[i];
        su.n[0] = scene->hnx[i]; su.n[1] = scene->hny[i]; su.n[2] = scene->hnz[i];
        su.msd = (double)(scene->hscale[i] * maxRel);
    };
    u64 candidates = 0;
    double taintMs = 0, fullMs = 0, specMs = 0, createMs = 0;
    using clock = std::chrono::steady_clock;
    auto msSince = [](clock::time_point t) { return std::chrono::duration<double, std::milli>(clock::now() - t).count(); };
    // the T0 leaf containing a subdivided node
    auto t0LeafOf = [&](u64 subdivided) {
        u64 l = subdivided;
        while (l > 1 && !t0.contains(l >> 3)) l >>= 3;
        return l;
    };
    // the units of [first, last) after `writer` become pending
    auto taintRange = [&](const u32 *first, const u32 *last, u32 writer) {
        const u32 *from = std::upper_bound(first, last, writer);
        candidates += (u64)(last - from);
        for (const u32 *it = from; it != last; ++it) {
            u64 &word = pendingBits[*it >> 6];
            const u64 bit = 1ull << (*it & 63);
            if (word & bit) continue;
            ++tainted;
            word |= bit;
            enqueue(*it);
        }
    };
    auto taint = [&](u64 subdivided, u32 writer) {
        const auto tt = clock::now();
        const u32 *first, *last;
        index.units(t0LeafOf(subdivided), first, last, tmp);
        taintRange(first, last, writer);
        taintMs += msSince(tt);
    };
    eventCount = 0;
    // nodes made inner since the current batch was speculated
    FlatSet createdSince;
    std::vector<u64> createdList;
    createdSince.init(1u << 10);   // a batch's creations; grown when needed
    // leaf n becomes an inner node of the evolving tree
    auto createNode = [&](u64 n) {
        if (inner.capacity() < 2 * (inner.count + 16)) {   // keep the load factor <= 1/2
            InnerSet grown;
            grown.init(2 * inner.count + 1024);
            for (const FlatMap::Slot &sl : inner.map.slots)
                if (sl.key) grown.map.put(sl.key, sl.val);
            grown.count = inner.count;
            inner = std::move(grown);
        }
        inner.insert(n);
        created.push_back(n);
        if (createdList.size() * 2 + 2 > createdSince.keys.size()) {
            FlatSet grown;
            grown.init(createdList.size() * 2 + 16);
            for (u64 k : createdList) grown.insert(k);
            createdSince = std::move(grown);
        }
        if (createdSince.insert(n)) createdList.push_back(n);
    };
    // the sequential semantics: evaluate `unit` on the current tree, refine if needed
    auto processFull = [&](u32 unit, bool knownToFail) {
        ++eventCount;
        const auto tf = clock::now();
        struct Acc {
            double &ms;
            clock::time_point t;
            ~Acc() { ms += std::chrono::duration<double, std::milli>(clock::now() - t).count(); }
        } acc{fullMs, tf};
        const bool positive = (unit & 1) == 0;
        SamplingUnit u;
        loadUnitHost(unit, u);
        if (!knownToFail && findReliableHost(f, searchStart(f, u, isInner), positive, u, isInner, noLeaf)) return;
        ++creating;
        const auto tc = clock::now();
        struct AccC {
            double &ms;
            clock::time_point t;
            ~AccC() { ms += std::chrono::duration<double, std::milli>(clock::now() - t).count(); }
        } accC{createMs, tc};
        // the nodes one descent creates form a chain below the first, so they all lie in
        // the same containment-tree leaf: after the first taint the others add no unit
        bool taintDone = false;
        throwRefine(refineUnit(f, u, positive, isInner, [&](u64 n) {
            createNode(n);
            if (!taintDone) {
                taint(n, unit);
                taintDone = true;
            }
        }));
    };
    // Batches of the next pending units (in order) are first evaluated in parallel on
    // the tree as it is when the batch starts: the reliable-leaf search, recording the
    // leaves it visits, and for a unit whose search fails, the refinement too, with the
    // nodes it creates kept in a per-unit overlay and every "not inner" answer of the
    // batch's tree recorded.  Committing in order, a unit none of whose recorded nodes has
    // been made inner since is settled by its speculation: a search only changes when a
    // leaf it visits becomes an inner node, and the refinement is a function of its
    // inner-ness answers, of which only "not inner" ones can change.  A settled refinement
    // inserts its overlay and taints as the sequential step would.  Every other unit, and
    // units made pending meanwhile that precede the rest of the batch, run the sequential
    // step above.
    constexpr u32 BATCH = 4096, SPEC_CAP = 48;
    struct Spec {
        // 0 = visit cap exceeded, 1 = reliable leaf found, 2 = none and the refinement
        // speculated, 3 = none and the refinement not speculated
        u32 state;
        u32 count;      // visited[0, count): the leaves the search visited
        u32 total;      // visited[count, total): the refinement's "not inner" answers
        u32 chainLen;   // the nodes the refinement creates, in order
        u64 visited[SPEC_CAP];
        u64 chain[MAX_DEPTH + 1];
        const u32 *taintFirst, *taintLast;   // the first creation's units, when in the hot index
        bool taintKnown;
    };
    std::vector<u32> batch;
    std::vector<Spec> spec(BATCH);
    u64 speculated = 0, specRefined = 0;
    // sequential steps by cause: made pending within the batch's range, search over the visit
    // cap, search invalidated, refinement not speculated, refinement invalidated
    u64 seqPending = 0, seqCap = 0, seqInvalid = 0, seqUnspeculated = 0, seqInvalidRefine = 0;
    const auto tLoop = std::chrono::steady_clock::now();
    u32 cursor = 0;   // every queued unit below it has run
    for (;;) {
        batch.clear();
        for (u32 from = cursor; batch.size() < BATCH;) {
            const u32 v = takeFrom(from, units);
            if (v == INVALID) break;
            batch.push_back(v);
            from = v + 1;
        }
        if (batch.empty()) break;
        const auto ts = clock::now();
        // units are taken in runs of 16 from a shared counter: their searches differ widely
        // in length, so a static split leaves threads idle
        std::atomic<size_t> nextUnit{0};
        const size_t batchSize = batch.size();
        parallelFor(hostThreads(), [&](size_t, size_t) {
            InnerView view{&inner.map};
            auto inView = [&](u64 loc) { return view.contains(loc); };
            for (size_t j0; (j0 = nextUnit.fetch_add(16, std::memory_order_relaxed)) < batchSize;)
            for (size_t j = j0; j < std::min(batchSize, j0 + 16); ++j) {
                Spec &sp = spec[j];
                sp.count = 0;
                bool over = false;
                auto record = [&](u64 loc) {
                    if (sp.count < SPEC_CAP) sp.visited[sp.count++] = loc;
                    else over = true;
                };
                SamplingUnit su;
                loadUnitHost(batch[j], su);
                const bool positive = (batch[j] & 1) == 0;
                const bool ok = findReliableHost(f, searchStart(f, su, inView), positive, su, inView, record);
                sp.state = over ? 0u : ok ? 1u : 2u;
                sp.total = sp.count;
                if (sp.state != 2u) continue;
                sp.chainLen = 0;
                auto specInner = [&](u64 loc) {
                    for (u32 k = 0; k < sp.chainLen; ++k)
                        if (sp.chain[k] == loc) return true;
                    if (sp.chainLen) {   // below the first created node nothing else is inner
                        const u32 d0 = locDepth(sp.chain[0]), d = locDepth(loc);
                        if (d > d0 && (loc >> (3 * (d - d0))) == sp.chain[0]) return false;
                    }
                    if (view.contains(loc)) return true;
                    if (sp.total < SPEC_CAP) sp.visited[sp.total++] = loc;
                    else over = true;
                    return false;
                };
                const Refine r = refineUnit(f, su, positive, specInner, [&](u64 n) {
                    if (sp.chainLen <= MAX_DEPTH) sp.chain[sp.chainLen++] = n;
                    else over = true;
                });
                if (over || r != Refine::DONE || sp.chainLen == 0) {
                    sp.state = 3u;
                    continue;
                }
                // the first creation's taint target (a leaf outside the hot part of the
                // index is looked up on the device at commit)
                sp.taintKnown = false;
                if (const RunSlot *rs = index.run.find(t0LeafOf(sp.chain[0]))) {
                    sp.taintFirst = sp.taintLast = nullptr;
                    if (rs->count) {
                        sp.taintFirst = index.unit + rs->first;
                        sp.taintLast = sp.taintFirst + rs->count;
                    }
                    sp.taintKnown = true;
                }
            }
        }, 1);
        specMs += msSince(ts);
        if (!createdList.empty()) {   // a new batch: nothing created since it was speculated
            std::fill(createdSince.keys.begin(), createdSince.keys.end(), 0ull);
            createdList.clear();
        }
        // the map slots a settled refinement's inserts touch, requested ahead of its commit
        auto prefetchChain = [&](const Spec &sp) {
            if (sp.state != 2u) return;
            const FlatMap::Slot *slots = inner.map.slots.data();
            const u64 mask = inner.map.mask;
            __builtin_prefetch(&slots[hashMix(sp.chain[0] >> 3) & mask], 1);
            for (u32 k = 0; k < sp.chainLen; ++k) __builtin_prefetch(&slots[hashMix(sp.chain[k]) & mask], 1);
        };
        constexpr size_t AHEAD = 8;
        for (size_t j = 0; j < std::min(AHEAD, batch.size()); ++j) prefetchChain(spec[j]);
        u32 done = cursor;   // units below `done` have run
        for (size_t j = 0; j < batch.size(); ++j) {
            const u32 u = batch[j];
            if (j + AHEAD < batch.size()) prefetchChain(spec[j + AHEAD]);
            for (;;) {   // made pending after the batch was formed, before u
                const u32 v = takeFrom(done, u);
                if (v == INVALID) break;
                ++seqPending;
                processFull(v, false);
                done = v + 1;
            }
            done = u + 1;
            const Spec &sp = spec[j];
            bool valid = sp.state != 0u;
            for (u32 i = 0; valid && i < sp.count && !createdList.empty(); ++i)
                if (createdSince.contains(sp.visited[i])) valid = false;
            if (valid && sp.state == 1u) {
                ++eventCount;
                ++speculated;
                continue;
            }
            if (valid && sp.state == 2u) {
                bool refined = true;
                for (u32 i = sp.count; refined && i < sp.total && !createdList.empty(); ++i)
                    if (createdSince.contains(sp.visited[i])) refined = false;
                if (refined) {
                    ++eventCount;
                    ++creating;
                    ++specRefined;
                    for (u32 k = 0; k < sp.chainLen; ++k) createNode(sp.chain[k]);
                    if (sp.taintKnown) {
                        const auto tt = clock::now();
                        if (sp.taintFirst) taintRange(sp.taintFirst, sp.taintLast, u);
                        taintMs += msSince(tt);
                    } else {
                        taint(sp.chain[0], u);
                    }
                    continue;
                }
            }
            ++(sp.state == 0u ? seqCap : !valid ? seqInvalid : sp.state == 3u ? seqUnspeculated : seqInvalidRefine);
            processFull(u, valid);   // valid here: the search is known to find no reliable leaf
        }
        cursor = batch.back() + 1;   // units made pending from here on come after the batch
    }
    if (std::getenv("TSR_TRACE"))
        std::fprintf(stderr, "[tsr tree]   replay set-up %.2f ms, in-order loop %.2f ms; initial events %zu, "
                     "creating %llu (refinement settled by speculation %llu), tainted %llu, searches settled by speculation %llu\n"
                     "[tsr tree]   parallel speculation %.1f ms, sequential steps %.1f ms (refinement %.1f ms, "
                     "taints %.1f ms over %llu candidates)\n",
                     std::chrono::duration<double, std::milli>(tLoop - tInit).count(),
                     std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tLoop).count(),
                     initialEvents, (unsigned long long)creating, (unsigned long long)specRefined,
                     (unsigned long long)tainted,
                     (unsigned long long)speculated, specMs, fullMs, createMs, taintMs,
                     (unsigned long long)candidates);
    if (std::getenv("TSR_TRACE"))
        std::fprintf(stderr, "[tsr tree]   sequential steps: pending in the batch range %llu, over the visit cap %llu, "
                     "search invalidated %llu, refinement not speculated %llu, refinement invalidated %llu\n",
                     (unsigned long long)seqPending, (unsigned long long)seqCap, (unsigned long long)seqInvalid,
                     (unsigned long long)seqUnspeculated, (unsigned long long)seqInvalidRefine);
}

// T0 leaves within HOT_RADIUS x max sampling distance of a unit the speculation could not
// settle: the writes of the in-order replay (and of short cascades) land there.
__global__ void k_markHot(Frame f, HashView inner, HashView hot, const float *__restrict__ px,
                          const float *__restrict__ py, const float *__restrict__ pz, const float *__restrict__ scale,
                          float maxRel, double hotRadius, const uint8_t *__restrict__ isEvent, u32 units, u32 *err) {
    const u32 unit = blockIdx.x * BLOCK + threadIdx.x;
    if (unit >= units || !isEvent[unit] || ((unit & 1) && isEvent[unit - 1])) return;
    const u32 i = unit >> 1;
    const double p[3] = {px[i], py[i], pz[i]};
    const double r = hotRadius * (double)(scale[i] * maxRel);
    auto overlaps = [&](u64 loc) {
        u32 k[3];
        locCoords(loc, k[0], k[1], k[2]);
        const u32 d = locDepth(loc);
        double d2 = 0.0;
        for (int a = 0; a < 3; ++a) {
            const double lo = f.B(a, k[a], d), hi = f.B(a, (u64)k[a] + 1, d);
            if (p[a] < lo) d2 += (lo - p[a]) * (lo - p[a]);
            else if (p[a] > hi) d2 += (p[a] - hi) * (p[a] - hi);
        }
        return d2 <= r * r;
    };
    if (!inner.contains(1ull)) {
        hot.insert(1ull, 0, err);
        return;
    }
    InnerMasks isInner{inner};
    u64 loc = 1ull;
    u32 next = 0;
    for (;;) {
        if (next == 8) {
            if (loc == 1ull) return;
            next = (u32)(loc & 7) + 1;
            loc >>= 3;
            continue;
        }
        const u64 ch = (loc << 3) | next;
        if (overlaps(ch)) {
            if (isInner(ch)) {
                loc = ch;
                next = 0;
                continue;
            }
            hot.insert(ch, 0, err);
        }
        ++next;
    }
}

// hot visits per unit (non-event units only) from the staged visits; mode 1 writes them
// The visits of settled units to hot leaves: mode 0 counts them per unit and records which
// of the unit's visits they are (bit j: visit j, VISIT_CAP <= 64); mode 1 writes them as
// (leaf, unit) pairs at the unit's offset, reading only those visits.
static_assert(VISIT_CAP <= 64, "hot-visit masks hold one bit per staged visit");
__global__ void k_hotVisits(HashView hot, const u64 *__restrict__ staged, const u32 *__restrict__ visitCount,
                            const uint8_t *__restrict__ isEvent, u32 units, int mode, u32 *__restrict__ hotCount,
                            u64 *__restrict__ hotMask, const u32 *__restrict__ hotOff, u64 *__restrict__ outLeaf,
                            u32 *__restrict__ outUnit) {
    const u32 u = blockIdx.x * BLOCK + threadIdx.x;
    if (u >= units) return;
    if (mode == 0) {
        const u32 c = isEvent[u] ? 0u : visitCount[u];
        u64 m = 0;
        for (u32 j = 0; j < c; ++j)
            if (hot.contains(staged[visitSlot(u, j)])) m |= 1ull << j;
        hotMask[u] = m;
        hotCount[u] = (u32)__popcll(m);
        return;
    }
    const size_t base = hotOff[u];
    u32 k = 0;
    for (u64 m = hotMask[u]; m; m &= m - 1, ++k) {
        outLeaf[base + k] = staged[visitSlot(u, (u32)__ffsll((long long)m) - 1)];
        outUnit[base + k] = u;
    }
}

// the hot-leaf table as RunSlots: each occupied slot's key with its run of units (a run
// index of INVALID: no unit visited the leaf)
__global__ void k_packRunSlots(const u64 *__restrict__ keys, const u32 *__restrict__ vals, size_t cap,
                               const u32 *__restrict__ runOff, const u32 *__restrict__ runLen,
                               RunSlot *__restrict__ out) {
    const size_t s = (size_t)blockIdx.x * BLOCK + threadIdx.x;
    if (s >= cap) return;
    const u64 k = keys[s];
    const u32 r = k ? vals[s] : INVALID;
    RunSlot o;
    o.key = k;
    o.first = r != INVALID ? runOff[r] : 0u;
    o.count = r != INVALID ? runLen[r] : 0u;
    out[s] = o;
}

// the event flags as one bit per unit (bit u & 63 of word u >> 6)
__global__ void k_eventBits(const uint8_t *__restrict__ isEvent, u32 units, u64 *__restrict__ bits) {
    const u32 u = blockIdx.x * BLOCK + threadIdx.x;
    const u64 b = __ballot(u < units && isEvent[u] != 0);
    if ((threadIdx.x & 63) == 0 && u < units) bits[u >> 6] = b;
}

// units (ascending) whose staged visits include `key`
__global__ void k_matchStaged(const u64 *__restrict__ staged, const u32 *__restrict__ visitCount,
                              const uint8_t *__restrict__ isEvent, u32 units, u64 key, uint8_t *__restrict__ flag) {
    const u32 u = blockIdx.x * BLOCK + threadIdx.x;
    if (u >= units) return;
    const u32 c = isEvent[u] ? 0u : visitCount[u];
    uint8_t hit = 0;
    for (u32 j = 0; j < c && !hit; ++j) hit = staged[visitSlot(u, j)] == key;
    flag[u] = hit;
}

__global__ void k_hotFlags(HashView hot, const u64 *__restrict__ leaf, size_t n, uint8_t *__restrict__ flag) {
    const size_t j = (size_t)blockIdx.x * BLOCK + threadIdx.x;
    if (j < n) flag[j] = hot.contains(leaf[j]) ? 1 : 0;
}

// [first, last) of `key` in the sorted array (one thread)
__global__ void k_insertIndexed(HashView h, const u64 *__restrict__ keys, u32 n, u32 *err) {
    const u32 i = blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) h.insert(keys[i], i, err);
}
__global__ void k_insertConst(HashView h, const u64 *__restrict__ keys, u32 n, u32 val, u32 *err) {
    const u32 i = blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) h.insert(keys[i], val, err);
}

void checkErr(u32 e) {
    if (e & ERR_OUTSIDE) throw tsr_error(TSR_ERR_SAMPLE_OUTSIDE_ROOT, "Cannot find a container node for a position which is not contained by the start node for the search.");
    if (e & ERR_TOO_DEEP) throw tsr_error(TSR_ERR_TREE_TOO_DEEP, "octree deeper than 19 levels");
    if (e & ERR_HASH_FULL) throw tsr_error(TSR_ERR_NODE_OVERFLOW, "octree node table full");
    if (e & ERR_NOT_INSERTED) throw tsr_error(TSR_ERR_SAMPLE_NOT_INSERTED, "There is a sample which was not inserted into the Tree object!");
    if (e & 16u) throw tsr_error(TSR_ERR_INVALID_SCALE, "Invalid (non-positive) sample scale detected.");
}

}  // namespace

void *pinnedHost(tsr_context *ctx, size_t bytes, int slot) {
    if (bytes > ctx->pinnedBytes[slot]) {
        if (ctx->pinned[slot]) HIPCHECK(hipHostFree(ctx->pinned[slot]));
        ctx->pinned[slot] = nullptr;
        ctx->pinnedBytes[slot] = 0;
        size_t cap = (size_t)1 << 26;
        while (cap < bytes) cap <<= 1;
        HIPCHECK(hipHostMalloc(&ctx->pinned[slot], cap, hipHostMallocDefault));
        ctx->pinnedBytes[slot] = cap;
    }
    return ctx->pinned[slot];
}

namespace {
// persistent host worker threads for parallelFor (spawning threads per call costs
// ~20 us each, and the replay calls it once per batch)
class HostPool {
public:
    explicit HostPool(unsigned workers) {
        for (unsigned w = 0; w < workers; ++w) threads_.emplace_back([this, w] { loop(w); });
    }
    ~HostPool() {
        {
            std::lock_guard<std::mutex> g(m_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto &t : threads_) t.join();
    }
    unsigned workers() const { return (unsigned)threads_.size(); }
    // runs job(c) for c in [1, chunks) on the workers; the caller runs chunk 0 itself
    void run(size_t chunks, const std::function<void(size_t)> &job) {
        std::unique_lock<std::mutex> g(m_);
        job_ = &job;
        chunks_ = chunks;
        next_ = 1;
        pending_ = chunks - 1;
        ++generation_;
        g.unlock();
        cv_.notify_all();
        job(0);
        g.lock();
        done_.wait(g, [this] { return pending_ == 0; });
        job_ = nullptr;
    }
    std::mutex busy;   // one run at a time; other callers spawn their own threads

private:
    void loop(unsigned) {
        size_t seen = 0;
        std::unique_lock<std::mutex> g(m_);
        for (;;) {
            cv_.wait(g, [&] { return stop_ || (generation_ != seen && next_ < chunks_); });
            if (stop_) return;
            while (job_ && next_ < chunks_) {
                const size_t c = next_++;
                const std::function<void(size_t)> *job = job_;
                g.unlock();
                (*job)(c);
                g.lock();
                if (--pending_ == 0) done_.notify_one();
            }
            seen = generation_;
        }
    }
    std::vector<std::thread> threads_;
    std::mutex m_;
    std::condition_variable cv_, done_;
    const std::function<void(size_t)> *job_ = nullptr;
    size_t chunks_ = 0, next_ = 0, pending_ = 0, generation_ = 0;
    bool stop_ = false;
};

unsigned hostThreads() {
    static const unsigned threads = [] {
        unsigned t = std::max(1u, std::thread::hardware_concurrency());
        if (const char *e = std::getenv("OMP_NUM_THREADS")) t = std::max(1, std::atoi(e));
        return std::min(t, 16u);
    }();
    return threads;
}

HostPool &hostPool() {
    static HostPool pool(hostThreads() - 1);
    return pool;
}
}  // namespace

void parallelFor(size_t n, const std::function<void(size_t, size_t)> &fn, size_t grain) {
    const size_t chunks = std::min<size_t>(hostThreads(), (n + grain - 1) / std::max<size_t>(grain, 1));
    if (chunks <= 1) {
        fn(0, n);
        return;
    }
    const std::function<void(size_t)> job = [&](size_t c) { fn(n * c / chunks, n * (c + 1) / chunks); };
    HostPool &pool = hostPool();
    std::unique_lock<std::mutex> mine(pool.busy, std::try_to_lock);
    if (mine.owns_lock() && pool.workers() + 1 >= chunks) {
        pool.run(chunks, job);
        return;
    }
    std::vector<std::thread> extra;
    for (size_t c = 1; c < chunks; ++c) extra.emplace_back([&job, c] { job(c); });
    job(0);
    for (auto &t : extra) t.join();
}

// Tree.cpp:70-101 getRootScope over Samples.cpp:290-298,442-449 (AABB of the samples
// inflated by their max sampling distance + EPSILON, float arithmetic as Real = float);
// sample scales are checked first (Samples.cpp:280-288).
Frame rootFrame(tsr_context *ctx, const float *px, const float *py, const float *pz, const float *scale, u32 n,
                float maxRel) {
    hipStream_t st = ctx->stream;
    if (n == 0) throw tsr_error(TSR_ERR_NO_SAMPLES, "no samples");
    DBuf<u32> err(1);
    HIPCHECK(hipMemsetAsync(err.p, 0, 4, st));
    k_checkScales<<<gridFor(n, BLOCK), BLOCK, 0, st>>>(scale, n, err.p);
    checkErr(readScalar(err.p, st));
    const u32 nb = std::min<u32>(1024, gridFor(n, BLOCK));
    DBuf<float> partial(6 * (size_t)nb);
    k_aabb<<<nb, BLOCK, 0, st>>>(px, py, pz, scale, maxRel, n, partial.p);
    std::vector<float> hp(6 * (size_t)nb);
    HIPCHECK(hipMemcpyAsync(hp.data(), partial.p, partial.bytes(), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    float mn[3] = {3.4028235e38f, 3.4028235e38f, 3.4028235e38f}, mx[3] = {-3.4028235e38f, -3.4028235e38f, -3.4028235e38f};
    for (u32 b = 0; b < nb; ++b)
        for (int a = 0; a < 3; ++a) {
            mn[a] = std::min(mn[a], hp[6 * b + a]);
            mx[a] = std::max(mx[a], hp[6 * b + 3 + a]);
        }
    Frame f{};
    float c[3], e[3];
    for (int a = 0; a < 3; ++a) { c[a] = (mn[a] + mx[a]) * 0.5f; e[a] = mx[a] - mn[a]; }
    const float size = std::max(std::max(e[0], e[1]), e[2]);
    const float half = size * 0.5f;
    f.rootSizeF = size;
    f.rootSize = (double)size;
    for (int a = 0; a < 3; ++a) {
        f.rootMinF[a] = c[a] - half;
        f.rootMin[a] = (double)f.rootMinF[a];
    }
    return f;
}

// container location code and depth of every sample (Nodes.cpp:231-269 findContainer's stop)
void sampleContainers(tsr_context *ctx, const Frame &f, const float *px, const float *py, const float *pz,
                      const float *scale, u32 n, float maxRel, u64 *loc, uint8_t *depth) {
    hipStream_t st = ctx->stream;
    DBuf<u32> err(1);
    HIPCHECK(hipMemsetAsync(err.p, 0, 4, st));
    k_containers<<<gridFor(n, BLOCK), BLOCK, 0, st>>>(f, px, py, pz, scale, maxRel, n, loc, depth, err.p);
    checkErr(readScalar(err.p, st));
}

void buildTree(tsr_context *ctx, const tsr_scene *scene, const tsr_params &prm, tsr_tree *tree) {
    hipStream_t st = ctx->stream;
    // TSR_TRACE=1: per-stage wall times on stderr (synchronises after each stage)
    static const bool trace = std::getenv("TSR_TRACE") != nullptr;
    auto t0 = std::chrono::steady_clock::now();
    auto mark = [&](const char *stage) {
        if (!trace) return;
        HIPCHECK(hipStreamSynchronize(st));
        const auto t1 = std::chrono::steady_clock::now();
        std::fprintf(stderr, "[tsr tree] %-22s %8.2f ms\n", stage, std::chrono::duration<double, std::milli>(t1 - t0).count());
        t0 = t1;
    };
    Scratch &sc = ctx->scratch;
    const u32 n = scene->n;
    if (n == 0) throw tsr_error(TSR_ERR_NO_SAMPLES, "no samples");
    const float maxRel = prm.max_relative_sampling_distance;
    tree->maxRel = maxRel;
    tree->sampleCount = n;
    DBuf<u32> err(1);
    HIPCHECK(hipMemsetAsync(err.p, 0, 4, st));
    const Frame f = rootFrame(ctx, scene->px.p, scene->py.p, scene->pz.p, scene->scale.p, n, maxRel);
    tree->frame = f;
    mark("root scope");

    // ---- containment
    DBuf<u64> containerLoc(n);
    DBuf<uint8_t> cdepth(n);
    k_containers<<<gridFor(n, BLOCK), BLOCK, 0, st>>>(f, scene->px.p, scene->py.p, scene->pz.p, scene->scale.p, maxRel,
                                                       n, containerLoc.p, cdepth.p, err.p);
    checkErr(readScalar(err.p, st));
    DHash inner;
    inner.alloc(std::max<u64>(4096, (u64)n), st);
    k_insertAncestors<<<gridFor(n, BLOCK), BLOCK, 0, st>>>(inner.view(), containerLoc.p, n, err.p);
    checkErr(readScalar(err.p, st));

    u32 maxContainerDepth = 0;
    {
        DBuf<uint8_t> md(1);
        reduceMax(sc, cdepth.p, md.p, n, st);
        maxContainerDepth = readScalar(md.p, st);
    }
    mark("containment");
    // T0 inner node list
    DBuf<u64> innerList(inner.capacity);
    DBuf<u32> t0Counter(1);
    HIPCHECK(hipMemsetAsync(t0Counter.p, 0, 4, st));
    compactKeys(sc, inner, innerList.p, t0Counter.p, st);
    const u32 t0Count = readScalar(t0Counter.p, st);
    if (t0Count) {   // each inner node's value: the mask of its inner children
        k_zeroVals<<<gridFor(t0Count, BLOCK), BLOCK, 0, st>>>(inner.view(), innerList.p, t0Count);
        k_childMasks<<<gridFor(t0Count, BLOCK), BLOCK, 0, st>>>(inner.view(), innerList.p, t0Count);
    }
    // the inner codes in pinned host memory (any order: the host sets insert every key
    // before they record children in the parents' masks)
    u64 *hInner = (u64 *)pinnedHost(ctx, std::max<size_t>((size_t)t0Count * 8, 8), 1);
    if (t0Count) HIPCHECK(hipMemcpyAsync(hInner, innerList.p, (size_t)t0Count * 8, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    // the replay's host sets are built on host threads while the GPU speculates
    std::future<ReplaySets> setsFuture = std::async(std::launch::async, buildReplaySets, hInner, (size_t)t0Count);

    // ---- sampling refinement: speculative evaluation on T0
    const u32 units = 2 * n;
    DBuf<u32> visitCount(units), offsets(units);
    DBuf<uint8_t> isEvent(units);
    // one traversal per unit: visit counts, creating units, and the visits staged
    DBuf<u64> staged(((size_t)units + 255) / 256 * 256 * VISIT_CAP);
    mark("speculation set-up");
    k_samplingEval<<<gridFor(units, BLOCK), BLOCK, 0, st>>>(f, inner.view(), scene->px.p, scene->py.p, scene->pz.p,
                                                            scene->nx.p, scene->ny.p, scene->nz.p, scene->scale.p,
                                                            maxRel, units, 2, visitCount.p, isEvent.p, nullptr,
                                                            staged.p, nullptr);
    mark("speculation pass");
    // transfer only the part of the index around unsettled units
    DHash hot;
    hot.alloc(std::max<u64>(4096, (u64)units / 2), st);
    k_markHot<<<gridFor(units, BLOCK), BLOCK, 0, st>>>(f, inner.view(), hot.view(), scene->px.p, scene->py.p,
                                                       scene->pz.p, scene->scale.p, maxRel, hotRadius(), isEvent.p,
                                                       units, err.p);
    DBuf<u32> counter(1);
    size_t hotPairs = 0;
    DBuf<u64> hotLeafSorted;
    DBuf<u32> hotUnitSorted;
    {   // hot (leaf, unit) pairs in unit order, then a stable sort by leaf code keeps units
        // ascending per leaf.  Visited leaves are T0 leaves, no deeper than the deepest container.
        DBuf<u64> hotMask(units);
        k_hotVisits<<<gridFor(units, BLOCK), BLOCK, 0, st>>>(hot.view(), staged.p, visitCount.p, isEvent.p, units, 0,
                                                             offsets.p, hotMask.p, nullptr, nullptr, nullptr);
        DBuf<u32> hotOff(units);
        exclusiveScan(sc, offsets.p, hotOff.p, units, st);
        hotPairs = (size_t)readScalar(hotOff.p + units - 1, st) + readScalar(offsets.p + units - 1, st);
        DBuf<u64> hotLeaf(std::max<size_t>(hotPairs, 1));
        DBuf<u32> hotUnit(std::max<size_t>(hotPairs, 1));
        k_hotVisits<<<gridFor(units, BLOCK), BLOCK, 0, st>>>(hot.view(), staged.p, visitCount.p, isEvent.p, units, 1,
                                                             nullptr, hotMask.p, hotOff.p, hotLeaf.p, hotUnit.p);
        hotLeafSorted.alloc(std::max<size_t>(hotPairs, 1));
        hotUnitSorted.alloc(std::max<size_t>(hotPairs, 1));
        if (hotPairs) sortPairs(sc, hotLeaf.p, hotLeafSorted.p, hotUnit.p, hotUnitSorted.p, hotPairs, 0,
                                (int)(3 * maxContainerDepth + 1), st);
    }
    size_t pairs = 0;   // all visits of settled units (statistics, traced runs only)
    if (trace) {
        DBuf<u32> tot(1);
        reduceSum(sc, visitCount.p, tot.p, units, st);
        pairs = readScalar(tot.p, st);
    }
    checkErr(readScalar(err.p, st));
    mark("hot pairs");
    // the event flags, one bit per unit, go to pinned host memory while the index is built:
    // the replay's initial queue
    VisitIndex index;
    const size_t eventWords = ((size_t)units + 63) / 64;
    u64 *hEventBits = (u64 *)pinnedHost(ctx, eventWords * 8, 2);
    {
        DBuf<u64> bits(eventWords);
        k_eventBits<<<gridFor(units, BLOCK), BLOCK, 0, st>>>(isEvent.p, units, bits.p);
        HIPCHECK(hipMemcpyAsync(hEventBits, bits.p, eventWords * 8, hipMemcpyDeviceToHost, st));
    }
    {
        // runs of equal leaves in the sorted hot pairs: unique leaf codes + run offsets + units
        u32 runs = 0;
        DBuf<u64> runLeaf(std::max<size_t>(hotPairs, 1));
        DBuf<u32> runLen(std::max<size_t>(hotPairs, 1)), runOff(std::max<size_t>(hotPairs, 1) + 1);
        if (hotPairs) {
            runLengthEncode(sc, hotLeafSorted.p, runLeaf.p, runLen.p, counter.p, hotPairs, st);
            runs = readScalar(counter.p, st);
            exclusiveScan(sc, runLen.p, runOff.p, runs, st);
        }
        // the hot-leaf -> run map is built on the GPU (run leaves first, then the other hot
        // leaves with INVALID) in the host map's layout, and copied with the run offsets and
        // units through the pinned staging buffer
        DBuf<u64> hotKeys(hot.capacity);
        HIPCHECK(hipMemsetAsync(counter.p + 0, 0, 4, st));
        compactKeys(sc, hot, hotKeys.p, counter.p, st);
        const u32 nh = readScalar(counter.p, st);
        DHash runMap;
        runMap.alloc(std::max<u64>(4096, std::max<u64>(nh, runs)), st);
        if (runs) k_insertIndexed<<<gridFor(runs, BLOCK), BLOCK, 0, st>>>(runMap.view(), runLeaf.p, runs, err.p);
        if (nh) k_insertConst<<<gridFor(nh, BLOCK), BLOCK, 0, st>>>(runMap.view(), hotKeys.p, nh, INVALID, err.p);
        const size_t cap = runMap.capacity;
        const size_t bytes = cap * sizeof(RunSlot) + hotPairs * 4;
        char *pin = (char *)pinnedHost(ctx, bytes + 64);
        RunSlot *hSlots = (RunSlot *)pin;
        u32 *hUnit = (u32 *)(hSlots + cap);
        {   // (key, run) slots packed on the device, then one copy
            DBuf<RunSlot> slots(cap);
            k_packRunSlots<<<gridFor(cap, BLOCK), BLOCK, 0, st>>>(runMap.keys.p, runMap.vals.p, cap, runOff.p,
                                                                 runLen.p, slots.p);
            HIPCHECK(hipMemcpyAsync(hSlots, slots.p, cap * sizeof(RunSlot), hipMemcpyDeviceToHost, st));
            if (runs) HIPCHECK(hipMemcpyAsync(hUnit, hotUnitSorted.p, hotPairs * 4, hipMemcpyDeviceToHost, st));
            HIPCHECK(hipStreamSynchronize(st));
        }
        checkErr(readScalar(err.p, st));
        index.unit = hUnit;
        index.run = RunSlotView{hSlots, cap - 1};
        mark("hot index transfer");
        if (trace) std::fprintf(stderr, "[tsr tree]   hot leaves %u, runs %u\n", nh, runs);
    }
    hotLeafSorted.release();
    hotUnitSorted.release();
    // a visited leaf outside the hot part: select the units whose staged visits include it
    DBuf<uint8_t> unitFlag;
    DBuf<u32> unitIota;
    index.fetch = [&](u64 loc, std::vector<u32> &out) {
        out.clear();
        if (!unitFlag.p) {
            unitFlag.alloc(units);
            unitIota.alloc(units);
            k_iota<<<gridFor(units, BLOCK), BLOCK, 0, st>>>(unitIota.p, units);
        }
        k_matchStaged<<<gridFor(units, BLOCK), BLOCK, 0, st>>>(staged.p, visitCount.p, isEvent.p, units, loc,
                                                               unitFlag.p);
        DBuf<u32> all(units);
        selectFlagged(sc, unitIota.p, unitFlag.p, all.p, counter.p, units, st);
        const u32 m = readScalar(counter.p, st);
        out.resize(m);
        if (m) HIPCHECK(hipMemcpyAsync(out.data(), all.p, (size_t)m * 4, hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
    };
    mark("sampling speculation");
    // The emptiness pass's direction index needs only the samples, their container depths and
    // the cameras: a worker thread builds it on the context's second stream while this thread
    // replays the sampling refinement (the GPU is otherwise idle then).  Its temporaries come
    // from the context's second block pool, which only work on the second stream uses, so no
    // block is shared between the two streams; the pool keeps them for the next build.
    struct EventGuard {   // destroyed after the index future below (declared before it)
        hipEvent_t e = nullptr;
        ~EventGuard() {
            if (e) (void)hipEventDestroy(e);
        }
    } readyGuard;
    HIPCHECK(hipEventCreateWithFlags(&readyGuard.e, hipEventDisableTiming));
    const hipEvent_t samplesReady = readyGuard.e;
    HIPCHECK(hipEventRecord(samplesReady, st));
    const EmptinessLinks inputLinks{scene->px.p, scene->py.p, scene->pz.p, scene->nx.p, scene->ny.p, scene->nz.p,
                                    scene->scale.p, scene->conf.p, scene->cams.p, scene->cx.p, scene->cy.p,
                                    scene->cz.p, cdepth.p, n, scene->maxCams, scene->camCount,
                                    prm.bandwidth_factor, scene->id};
    std::future<std::shared_ptr<EmptinessIndex>> indexFuture =
        std::async(std::launch::async, [ctx, inputLinks, samplesReady]() -> std::shared_ptr<EmptinessIndex> {
            HIPCHECK(hipSetDevice(ctx->device));
            // the context's second pool: only stream2 work takes and returns its blocks
            PoolScope scope(ctx->pool2);
            HIPCHECK(hipStreamWaitEvent(ctx->stream2, samplesReady, 0));
            std::shared_ptr<EmptinessIndex> ix;
            {
                Scratch sc2;
                try {
                    ix = buildEmptinessIndex(ctx->stream2, sc2, inputLinks);
                } catch (const tsr_error &e) {
                    if (e.status != TSR_ERR_OUT_OF_MEMORY) throw;
                    ix.reset();   // the occupancy builds its own
                }
                HIPCHECK(hipStreamSynchronize(ctx->stream2));
            }
            return ix;
        });
    std::vector<u64> created;
    u64 events = 0;
    ReplaySets sets = setsFuture.get();
    mark("replay sets (host, overlapped)");
    replaySampling(f, scene, maxRel, sets, hEventBits, units, index, created, events);
    tree->eIndex = indexFuture.get();
    if (trace)
        std::fprintf(stderr, "[tsr tree] units %u pairs %zu hot pairs %zu t0-inner %u events %llu created nodes %zu "
                     "index misses %llu\n", units, pairs, hotPairs, t0Count, (unsigned long long)events,
                     8 * created.size(), (unsigned long long)index.misses);
    staged.release();
    mark("sampling replay (host)");
    tree->samplingEvents = events;
    tree->samplingCreated = 8 * (u64)created.size();
    tree->samplingUnits = units;
    tree->samplingMisses = index.misses;

    // re-size the inner hash with room for balancing, insert T0 + created
    {
        const u64 need = (u64)t0Count + created.size();
        DHash grown;
        grown.alloc(std::max<u64>(4096, 4 * need), st);
        if (t0Count)
            k_insertList<<<gridFor(t0Count, BLOCK), BLOCK, 0, st>>>(grown.view(), innerList.p, t0Count, err.p);
        if (!created.empty()) {
            DBuf<u64> dc(created.size());
            HIPCHECK(hipMemcpyAsync(dc.p, created.data(), created.size() * 8, hipMemcpyHostToDevice, st));
            k_insertList<<<gridFor(created.size(), BLOCK), BLOCK, 0, st>>>(grown.view(), dc.p, (u32)created.size(), err.p);
            HIPCHECK(hipStreamSynchronize(st));
        }
        inner = std::move(grown);
    }

    mark("inner set rebuild");
    // ---- balancing fixpoint: all inner nodes once, then only the nodes each pass added
    {
        const u32 cap = (u32)std::min<u64>(inner.capacity, 0xffffffffull);
        DBuf<u64> listA(cap), listB(cap);
        compactKeys(sc, inner, listA.p, counter.p, st);
        u32 n = readScalar(counter.p, st);
        for (int iter = 0; n; ++iter) {
            if (iter > 64) throw tsr_error(TSR_ERR_HIP, "balancing did not converge");
            HIPCHECK(hipMemsetAsync(counter.p, 0, 4, st));
            k_balance<<<gridFor(n, BLOCK), BLOCK, 0, st>>>(inner.view(), listA.p, n, listB.p, counter.p, cap, err.p);
            n = readScalar(counter.p, st);
            checkErr(readScalar(err.p, st));
            std::swap(listA, listB);
        }
    }

    mark("balancing");
    // ---- final inner list sorted in pre-order
    innerList.alloc(inner.capacity);
    HIPCHECK(hipMemsetAsync(counter.p, 0, 4, st));
    compactKeys(sc, inner, innerList.p, counter.p, st);
    const u32 I = readScalar(counter.p, st);
    if ((u64)1 + 8ull * I >= (u64)INVALID) throw tsr_error(TSR_ERR_NODE_OVERFLOW, "Required scene tree node count is larger than supported.");
    DBuf<u64> keys(std::max<u32>(I, 1)), keysSorted(std::max<u32>(I, 1)), innerSorted(std::max<u32>(I, 1));
    if (I) {
        k_dfsKeys<<<gridFor(I, BLOCK), BLOCK, 0, st>>>(innerList.p, I, keys.p);
        sortPairs(sc, keys.p, keysSorted.p, innerList.p, innerSorted.p, I, 0, 64, st);
    }
    const u32 N = 1 + 8 * I;
    tree->innerCount = I;
    tree->nodeCount = N;
    if (I) k_innerRanks<<<gridFor(I, BLOCK), BLOCK, 0, st>>>(inner.view(), innerSorted.p, I);
    tree->child.alloc(N);
    tree->parents.alloc(N);
    tree->nodeDepth.alloc(N);
    tree->nodeLoc.alloc(N);
    tree->nodeToLeaf.alloc(N);
    k_fillU32<<<gridFor(N, BLOCK), BLOCK, 0, st>>>(tree->nodeToLeaf.p, N, INVALID);
    const u32 Lmax = I ? 7 * I + 1 : 1;
    DBuf<u64> leafKeys(Lmax), leafKeysSorted(Lmax);
    DBuf<u32> leafIds(Lmax), leafIdsSorted(Lmax);
    HIPCHECK(hipMemsetAsync(counter.p, 0, 4, st));
    k_nodeArrays<<<gridFor(std::max<u32>(I, 1), BLOCK), BLOCK, 0, st>>>(
        inner.view(), innerSorted.p, I, tree->child.p, tree->parents.p, tree->nodeDepth.p,
        tree->nodeLoc.p, leafKeys.p, leafIds.p, counter.p);
    const u32 L = readScalar(counter.p, st);
    checkErr(readScalar(err.p, st));
    sortPairs(sc, leafKeys.p, leafKeysSorted.p, leafIds.p, leafIdsSorted.p, L, 0, 64, st);
    tree->leafCount = L;
    tree->leafLoc.alloc(L);
    tree->leafNode.alloc(L);
    k_leafArrays<<<gridFor(L, BLOCK), BLOCK, 0, st>>>(leafIdsSorted.p, L, tree->nodeLoc.p, tree->child.p,
                                                      tree->nodeToLeaf.p, tree->leafLoc.p, tree->leafNode.p);
    {   // max depth
        DBuf<uint8_t> md(1);
        reduceMax(sc, tree->nodeDepth.p, md.p, N, st);
        tree->maxDepth = readScalar(md.p, st);
    }

    mark("node/leaf tables");
    // ---- samples: node ids, counts, starts, stable order by node id
    DBuf<u32> nodeOf(n), orig(n), nodeOfSorted(n), origSorted(n);
    tree->sampleCounts.alloc(N);
    tree->sampleStarts.alloc(N);
    HIPCHECK(hipMemsetAsync(tree->sampleCounts.p, 0, (size_t)N * 4, st));
    k_sampleNodes<<<gridFor(n, BLOCK), BLOCK, 0, st>>>(inner.view(), containerLoc.p, n, nodeOf.p,
                                                       tree->sampleCounts.p, err.p);
    checkErr(readScalar(err.p, st));
    exclusiveScan(sc, tree->sampleCounts.p, tree->sampleStarts.p, N, st);
    k_iota<<<gridFor(n, BLOCK), BLOCK, 0, st>>>(orig.p, n);
    int bits = 1;
    while (bits < 32 && (1ull << bits) < N) ++bits;
    sortPairs(sc, nodeOf.p, nodeOfSorted.p, orig.p, origSorted.p, n, 0, bits, st);
    tree->sampleTarget.alloc(n);
    k_scatterTarget<<<gridFor(n, BLOCK), BLOCK, 0, st>>>(origSorted.p, n, tree->sampleTarget.p);
    // the samples in node order (Samples::reorder); the scene keeps the input order
    auto permuteF = [&](const DBuf<float> &src, DBuf<float> &dst) {
        dst.alloc(n);
        k_gather<float><<<gridFor(n, BLOCK), BLOCK, 0, st>>>(src.p, origSorted.p, n, 1, dst.p);
    };
    permuteF(scene->px, tree->px); permuteF(scene->py, tree->py); permuteF(scene->pz, tree->pz);
    permuteF(scene->nx, tree->nx); permuteF(scene->ny, tree->ny); permuteF(scene->nz, tree->nz);
    permuteF(scene->scale, tree->scale); permuteF(scene->conf, tree->conf);
    tree->maxCams = scene->maxCams;
    tree->cams.alloc((size_t)n * scene->maxCams);
    k_gather<u32><<<gridFor(n, BLOCK), BLOCK, 0, st>>>(scene->cams.p, origSorted.p, n, scene->maxCams, tree->cams.p);
    tree->sampleDepth.alloc(n);
    k_gather<uint8_t><<<gridFor(n, BLOCK), BLOCK, 0, st>>>(cdepth.p, origSorted.p, n, 1, tree->sampleDepth.p);

    mark("sample reorder");
    // ---- dual cells
    if (L > 0) {
        const size_t corners = 8 * (size_t)L;
        DBuf<u64> codes(corners), codesSorted(corners), uniq(corners);
        const u32 D = std::max<u32>(tree->maxDepth, 1);
        k_cornerCodes<<<gridFor(L, BLOCK), BLOCK, 0, st>>>(tree->leafLoc.p, L, D, codes.p);
        sortKeys(sc, codes.p, codesSorted.p, corners, 0, (int)(3 * D + 1), st);
        uniqueKeys(sc, codesSorted.p, uniq.p, counter.p, corners, st);
        u32 V = readScalar(counter.p, st);
        if (V > 0) {
            const u64 last = readScalar(uniq.p + V - 1, st);
            if (last == ~0ull) --V;   // drop the boundary marker
        }
        codes.release(); codesSorted.release();
        tree->cellCount = V;
        if (V > 0) {
            DBuf<u32> cellsUnsorted(8 * (size_t)V), idx(V), idxA(V), idxB(V);
            DBuf<u64> kh(V), kl(V), khS(V), klS(V);
            // lexicographic (hi, lo): stable sort by lo, then by hi.  A vertex on the grid of
            // depth D has a path of at most D levels of 5 bits: levels 0-11 at hi bits
            // 4 + 5 (11 - lvl), levels 12-18 at lo bits 4 + 5 (18 - lvl); the key halves are
            // shifted down past the bits no level < D sets, and the sorts cover the rest.
            const u32 hiShift = D > 12 ? 4u : 4u + 5u * (12u - D), loShift = D > 12 ? 4u + 5u * (19u - D) : 39u;
            k_cells<<<gridFor(V, BLOCK), BLOCK, 0, st>>>(inner.view(), tree->nodeToLeaf.p,
                                                         uniq.p, V, D, hiShift, loShift, cellsUnsorted.p, kh.p,
                                                         kl.p, idx.p);
            if (D > 12) {
                sortPairs(sc, kl.p, klS.p, idx.p, idxA.p, V, 0, (int)(39 - loShift), st);
                k_gather<u64><<<gridFor(V, BLOCK), BLOCK, 0, st>>>(kh.p, idxA.p, V, 1, khS.p);
                sortPairs(sc, khS.p, kh.p, idxA.p, idxB.p, V, 0, (int)(64 - hiShift), st);
            } else {   // lo is zero
                sortPairs(sc, kh.p, khS.p, idx.p, idxB.p, V, 0, (int)(64 - hiShift), st);
            }
            tree->cells.alloc(8 * (size_t)V);
            k_gather<u32><<<gridFor(V, BLOCK), BLOCK, 0, st>>>(cellsUnsorted.p, idxB.p, V, 8, tree->cells.p);
            k_checkCells<<<gridFor(8 * (size_t)V, BLOCK), BLOCK, 0, st>>>(tree->cells.p, 8 * (size_t)V, err.p);
        }
    }
    HIPCHECK(hipStreamSynchronize(st));
    checkErr(readScalar(err.p, st));
    mark("dual cells");
}

}  // namespace tsr
