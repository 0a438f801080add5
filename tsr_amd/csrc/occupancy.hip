// Occupancy scalar field on the GPU (Occupancy::Occupancy(const Tree*),
// SR/SurfaceExtraction/Occupancy.cpp:78-127).
//
//   S pass   Occupancy.cpp:138-248 (SAMPLENESS)   one lane per camera->sample link; stackless
//                                                 DFS over the node arrays (first child / parent
//                                                 links), FP32 cone culling with a safety margin,
//                                                 FP64 kernel evaluation at leaves
//   flags    Occupancy.cpp:491-516                leaf S flags from the reduced hits, OR up the tree
//   E pass   Occupancy.cpp:138-248 (EMPTINESS)    same traversal, gated by the S flag
//   priors   Occupancy.cpp:530-568                per-leaf sphere query with subtree sums for nodes
//                                                 entirely inside the sphere
//   ratio    Occupancy.cpp:570-587                FP64 occupancy per leaf
//
// Accumulation (default): every accepted (link, leaf) contribution is written
// as a record in link order (count pass + scan + write pass), records are
// stably radix-sorted by leaf, and each leaf's records are summed in FP32 with
// Kahan compensation in that fixed order: deterministic, no float atomics.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "handles.h"
#include "link_ranges.h"

namespace tsr {

namespace {

constexpr u32 BLOCK = 256;
constexpr u32 XCD_COUNT = 8;

inline unsigned xcdGrid(size_t n, unsigned block) {
    const unsigned g = gridFor(n, block);
    return (g + XCD_COUNT - 1) / XCD_COUNT * XCD_COUNT;
}

// staging slots per link of the single-traversal record pass (as many as a memory
// budget allows, at most CONE_SLOTS_MAX); links with more accepted hits are traversed
// again (k_coneLinkWave) into their final place
constexpr u32 CONE_SLOTS_MAX = 512;
constexpr size_t CONE_STAGING_BUDGET = (size_t)64 << 30;   // bytes
constexpr u64 CONE_MAX_RECORDS = 1536ull << 20;              // records of one link range (one sort)
// links per chunk of a cone pass (TSR_CONE_CHUNK overrides, e.g. to exercise the
// multi-chunk merge on small scenes)
inline u32 coneChunkLinks() {   // read at every occupancy computation
    const char *e = std::getenv("TSR_CONE_CHUNK");
    const long v = e ? std::atol(e) : 0;
    return v > 0 ? (u32)v : (1u << 24);
}

// Staging layout of the record pass: the records of link position idx are contiguous,
// record j at idx * slots + j, so that the gather into link order reads and writes whole runs.
constexpr u32 STAGE_GROUP = 256;   // positions per gather block (the staging area is padded to it)
__device__ __forceinline__ size_t stagingIndex(u32 idx, u32 j, u32 slots) { return (size_t)idx * slots + j; }

// Traversal copies of the node arrays store node i at index i + NODE_PAD: the 8
// children of the r-th inner node (ids 1 + 8r .. 8 + 8r) then start on a 32-byte
// (child links) or 64-byte ((child, state) pairs) boundary and a sibling scan reads
// one cache line.  Parents are kept per child block: blockParent[(n - 1) / 8] is the
// parent of node n (8 times fewer entries than a per-node table).
constexpr u32 NODE_PAD = 7;   // PriorRecords (handles.h) uses the same padding

struct TraversalTree {
    const u32 *child;        // first child id, or LEAF_BIT | leaf index (node-padded copy)
    const u32 *blockParent;  // parent per child block
    const u32 *states;       // node state flags (E pass gate), may be null
    Frame f;
    float margin;            // absolute culling margin
    const uint2 *packed;     // (child, state) per node for gated passes, node-padded: one load per visit; may be null
};

__global__ void k_packChildState(const u32 *__restrict__ child, const u32 *__restrict__ states, u32 n,
                                 uint2 *__restrict__ out) {
    const u32 i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) out[i] = make_uint2(child[i], states[i]);
}

__global__ void k_traversalLayout(const u32 *__restrict__ child, const u32 *__restrict__ parents, u32 n,
                                  u32 *__restrict__ childOut, u32 *__restrict__ blockParent) {
    const u32 i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    childOut[i] = child[i];
    if (i >= 1 && ((i - 1) & 7u) == 0) blockParent[(i - 1) >> 3] = parents[i];
}

// node-padded child links and block parents of `tree` into the given buffers
void traversalLayout(hipStream_t st, const tsr_tree *tree, DBuf<u32> &childPad, DBuf<u32> &blockParent,
                     TraversalTree &tt) {
    const u32 N = tree->nodeCount;
    u32 *c = childPad.ensure((size_t)N + NODE_PAD);
    u32 *bp = blockParent.ensure((size_t)N / 8 + 1);
    k_traversalLayout<<<gridFor(N, 256), 256, 0, st>>>(tree->child.p, tree->parents.p, N, c + NODE_PAD, bp);
    tt.child = c + NODE_PAD;
    tt.blockParent = bp;
}

struct LinkInputs {
    const float *px, *py, *pz, *nx, *ny, *nz, *scale, *conf;
    const u32 *cams;
    const float *cx, *cy, *cz;
    const uint8_t *sampleDepth;
    u32 maxCams, camCount;
    float bandwidth, relConeLength;
    u32 linkStride;            // process every linkStride-th link (1 = all)
    const u32 *order;          // thread index -> link (links grouped by camera), may be null
    u32 slots;                 // record pass: staging slots per link (layout: stagingIndex)
    u32 linkBase;              // without `order`: thread index i -> link linkBase + i
};

// Cone of link `link` (Occupancy.cpp:174-201, getCameraToSampleLinkData :250-277)
__device__ bool setupLink(const LinkInputs &in, u32 link, bool sampleness, ConeD &cone, u32 &sample) {
    const u32 i = link / in.maxCams;
    const u32 cam = in.cams[link];
    sample = i;
    if (cam >= in.camCount) return false;   // Cameras::isValid
    const double apex[3] = {in.cx[cam], in.cy[cam], in.cz[cam]};
    const double p[3] = {in.px[i], in.py[i], in.pz[i]};
    const double nrm[3] = {-(double)in.nx[i], -(double)in.ny[i], -(double)in.nz[i]};
    const double R = (double)in.bandwidth * (double)in.scale[i];
    double L = 3.4028234663852886e38;   // REAL_MAX
    double end[3] = {p[0], p[1], p[2]};
    if (sampleness) {
        const double cs[3] = {p[0] - apex[0], p[1] - apex[1], p[2] - apex[2]};
        const double total = sqrt(cs[0] * cs[0] + cs[1] * cs[1] + cs[2] * cs[2]);
        L = R * (double)in.relConeLength;
        for (int a = 0; a < 3; ++a) end[a] += (cs[a] / total) * (L * 0.5);
    }
    coneInit(cone, apex, nrm, end, R, L);
    return !cone.degenerate;
}

// Depth-first traversal (NodesIterator.cpp:38-86 / LeavesIterator.cpp:24-31):
// calls onLeaf(leafIndex, centre) for every leaf that passes the depth, flag and
// cone tests.  Inner nodes failing a test are skipped with their subtrees.  With
// `path` (this thread's column of an LDS array [MAX_DEPTH + 2][BLOCK]) the ancestors
// are kept there on the way down and climbing reads them back; without it, climbing
// reads the block-parent table.
template <typename OnLeaf>
__device__ void traverseCone(const TraversalTree &t, const ConeCull &cull, u32 maxDepth, u32 required,
                             u32 &visits, u32 *path, OnLeaf onLeaf) {
    u32 node = 0, d = 0, k[3] = {0, 0, 0};
    bool descend = true;   // whether to test `node` (true) or move on from it (false)
    for (;;) {
        if (descend) {
            ++visits;
            bool accept = d <= maxDepth;
            u32 c = 0u;
            if (t.packed) {
                const uint2 cs = t.packed[node];
                if (accept && required) accept = (cs.y & required) == required;
                c = cs.x;
            } else if (accept && required) {
                accept = (t.states[node] & required) == required;
            }
            if (!t.packed) c = accept ? t.child[node] : 0u;
            if (accept) {
                if (c & LEAF_BIT) {
                    double x[3];
                    for (int a = 0; a < 3; ++a) x[a] = t.f.B(a, 2ull * k[a] + 1, d + 1);
                    onLeaf(c & ~LEAF_BIT, x);
                } else if (c != INVALID) {
                    const float h = ldexpf(t.f.rootSizeF, -(int)d);
                    const float lo0 = t.f.rootMinF[0] + (float)k[0] * h, lo1 = t.f.rootMinF[1] + (float)k[1] * h,
                                lo2 = t.f.rootMinF[2] + (float)k[2] * h;
                    if (!cullSeparated(cull, lo0, lo1, lo2, h, t.margin)) {
                        if (path) path[(d + 1) * BLOCK] = node;
                        node = c;   // first child
                        ++d;
                        k[0] *= 2; k[1] *= 2; k[2] *= 2;
                        continue;
                    }
                }
            }
        }
        // advance: next sibling, or climb
        for (;;) {
            if (node == 0) return;
            const u32 slot = (node - 1) & 7;
            if (slot < 7) {
                ++node;
                const u32 s1 = slot + 1;
                k[0] = (k[0] & ~1u) | (s1 & 1);
                k[1] = (k[1] & ~1u) | ((s1 >> 1) & 1);
                k[2] = (k[2] & ~1u) | ((s1 >> 2) & 1);
                descend = true;
                break;
            }
            node = path ? path[d * BLOCK] : t.blockParent[(node - 1) >> 3];
            --d;
            k[0] >>= 1; k[1] >>= 1; k[2] >>= 1;
        }
    }
}

__device__ void coneLink(const TraversalTree &t, const LinkInputs &in, u32 link, u32 idx, u32 linkCount,
                         int sampleness, int mode,
                         u32 *__restrict__ counts, float *__restrict__ atomicSum, float *__restrict__ atomicLen,
                         u32 *__restrict__ atomicStates, double *__restrict__ confOut, u32 &visits, u32 &count,
                         u32 *path);

// Deepest node whose subtree holds every leaf centre the cone can reach: a centre with a
// non-zero kernel lies in a ball B(a + v t, R t) with t in [tStart, 1], so inside the box
// of the two end balls; the descent follows the child whose half contains that box on
// every axis (a lower child's centres are < the splitting plane, an upper child's >) and
// stops where the box straddles a plane or at a leaf.  Any node above it that the
// traversal would accept contains the same contributing centres, and the culling never
// rejects a node holding one, so a traversal started here meets the same leaves.
__device__ void coneStart(const TraversalTree &t, const ConeD &c, u32 &node, u32 &d, u32 &k0, u32 &k1, u32 &k2) {
    double lo[3], hi[3];
    const double m = t.f.rootSize * 1e-9;   // far above the FP64 rounding of the kernel's tests
    for (int a = 0; a < 3; ++a) {
        const double c1 = c.a[a] + c.v[a] * c.tStart, r1 = c.R * c.tStart;
        const double c2 = c.a[a] + c.v[a], r2 = c.R;
        lo[a] = fmin(c1 - r1, c2 - r2) - m;
        hi[a] = fmax(c1 + r1, c2 + r2) + m;
    }
    node = 0;
    d = 0;
    u32 k[3] = {0, 0, 0};
    for (;;) {
        const u32 ch = t.packed ? t.packed[node].x : t.child[node];
        if ((ch & LEAF_BIT) || ch == INVALID) break;
        u32 slot = 0;
        bool straddles = false;
        for (int a = 0; a < 3; ++a) {
            const double mid = t.f.B(a, 2u * k[a] + 1u, d + 1);
            if (lo[a] >= mid) slot |= 1u << a;
            else if (!(hi[a] < mid)) straddles = true;
        }
        if (straddles) break;
        node = ch + slot;
        ++d;
        for (int a = 0; a < 3; ++a) k[a] = 2 * k[a] + ((slot >> a) & 1u);
    }
    k0 = k[0];
    k1 = k[1];
    k2 = k[2];
}

// set bits of the wave-wide mask m below this lane (v_mbcnt: no per-lane mask register)
__device__ __forceinline__ u32 lanesBelow(u64 m) {
    return __builtin_amdgcn_mbcnt_hi((u32)(m >> 32), __builtin_amdgcn_mbcnt_lo((u32)m, 0u));
}

// wave sum of a per-lane count, added once per wave to *acc
__device__ void waveAccumulate(unsigned long long *acc, u32 v) {
    unsigned long long x = v;
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
    if ((threadIdx.x & 63) == 0) atomicAdd(acc, x);
}

// mode 0: count accepted hits per link (with the confidence of each link); 2: FP32 atomics
// into the leaf sums (precision study)
// stats[0] += node visits, stats[1] += accepted hits
// Register use capped for 6 waves per SIMD (80 VGPRs, no spill).
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6))) void k_conePass(TraversalTree t, LinkInputs in, u32 linkCount, int sampleness, int mode,
                           u32 *__restrict__ counts, float *__restrict__ atomicSum, float *__restrict__ atomicLen,
                           u32 *__restrict__ atomicStates, double *__restrict__ confOut,
                           unsigned long long *__restrict__ stats) {
    __shared__ u32 pathL[MAX_DEPTH + 2][BLOCK];   // ancestors per thread (launched with BLOCK threads)
    const u32 idx = (blockIdx.x * blockDim.x + threadIdx.x) * in.linkStride;
    u32 visits = 0, count = 0;
    const u32 link = idx < linkCount && in.order ? in.order[idx] : in.linkBase + idx;
    if (idx < linkCount) coneLink(t, in, link, idx, linkCount, sampleness, mode, counts, atomicSum, atomicLen,
                                   atomicStates, confOut, visits, count, &pathL[0][threadIdx.x]);
    if (stats) {
        waveAccumulate(&stats[0], visits);
        waveAccumulate(&stats[1], count);
    }
}

__device__ void coneLink(const TraversalTree &t, const LinkInputs &in, u32 link, u32 idx, u32 linkCount,
                         int sampleness, int mode,
                         u32 *__restrict__ counts, float *__restrict__ atomicSum, float *__restrict__ atomicLen,
                         u32 *__restrict__ atomicStates, double *__restrict__ confOut, u32 &visits, u32 &count,
                         u32 *path) {
    ConeD cone;
    u32 i;
    const bool ok = setupLink(in, link, sampleness != 0, cone, i);
    if (mode == 0 && confOut) confOut[link] = ok ? (double)in.conf[i] : 0.0;
    if (!ok) {
        if (mode == 0) counts[link] = 0;
        return;
    }
    ConeCull cull;
    cullInit(cull, cone);
    const double conf = in.conf[i];
    const float weightedLength = (float)(conf * cone.length);
    const u32 maxDepth = (u32)in.sampleDepth[i] + MAX_DEPTH_DIFFERENCE;
    const u32 required = sampleness ? 0u : FLAG_S;
    traverseCone(t, cull, maxDepth, required, visits, path,
                 [&](u32 leaf, const double x[3]) {
        const double kv = coneKernel(cone, x, sampleness != 0, conf);
        if (kv == 0.0) return;
        if (mode == 2) {
            atomicAdd(&atomicSum[leaf], (float)kv);
            atomicAdd(&atomicLen[leaf], weightedLength);
            atomicOr(&atomicStates[leaf], 1u);
        }
        ++count;
    });
    if (mode == 0) counts[link] = count;
}

// The record pass with persistent threads (after Aila & Laine, "Understanding the
// efficiency of ray traversal on GPUs", 2009): every lane runs the traversal of
// traverseCone as an explicit state machine, from the deepest node enclosing its cone
// (coneStart) instead of the root, and, when its link is done, takes the next
// link index from a shared counter (one atomic per wavefront for all idle lanes).  Lanes
// no longer idle at the end of a wave's longest traversal.  Per link it counts the
// accepted hits and stages the first in.slots records (stagingIndex).
//
// With `slices` = 8 the link list is cut into 8 contiguous slices with a counter
// each (nextIdx[0..7]); a wave starts on the slice of its XCD (blocks are dispatched
// round-robin over the 8 XCDs) and moves on to the next slice when its own is used
// up.  Each XCD's L2 then serves cones of a few cameras towards one part of the
// scene instead of the whole GPU's window of links.
template <bool SAMPLENESS>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_conePassPT(TraversalTree t, LinkInputs in, u32 linkCount,
                                                    u32 *__restrict__ counts, u32 *__restrict__ stLeaf,
                                                    uint2 *__restrict__ stVal, double *__restrict__ confOut,
                                                    u32 *__restrict__ nextIdx, u32 slices,
                                                    unsigned long long *__restrict__ stats) {
    // ancestors of the current node per lane: pathL[d][x] = parent of thread x's node at depth d
    // (climbing reads LDS instead of the block-parent table)
    __shared__ u32 pathL[MAX_DEPTH + 2][256];
    const u32 lane = threadIdx.x & 63;
    const u32 sliceLen = (linkCount + slices - 1) / slices;
    u32 cur = blockIdx.x % slices, tried = 0;   // wave-uniform
    constexpr u32 required = SAMPLENESS ? 0u : FLAG_S;
    bool active = false, more = true;
    u32 idx = 0, link = 0, count = 0, maxDepth = 0, node = 0, d = 0, k0 = 0, k1 = 0, k2 = 0, start = 0;
    bool descend = true;
    ConeD cone;
    ConeCull cull;
    double conf = 0.0;
    float wl = 0.f;
    u32 visits = 0, hits = 0;
    for (;;) {
        const u64 idle = __ballot(!active && more);
        // refill the idle lanes with consecutive link indices of the current slice, once at
        // least 4 are idle (or none is active): a refill costs the wavefront the cone set-up
        // and a copy of the lanes' traversal state, so fewer, larger refills pay
        if (idle && (__popcll(idle) >= 4 || !__ballot(active))) {
            const int leader = __ffsll((long long)idle) - 1;
            u32 base = 0;
            if ((int)lane == leader) base = atomicAdd(&nextIdx[cur], (u32)__popcll(idle));
            base = __shfl(base, leader);
            const u32 sBegin = min(linkCount, cur * sliceLen), sEnd = min(linkCount, sBegin + sliceLen);
            if (!__ballot(!active && more && sBegin + base + lanesBelow(idle) < sEnd)) {
                // this slice is used up: move on; done after the last one
                cur = cur + 1 == slices ? 0u : cur + 1;
                if (++tried == slices) more = false;
                continue;
            }
            if (!active && more) {
                idx = sBegin + base + lanesBelow(idle);
                if (idx >= sEnd) {
                    // past the slice end: stays idle and refills from the next slice
                } else {
                    link = in.order ? in.order[idx] : in.linkBase + idx;
                    u32 i;
                    const bool ok = setupLink(in, link, SAMPLENESS, cone, i);
                    if (!SAMPLENESS) confOut[link] = ok ? (double)in.conf[i] : 0.0;
                    if (!ok) {
                        counts[link] = 0;
                    } else {
                        cullInit(cull, cone);
                        conf = in.conf[i];
                        wl = (float)(conf * cone.length);
                        maxDepth = (u32)in.sampleDepth[i] + MAX_DEPTH_DIFFERENCE;
                        count = 0;
                        coneStart(t, cone, node, d, k0, k1, k2);
                        start = node;
                        descend = true;
                        active = true;
                    }
                }
            }
        }
        if (!__ballot(active)) {
            if (!__ballot(more)) break;
            continue;
        }
        if (!active) continue;
        // one step of traverseCone
        bool advance = true;
        if (descend) {
            ++visits;
            bool accept = d <= maxDepth;
            u32 c = 0u;
            if (t.packed) {
                const uint2 cs = t.packed[node];
                if (accept && required) accept = (cs.y & required) == required;
                c = cs.x;
            } else {
                if (accept && required) accept = (t.states[node] & required) == required;
                c = t.child[node];
            }
            if (accept) {
                if (c & LEAF_BIT) {
                    const u32 k[3] = {k0, k1, k2};
                    double x[3];
                    for (int a = 0; a < 3; ++a) x[a] = t.f.B(a, 2ull * k[a] + 1, d + 1);
                    const double kv = coneKernel(cone, x, SAMPLENESS, conf);
                    if (kv != 0.0) {
                        if (count < in.slots) {
                            const size_t at = stagingIndex(idx, count, in.slots);
                            stLeaf[at] = c & ~LEAF_BIT;
                            stVal[at] = make_uint2(__float_as_uint((float)kv), __float_as_uint(wl));
                        }
                        ++count;
                    }
                } else if (c != INVALID) {
                    const float h = ldexpf(t.f.rootSizeF, -(int)d);
                    const float lo0 = t.f.rootMinF[0] + (float)k0 * h, lo1 = t.f.rootMinF[1] + (float)k1 * h,
                                lo2 = t.f.rootMinF[2] + (float)k2 * h;
                    if (!cullSeparated(cull, lo0, lo1, lo2, h, t.margin)) {
                        pathL[d + 1][threadIdx.x] = node;
                        node = c;   // first child
                        ++d;
                        k0 *= 2; k1 *= 2; k2 *= 2;
                        advance = false;
                    }
                }
            }
        }
        if (advance) {   // next sibling, or climb; done when back at the start node
            for (;;) {
                if (node == start) {
                    counts[link] = count;
                    hits += count;
                    active = false;
                    break;
                }
                const u32 slot = (node - 1) & 7;
                if (slot < 7) {
                    ++node;
                    const u32 s1 = slot + 1;
                    k0 = (k0 & ~1u) | (s1 & 1);
                    k1 = (k1 & ~1u) | ((s1 >> 1) & 1);
                    k2 = (k2 & ~1u) | ((s1 >> 2) & 1);
                    descend = true;
                    break;
                }
                node = pathL[d][threadIdx.x];
                --d;
                k0 >>= 1; k1 >>= 1; k2 >>= 1;
            }
        } else {
            descend = true;
        }
    }
    if (stats) {
        waveAccumulate(&stats[0], visits);
        waveAccumulate(&stats[1], hits);
    }
}

// Re-traversal of one heavy link per wavefront (the links with more records than
// staging slots): an LDS stack of inner nodes still to open; each step pops up to 8
// and tests their 8 children on the 64 lanes with traverseCone's per-node rules, and
// evaluates the kernel at accepted leaves.  A link meets every leaf at most once, so
// its records can be written in any order within its segment [offsets[link], +count):
// the per-leaf order after the stable sort by leaf is still the link order.
constexpr u32 CONE_WAVE_STACK = 1344;   // >= 64 * (MAX_DEPTH + 1)
__global__ __launch_bounds__(64) void k_coneLinkWave(TraversalTree t, LinkInputs in, const u32 *__restrict__ links,
                                                     u32 count, int sampleness, const u64 *__restrict__ nodeLoc,
                                                     const u32 *__restrict__ offsets, const u32 *__restrict__ counts,
                                                     u32 *__restrict__ recLeaf, uint2 *__restrict__ recVal,
                                                     u32 *__restrict__ err, unsigned long long *__restrict__ stats) {
    __shared__ u32 stack[CONE_WAVE_STACK];
    __shared__ u32 sp, written;
    const u32 lane = threadIdx.x;
    const u32 required = sampleness ? 0u : FLAG_S;
    u32 visits = 0;
    for (u32 w = blockIdx.x; w < count; w += gridDim.x) {
        const u32 link = links[w];
        ConeD cone;
        u32 i;
        const bool ok = setupLink(in, link, sampleness != 0, cone, i);
        if (!ok) continue;   // (no records: such links are never listed)
        ConeCull cull;
        cullInit(cull, cone);
        const double conf = in.conf[i];
        const float wl = (float)(conf * cone.length);
        const u32 maxDepth = (u32)in.sampleDepth[i] + MAX_DEPTH_DIFFERENCE;
        const size_t base = offsets[link];
        // node accept rules of traverseCone
        auto accepts = [&](u32 node, u32 d, u32 &c) {
            bool accept = d <= maxDepth;
            if (accept && required) accept = (t.states[node] & required) == required;
            c = t.child[node];
            return accept;
        };
        auto leafRecord = [&](u32 c, u64 loc) {
            u32 k[3];
            locCoords(loc, k[0], k[1], k[2]);
            const u32 d = locDepth(loc);
            double x[3];
            for (int a = 0; a < 3; ++a) x[a] = t.f.B(a, 2ull * k[a] + 1, d + 1);
            const double kv = coneKernel(cone, x, sampleness != 0, conf);
            if (kv != 0.0) {
                const u32 pos = atomicAdd(&written, 1u);
                recLeaf[base + pos] = c & ~LEAF_BIT;
                recVal[base + pos] = make_uint2(__float_as_uint((float)kv), __float_as_uint(wl));
            }
        };
        auto opens = [&](u32 c, u64 loc) {
            u32 k[3];
            locCoords(loc, k[0], k[1], k[2]);
            const u32 d = locDepth(loc);
            const float h = ldexpf(t.f.rootSizeF, -(int)d);
            return !cullSeparated(cull, t.f.rootMinF[0] + (float)k[0] * h, t.f.rootMinF[1] + (float)k[1] * h,
                                  t.f.rootMinF[2] + (float)k[2] * h, h, t.margin);
        };
        if (lane == 0) {
            sp = 0;
            written = 0;
            ++visits;
            u32 c;
            if (accepts(0, 0, c)) {
                if (c & LEAF_BIT) leafRecord(c, 1ull);
                else if (c != INVALID && opens(c, 1ull)) {
                    stack[0] = 0;
                    sp = 1;
                }
            }
        }
        __syncthreads();
        for (;;) {
            const u32 top = sp;
            if (top == 0) break;
            const u32 take = min(top, 8u);
            const u32 b = top - take;
            const u32 e = lane >> 3, slot = lane & 7;
            bool push = false;
            u32 cid = 0;
            if (e < take) {
                const u32 parent = stack[b + e];
                cid = t.child[parent] + slot;
                const u64 loc = (nodeLoc[parent] << 3) | slot;
                ++visits;
                u32 c;
                if (accepts(cid, locDepth(loc), c)) {
                    if (c & LEAF_BIT) leafRecord(c, loc);
                    else if (c != INVALID) push = opens(c, loc);
                }
            }
            __syncthreads();   // every lane has read its entry before the slots are reused
            const u64 mask = __ballot(push);
            const u32 pos = lanesBelow(mask);
            const u32 newTop = b + (u32)__popcll(mask);
            if (newTop > CONE_WAVE_STACK) {
                if (lane == 0) { atomicOr(err, 1u); sp = 0; }
            } else {
                if (push) stack[b + pos] = cid;
                if (lane == 0) sp = newTop;
            }
            __syncthreads();
        }
        if (lane == 0 && written != counts[link]) atomicOr(err, 2u);   // must match the counting pass
        __syncthreads();
    }
    if (stats) waveAccumulate(&stats[0], visits);
}

// staged records of the links that fit their slots -> link order (offsets = exclusive
// scan of the counts); a link's records stay in traversal order.  A block of STAGE_GROUP
// threads moves the records of STAGE_GROUP consecutive positions together: after a
// block-wide scan of their counts, thread t copies records t, t + STAGE_GROUP, ... of the
// block's concatenated runs, so consecutive threads read and write consecutive records.
__global__ __launch_bounds__(STAGE_GROUP) void k_gatherSlots(const u32 *__restrict__ counts,
                                                             const u32 *__restrict__ offsets,
                                                             const u32 *__restrict__ order, u32 base, u32 links,
                                                             u32 slots, const u32 *__restrict__ stLeaf,
                                                             const uint2 *__restrict__ stVal, u32 *__restrict__ recLeaf,
                                                             uint2 *__restrict__ recVal) {
    __shared__ u32 pre[STAGE_GROUP + 1], dst[STAGE_GROUP];
    const u32 t = threadIdx.x, i = blockIdx.x * STAGE_GROUP + t;
    u32 c = 0;
    if (i < links) {
        const u32 l = order ? order[i] : base + i;
        c = counts[l];
        if (c > slots) c = 0;   // re-traversed into place
        dst[t] = offsets[l];
    }
    pre[t + 1] = c;
    if (t == 0) pre[0] = 0;
    __syncthreads();
    for (u32 w = 1; w < STAGE_GROUP; w <<= 1) {   // inclusive scan of pre[1..]
        const u32 v = t + 1 > w ? pre[t + 1 - w] : 0u;
        __syncthreads();
        pre[t + 1] += v;
        __syncthreads();
    }
    const u32 total = pre[STAGE_GROUP];
    const size_t src0 = (size_t)blockIdx.x * STAGE_GROUP * slots;
    for (u32 r = t; r < total; r += STAGE_GROUP) {
        u32 lo = 0, hi = STAGE_GROUP;   // position s with pre[s] <= r < pre[s + 1]
        while (hi - lo > 1) {
            const u32 mid = (lo + hi) >> 1;
            if (pre[mid] <= r) lo = mid;
            else hi = mid;
        }
        const u32 j = r - pre[lo];
        const size_t src = src0 + (size_t)lo * slots + j;
        recLeaf[dst[lo] + j] = stLeaf[src];
        recVal[dst[lo] + j] = stVal[src];
    }
}

__global__ void k_overflowFlags(const u32 *__restrict__ counts, u32 links, u32 slots, u32 *__restrict__ flag) {
    const u32 l = blockIdx.x * BLOCK + threadIdx.x;
    if (l < links) flag[l] = counts[l] > slots ? 1u : 0u;
}

// first record of every leaf segment in the sorted record stream
// first and one-past-last record of each leaf's run in the sorted records
__global__ void k_segmentStarts(const u32 *__restrict__ leafSorted, size_t n, u32 *__restrict__ segStart,
                                u32 *__restrict__ segEnd) {
    const size_t j = (size_t)blockIdx.x * BLOCK + threadIdx.x;
    if (j >= n) return;
    const u32 l = leafSorted[j];
    if (j == 0 || leafSorted[j - 1] != l) segStart[l] = (u32)j;
    if (j + 1 == n || leafSorted[j + 1] != l) segEnd[l] = (u32)(j + 1);
}

// Kahan sums of each leaf's records in record order (deterministic)
__global__ void k_segmentReduce(const uint2 *__restrict__ valSorted, const u32 *__restrict__ segStart,
                                const u32 *__restrict__ segEnd, u32 leafCount, float *__restrict__ ksum,
                                float *__restrict__ clen, u32 *__restrict__ hitCount, float *__restrict__ kcomp,
                                float *__restrict__ lcomp) {
    const u32 l = blockIdx.x * BLOCK + threadIdx.x;
    if (l >= leafCount) return;
    const u32 s = segStart[l];
    float sk = 0.f, ck = 0.f, sl = 0.f, cl = 0.f;
    u32 cnt = 0;
    if (s != INVALID) {
        const u32 e = segEnd[l];
        for (u32 j = s; j < e; ++j, ++cnt) {
            const uint2 v = valSorted[j];
            const float a = __uint_as_float(v.x) - ck;
            const float ta = sk + a;
            ck = (ta - sk) - a;
            sk = ta;
            const float b = __uint_as_float(v.y) - cl;
            const float tb = sl + b;
            cl = (tb - sl) - b;
            sl = tb;
        }
    }
    ksum[l] = sk;
    clen[l] = sl;
    hitCount[l] = cnt;
    if (kcomp) {   // the Kahan sums approximate sk - ck and sl - cl
        kcomp[l] = -ck;
        lcomp[l] = -cl;
    }
}

TSR_HD void twoSumAdd(float &hi, float &lo, float x) {   // hi + lo += x (Kahan-Babuska)
    const float s = hi + x;
    const float bp = s - hi;
    const float err = (hi - (s - bp)) + (x - bp);
    hi = s;
    lo += err;
}

// fold one link chunk's per-leaf Kahan sums (sum, compensation) into running
// (hi, lo) pairs, chunk after chunk (fixed order: deterministic)
__global__ void k_mergeChunk(u32 leafCount, const float *__restrict__ ks, const float *__restrict__ kc,
                             const float *__restrict__ ls, const float *__restrict__ lc, const u32 *__restrict__ hits,
                             float *__restrict__ kHi, float *__restrict__ kLo, float *__restrict__ lHi,
                             float *__restrict__ lLo, u32 *__restrict__ hitSum) {
    const u32 l = blockIdx.x * BLOCK + threadIdx.x;
    if (l >= leafCount) return;
    float h = kHi[l], lo = kLo[l];
    twoSumAdd(h, lo, ks[l]);
    kHi[l] = h;
    kLo[l] = lo + kc[l];
    h = lHi[l];
    lo = lLo[l];
    twoSumAdd(h, lo, ls[l]);
    lHi[l] = h;
    lLo[l] = lo + lc[l];
    hitSum[l] += hits[l];
}

__global__ void k_finishChunks(u32 leafCount, const float *__restrict__ kHi, const float *__restrict__ kLo,
                               const float *__restrict__ lHi, const float *__restrict__ lLo, float *__restrict__ ksum,
                               float *__restrict__ clen) {
    const u32 l = blockIdx.x * BLOCK + threadIdx.x;
    if (l >= leafCount) return;
    ksum[l] = kHi[l] + kLo[l];
    clen[l] = lHi[l] + lLo[l];
}

// leaf flags from hit counts, then OR into ancestors (Occupancy.cpp:491-516)
__global__ void k_setLeafFlags(const u32 *__restrict__ hitCount, const u32 *__restrict__ leafNode, u32 leafCount,
                               u32 flag, u32 *__restrict__ states) {
    const u32 l = blockIdx.x * BLOCK + threadIdx.x;
    if (l >= leafCount || hitCount[l] == 0) return;
    atomicOr(&states[leafNode[l]], flag);
}

__global__ void k_propagateFlag(const u32 *__restrict__ leafNode, u32 leafCount, const u32 *__restrict__ parents,
                                u32 flag, u32 *__restrict__ states) {
    const u32 l = blockIdx.x * BLOCK + threadIdx.x;
    if (l >= leafCount) return;
    u32 nd = leafNode[l];
    if (!(states[nd] & flag)) return;
    for (nd = parents[nd]; nd != INVALID; nd = parents[nd]) {
        const u32 old = atomicOr(&states[nd], flag);
        if (old & flag) break;
    }
}

// ---------------------------------------------------------------------------
// priors (Occupancy.cpp:530-568)
// ---------------------------------------------------------------------------

// Subtree sums (hi, lo) of E and S cone lengths per node, plus aux = node state
// flags | deepest leaf depth << 8, built bottom-up one depth level at a time.
__global__ void k_subtreeLeaves(const u32 *__restrict__ leafNode, u32 leafCount, const float *__restrict__ clen,
                                const uint8_t *__restrict__ depth, const u32 *__restrict__ states,
                                float4 *__restrict__ sums, u32 *__restrict__ aux) {
    const u32 l = blockIdx.x * BLOCK + threadIdx.x;
    if (l >= leafCount) return;
    const u32 nd = leafNode[l];
    sums[nd] = make_float4(clen[l], 0.f, clen[leafCount + l], 0.f);
    aux[nd] = states[nd] | ((u32)depth[nd] << 8);
}

__global__ void k_subtreeLevel(const u32 *__restrict__ child, const uint8_t *__restrict__ depth, u32 nodeCount,
                               u32 level, const u32 *__restrict__ states, float4 *__restrict__ sums,
                               u32 *__restrict__ aux) {
    const u32 nd = blockIdx.x * BLOCK + threadIdx.x;
    if (nd >= nodeCount || depth[nd] != level) return;
    const u32 c = child[nd];
    if (c & LEAF_BIT) return;
    float eh = 0.f, el = 0.f, sh = 0.f, sl = 0.f;
    u32 md = 0;
    for (u32 s = 0; s < 8; ++s) {
        const float4 v = sums[c + s];
        twoSumAdd(eh, el, v.x);
        el += v.y;
        twoSumAdd(sh, sl, v.z);
        sl += v.w;
        md = max(md, aux[c + s] >> 8);
    }
    sums[nd] = make_float4(eh, el, sh, sl);
    aux[nd] = states[nd] | (md << 8);
}

// leaves whose prior the occupancy ratio reads: E + S kernel sums >= confidence
// threshold (Occupancy.cpp:575-576); all leaves when `all` is set.  With `stage`, the
// certified modes start every query at stage 0 and mark the others PRIOR_NOT_QUERIED.
constexpr uint8_t PRIOR_NOT_QUERIED = 255, PRIOR_DONE = 254;
__global__ void k_priorQueryFlags(const float *__restrict__ ksum, u32 leafCount, double threshold, int all,
                                  double *__restrict__ prior, u32 *__restrict__ flag, uint8_t *__restrict__ stage) {
    const u32 l = blockIdx.x * BLOCK + threadIdx.x;
    if (l >= leafCount) return;
    const bool used = all || ((double)ksum[l] + (double)ksum[leafCount + l] >= threshold);
    flag[l] = used ? 1u : 0u;
    if (stage) stage[l] = used ? 0 : PRIOR_NOT_QUERIED;
    if (!used) prior[l] = __longlong_as_double(0x7ff8000000000000ll);   // not needed by the field: NaN
}

// flag = 1 for the leaves at stage s of the certified cascade
__global__ void k_stageFlags(const uint8_t *__restrict__ stage, u32 leafCount, uint8_t s, u32 *__restrict__ flag) {
    const u32 l = blockIdx.x * BLOCK + threadIdx.x;
    if (l < leafCount) flag[l] = stage[l] == s ? 1u : 0u;
}

// queries in leaf (DFS = spatial) order
__global__ void k_priorQueryList(const u32 *__restrict__ flag, const u32 *__restrict__ pos, u32 leafCount,
                                 u32 *__restrict__ queries) {
    const u32 l = blockIdx.x * BLOCK + threadIdx.x;
    if (l < leafCount && flag[l]) queries[pos[l]] = l;
}

// Priors traversal records, in two node-padded arrays (node i at index i + NODE_PAD)
// so that one popped entry's 8 children are read as one aligned 64-byte run of
// (child link, state | deepest leaf depth << 8) and one aligned 128-byte run of
// subtree sums (E, S) in FP64.

// Also turns each (E hi, E lo, S hi, S lo) float4 of subtree sums, in place, into the FP64
// pair ((double)E hi + (double)E lo, (double)S hi + (double)S lo), once per node: the query
// traversal adds these doubles per summed child instead of converting per (child, query).
__global__ void k_packPriorLinks(const u32 *__restrict__ child, const u32 *__restrict__ aux, u32 nodeCount,
                                 uint2 *__restrict__ links, float4 *__restrict__ sums) {
    const u32 nd = blockIdx.x * BLOCK + threadIdx.x;
    if (nd >= nodeCount) return;
    links[nd] = make_uint2(child[nd], aux[nd]);
    const float4 v = sums[nd];
    reinterpret_cast<double2 *>(sums)[nd] = make_double2((double)v.x + (double)v.y, (double)v.z + (double)v.w);
}

// The priors traversal keeps an LDS stack of inner nodes still to open.  The first
// pass uses a short stack; a group that would overflow it is abandoned and listed,
// and a second pass redoes the listed groups with a stack of the proven bound.  The
// short stack (14 B per entry) leaves room in the 160 KB of LDS per CU for 8 waves per
// SIMD of one-wave blocks.
constexpr u32 PRIOR_STACK_SHORT = 320;
constexpr u32 XCDS = 8;   // MI355X: blocks are dispatched round-robin over 8 XCDs, each with its own L2
constexpr u32 PRIOR_STACK = 1344;   // >= 64 * (MAX_DEPTH + 1): one step pushes <= 56 more than it pops

// Group form of SphereNodeStatesChecker + LeavesIterator (Occupancy.cpp:541-566):
// one wavefront serves GROUP_Q consecutive query leaves (DFS order: spatial
// neighbours, whose spheres overlap almost entirely).  A stack entry holds a node
// to open (its first child id, its grid coordinates and depth) and a mask of the
// queries that still need its children; each step pops up to 8 entries and tests their 8
// children on 64 lanes (entry = lane / 8, child = lane % 8), every lane looping
// over the queries of its entry's mask.  Each child record is read once for all
// queries of the group, and only the child records: the entry already carries
// what the parent record would give.  Children entirely inside the sphere whose
// deepest leaf respects the depth limit contribute their subtree sums; straddling
// inner children are pushed.  Every query applies the FP64 predicates of the
// reference (behind an FP32 pre-test with a proven margin); sums are FP64 per lane
// and query, reduced across the wave in a fixed order (deterministic).
//
// Depth-capped bound (CAPPED, capLevels = c): a query of depth dq does not open
// straddling children at depth >= dq - c.  The part of such a child that the exact
// traversal sums is a sub-sum of the child's subtree sums (all terms are >= 0), so those
// go to upper-bound accumulators (Eu, Su) instead, and the query gets the interval
// [E / (E + S + Su), (E + Eu) / (E + Eu + S)] (interval[query index]), which holds the exact
// prior.  The capped form keeps four FP64 sums per query, the exact form two.
#define PRIORS_GROUP_PARAMS                                                                                \
    Frame f, const uint2 *__restrict__ links, const double2 *__restrict__ sums, const u64 *__restrict__ leafLoc, \
        const u32 *__restrict__ queries, u32 queryCount, float relRadius, u32 capLevels,                       \
        double *__restrict__ prior, double2 *__restrict__ interval, u32 *err,                                  \
        unsigned long long *__restrict__ stats, u32 *__restrict__ nextGroup,                                   \
        const u32 *__restrict__ groupList, const u32 *__restrict__ groupListCount,                              \
        u32 *__restrict__ overflowList, u32 *__restrict__ overflowCount
#define PRIORS_GROUP_ARGS                                                                                  \
    f, links, sums, leafLoc, queries, queryCount, relRadius, capLevels, prior, interval, err, stats, nextGroup, \
        groupList, groupListCount, overflowList, overflowCount

template <u32 GROUP_Q, u32 STACK, bool CAPPED>
__device__ __forceinline__ void priorsGroup(PRIORS_GROUP_PARAMS) {
    __shared__ u32 stFirst[STACK];
    __shared__ u64 stK[STACK];          // kx | ky << 21 | kz << 42
    __shared__ uint16_t stMD[STACK];    // query mask | depth << 8
    __shared__ double qc[GROUP_Q][3], qr2[GROUP_Q], qri2[GROUP_Q];
    // FP32 query data packed for two LDS reads per query: (centre, r^2 + margin),
    // (r^2 - margin, inner r^2 - margin, inner r^2 + margin, depth limit)
    __shared__ float4 qA[GROUP_Q], qB[GROUP_Q];
    __shared__ u32 sp, abandoned;
    const u32 lane = threadIdx.x;
    u32 nPops = 0, nTests = 0, nWhole = 0;
    // FP32 pre-test of the FP64 sphere predicates, in coordinates relative to the root
    // corner (|coordinate| <= S, the root side; u = 2^-24 is the float rounding unit; M
    // bounds every absolute FP64 coordinate, whose own roundings are far below u S).  A
    // child cube of side h is taken as centre c = (2k+1) * h/2 and half side h/2 in float:
    // the half side is within u h/2 of its FP64 value, so (2k+1) h/2 (the odd factor is
    // exact) is within u S, and the product rounds once (u S).  The query centre rounds
    // once (u S).  Near a decision the centre distance t is below r + h/2 and rounds once
    // (u (r + h/2)), the nearest / farthest offsets t -+ h/2 once more (u r) and the half
    // side is off by u h/2.  With h <= S/2, the offsets per axis are within
    // e = u (3.5 S + 2 r) (+ 2^-50 M for the FP64 inputs) of the FP64 ones, and the
    // squared distance near radius r within 2 sqrt(3) e r + 3 e^2 + 5 u r^2 (three fused
    // multiply-adds, thresholds rounded to float), below the margin
    // 4 e r + 4 e^2 + 2^-20 r^2.  Outside that margin the FP32 result decides, inside it
    // the exact FP64 test does.
    double M = 0.0;
    for (int a = 0; a < 3; ++a) M = fmax(M, fabs(f.rootMin[a]) + f.rootSize);
    // Groups are taken dynamically from a counter per XCD slice (blocks are dispatched
    // round-robin over the 8 XCDs, so block b runs on XCD b mod 8): each XCD's L2 sees
    // one contiguous, spatially coherent range, and waves with cheap groups take more.
    // A wave whose slice is used up moves on to the next slice.  With a group list,
    // the listed groups are taken from one counter.
    const u32 groups = groupList ? *groupListCount : (queryCount + GROUP_Q - 1) / GROUP_Q;
    const u32 slices = groupList ? 1u : XCDS;
    const u32 slice = (groups + slices - 1) / slices;
    u32 cur = blockIdx.x % slices, tried = 0;
    __shared__ u32 gNext;
    for (;;) {
        const u32 gBegin = min(groups, cur * slice), gEnd = min(groups, gBegin + slice);
        if (lane == 0) gNext = gBegin + atomicAdd(&nextGroup[cur], 1u);
        __syncthreads();
        const u32 gi = gNext;
        __syncthreads();
        if (gi >= gEnd) {
            if (++tried == slices) break;
            cur = cur + 1 == slices ? 0u : cur + 1;
            continue;
        }
        const u32 g = groupList ? groupList[gi] : gi;
        const u32 nQ = min(GROUP_Q, queryCount - g * GROUP_Q);
        if (lane < GROUP_Q) {
            const u32 l = lane < nQ ? queries[g * GROUP_Q + lane] : 0u;
            const u64 qloc = lane < nQ ? leafLoc[l] : 1ull;
            u32 kq[3];
            locCoords(qloc, kq[0], kq[1], kq[2]);
            const u32 dq = locDepth(qloc);
            for (int a = 0; a < 3; ++a) qc[lane][a] = f.centre(a, kq[a], dq);
            const double radius = f.h(dq) * (double)relRadius;
            qr2[lane] = radius * radius;
            qri2[lane] = radius * radius * (1.0 - 1e-9);
            // decision thresholds in float, formed once per query: r^2 +- margin, inner r^2 +- margin
            const float r2F = (float)qr2[lane], ri2F = (float)qri2[lane];
            const double eps = ldexp(3.5 * f.rootSize + 2.0 * radius, -24) + ldexp(M, -50);
            const float mg = (float)(4.0 * eps * radius + 4.0 * eps * eps + ldexp(radius * radius, -20));
            qA[lane] = make_float4((float)(qc[lane][0] - f.rootMin[0]), (float)(qc[lane][1] - f.rootMin[1]),
                                   (float)(qc[lane][2] - f.rootMin[2]), r2F + mg);
            qB[lane] = make_float4(r2F - mg, ri2F - mg, ri2F + mg, __uint_as_float(dq + MAX_DEPTH_DIFFERENCE));
        }
        double accE[GROUP_Q], accS[GROUP_Q], upE[GROUP_Q], upS[GROUP_Q];
#pragma unroll
        for (u32 j = 0; j < GROUP_Q; ++j) accE[j] = accS[j] = upE[j] = upS[j] = 0.0;
        const uint2 root = links[0];
        const uint8_t all = (uint8_t)((1u << nQ) - 1u);
        if (lane == 0) {
            sp = 0;
            abandoned = 0;
            if ((root.y & FLAG_S) && !(root.x & LEAF_BIT)) {
                stFirst[0] = root.x;
                stK[0] = 0ull;
                stMD[0] = all;
                sp = 1;
            }
        }
        if ((root.y & FLAG_S) && (root.x & LEAF_BIT) && lane == 0) {   // every query sums the root
            const double2 rs = sums[0];
#pragma unroll
            for (u32 j = 0; j < GROUP_Q; ++j) {
                accE[j] = rs.x;
                accS[j] = rs.y;
            }
        }
        __syncthreads();
        for (;;) {
            const u32 top = sp;
            if (top == 0) break;
            const u32 take = min(top, 8u);
            const u32 base = top - take;
            const u32 e = lane >> 3, slot = lane & 7;
            uint8_t pushMask = 0;
            uint2 cn = make_uint2(0u, 0u);
            u32 k[3] = {0u, 0u, 0u}, d = 0;
            nPops += lane == 0 ? take : 0;
            if (e < take) {
                const u32 cid = stFirst[base + e] + slot;
                const u64 pk = stK[base + e];
                const u32 md16 = stMD[base + e];
                cn = links[cid];
                const double2 cs = sums[cid];
                // both loads in flight together: keep the sums load ahead of the state test
                asm volatile("" ::"v"(cs.x), "v"(cs.y));
                const u32 mask = md16 & 0xffu;
                d = (md16 >> 8) + 1;
                // child coordinates: slot bit 0 -> x, bit 1 -> y, bit 2 -> z (location code order)
                k[0] = 2u * (u32)(pk & 0x1fffffull) + (slot & 1u);
                k[1] = 2u * (u32)((pk >> 21) & 0x1fffffull) + ((slot >> 1) & 1u);
                k[2] = 2u * (u32)(pk >> 42) + (slot >> 2);
                if (cn.y & FLAG_S) {
                    ++nTests;
                    // the child cube as centre and half side, in float, relative to the root corner
                    const float halfF = ldexpf(f.rootSizeF, -(int)d - 1);
                    float cenF[3];
                    for (int a = 0; a < 3; ++a) cenF[a] = (float)(2u * k[a] + 1u) * halfF;
                    const bool leaf = (cn.x & LEAF_BIT) != 0;
                    const u32 mld = cn.y >> 8;
                    bool summed = false;
#pragma unroll
                    for (u32 j = 0; j < GROUP_Q; ++j) {
                        if (!((mask >> j) & 1u)) continue;
                        // an entry carries query j only if its children are within j's depth limit
                        const float4 A = qA[j], B = qB[j];
                        // both LDS reads in flight together (one wait, not one per branch)
                        asm volatile("" ::"v"(A.x), "v"(B.x), "v"(B.y), "v"(B.z), "v"(B.w));
                        const u32 md = __float_as_uint(B.w);
                        const float cq[3] = {A.x, A.y, A.z};
                        // per axis: distance of the centres t, nearest offset t - half (or 0) and
                        // farthest offset t + half; fused multiply-adds, one rounding each
                        float dminF = 0.f, dmaxF = 0.f;
                        for (int q3 = 0; q3 < 3; ++q3) {
                            const float t = fabsf(cq[q3] - cenF[q3]);
                            const float o = fmaxf(t - halfF, 0.f);
                            dminF = __builtin_fmaf(o, o, dminF);
                            const float far = t + halfF;
                            dmaxF = __builtin_fmaf(far, far, dmaxF);
                        }
                        if (dminF > A.w) continue;      // surely outside
                        if (!(dminF < B.x)) {           // undecided: the FP64 test
                            double dmin = 0.0;
                            for (int q3 = 0; q3 < 3; ++q3) {   // SphereNodesChecker.cpp:29-40
                                const double c = qc[j][q3];
                                const double lo = f.B(q3, k[q3], d), hi = f.B(q3, (u64)k[q3] + 1, d);
                                if (c < lo) dmin += (lo - c) * (lo - c);
                                else if (c > hi) dmin += (c - hi) * (c - hi);
                            }
                            if (!(dmin <= qr2[j])) continue;
                        }
                        bool whole = leaf;
                        if (!whole && mld <= md) {
                            if (dmaxF < B.y) whole = true;
                            else if (!(dmaxF > B.z)) {
                                    double dmax = 0.0;
                                for (int q3 = 0; q3 < 3; ++q3) {
                                    const double c = qc[j][q3];
                                    const double lo = f.B(q3, k[q3], d), hi = f.B(q3, (u64)k[q3] + 1, d);
                                    const double far = fmax(fabs(c - lo), fabs(hi - c));
                                    dmax += far * far;
                                }
                                whole = dmax <= qri2[j];
                            }
                        }
                        if (whole) {
                            accE[j] += cs.x;
                            accS[j] += cs.y;
                            summed = true;
                        } else if (d < md) {   // children at depth d + 1 <= md still count
                            // capped: children at depth >= dq - capLevels are bounded, not opened
                            // (md = dq + MAX_DEPTH_DIFFERENCE)
                            if (CAPPED && d + MAX_DEPTH_DIFFERENCE + capLevels >= md) {
                                upE[j] += cs.x;
                                upS[j] += cs.y;
                                summed = true;
                            } else {
                                pushMask |= (uint8_t)(1u << j);
                            }
                        }
                    }
                    nWhole += summed ? 1u : 0u;
                }
            }
            __syncthreads();   // every lane has read its entry before the slots are reused
            const u64 pm = __ballot(pushMask != 0);
            const u32 pos = lanesBelow(pm);
            const u32 newTop = base + (u32)__popcll(pm);
            if (newTop > STACK) {
                if (lane == 0) {
                    if (overflowList) overflowList[atomicAdd(overflowCount, 1u)] = gi;
                    else atomicOr(err, 1u);
                    abandoned = 1;
                    sp = 0;
                }
            } else {
                if (pushMask) {
                    stFirst[base + pos] = cn.x;
                    stK[base + pos] = (u64)k[0] | ((u64)k[1] << 21) | ((u64)k[2] << 42);
                    stMD[base + pos] = (uint16_t)(pushMask | (d << 8));
                }
                if (lane == 0) sp = newTop;
            }
            __syncthreads();
        }
        // fixed-order wave reduction per query (FP64)
        const bool write = abandoned == 0;
#pragma unroll
        for (u32 j = 0; j < GROUP_Q; ++j) {
            double ee = accE[j], ss = accS[j], ue = upE[j], us = upS[j];
            for (int off = 32; off > 0; off >>= 1) {
                ee += __shfl_xor(ee, off);
                ss += __shfl_xor(ss, off);
                if (CAPPED) {
                    ue += __shfl_xor(ue, off);
                    us += __shfl_xor(us, off);
                }
            }
            if (lane == 0 && j < nQ && write) {
                if (CAPPED)
                    interval[g * GROUP_Q + j] = make_double2(ee / (ee + ss + us), (ee + ue) / (ee + ue + ss));
                else
                    prior[queries[g * GROUP_Q + j]] = ee / (ee + ss);
            }
        }
        __syncthreads();
    }
    if (stats) {
        waveAccumulate(&stats[0], nPops);
        waveAccumulate(&stats[1], nTests);
        waveAccumulate(&stats[2], nWhole);
    }
}

// register use capped for 8 waves per SIMD (64 VGPRs) in the exact form: the step's memory
// and LDS round trips are hidden better than with fewer waves (7: 488 ms, 8: 474 ms on 36.5M
// samples); 9 waves spill and run slower.  The capped form, with twice the FP64 sums, is
// capped at 7 waves (72 VGPRs; at 8 its spill takes the certified priors 207 -> 254 ms).
// The long-stack instantiation is held to few waves by its LDS stack anyway and is not capped.
template <u32 GROUP_Q, u32 STACK, bool CAPPED>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(STACK == PRIOR_STACK_SHORT ? (CAPPED ? 7 : 8) : 1))) void k_priorsGroup(PRIORS_GROUP_PARAMS) {
    priorsGroup<GROUP_Q, STACK, CAPPED>(PRIORS_GROUP_ARGS);
}
// Queries per wave, exact / capped.  Each step runs the per-query loop once per query of
// the group whose bit some popped entry carries, while a child lane is busy only for the
// queries of its own entry (about 3 of 8): 4 queries per group run fewer loop iterations in
// all than 8 (6.7G against 8.1G on 36.5M samples) although they pop more entries (13.4G
// against 8.1G): 474 against 516 ms; 2 queries per group pop 22.3G entries (500 ms).
constexpr u32 PRIOR_GROUP = 4, PRIOR_GROUP_CAPPED = 4;

// Occupancy.cpp:570-587, FP64
__global__ void k_occupancy(const float *__restrict__ ksum, const double *__restrict__ prior, u32 leafCount,
                            double coneCount, double threshold, double *__restrict__ occ) {
    const u32 l = blockIdx.x * BLOCK + threadIdx.x;
    if (l >= leafCount) return;
    const double e = ksum[l], s = ksum[leafCount + l];
    if (e + s < threshold) {
        occ[l] = -EPS_D;
        return;
    }
    const double p = prior[l];
    const double pE = p * (e / coneCount);
    const double pS = (1.0 - p) * (s / coneCount);
    occ[l] = pE >= pS ? -EPS_D : pS - pE;
}

// What the mesh reads of a leaf's occupancy v (Occupancy.cpp:570-587) at prior p: class 0
// for v = -EPSILON (pE >= pS), 1 for 0 < v <= EPSILON, 2 for v > EPSILON.  DMC cell
// cases read the sign, and a crossing vertex lies at the corner whose |v| <= EPSILON is
// tested first (DualMarchingCell.cpp:51-71: the negative corner, which is always at
// -EPSILON, or a class-1 positive one), so two fields with equal classes give the same
// mesh.  Returns -1 where rounding could decide the class: within 1e-9 T of a class
// boundary (T = max(E, S) / M bounds pE + pS; FP64 roundings of the occupancy formula
// are near 1e-16 T) or below 1e-18 (where v + EPSILON would round to EPSILON).
__device__ int occupancyClass(double p, double e, double s, double M) {
    const double pE = p * (e / M), pS = (1.0 - p) * (s / M);
    const double tol = 1e-9 * (fmax(e, s) / M);
    const double v = pS - pE;
    if (!(v == v)) return -1;
    if (v < -tol) return 0;
    if (v <= fmax(tol, 1e-18)) return -1;
    if (v < EPS_D - tol) return 1;
    if (v > EPS_D + tol) return 2;
    return -1;
}

// The judge of the certified cascade: the occupancy formula on both ends of a query's
// prior interval (widened by 1e-9 relative, far above the difference between the sums of
// two traversal orders).  The occupancy is monotone in the prior, so equal classes at both
// ends mean the exact prior gives that class too: the leaf keeps the interval midpoint as
// its prior (so its occupancy has the class of the exact one) and is done at stage s;
// otherwise it moves on to stage s + 1.
__global__ void k_certify(const u32 *__restrict__ list, u32 n, const double2 *__restrict__ interval,
                          const float *__restrict__ ksum, u32 leafCount, double coneCount, uint8_t s,
                          uint8_t *__restrict__ stage, uint8_t *__restrict__ level, double *__restrict__ prior) {
    const u32 i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    const u32 l = list[i];
    const double2 iv = interval[i];
    const double e = ksum[l], sm = ksum[leafCount + l];
    const double lo = iv.x * (1.0 - 1e-9), hi = fmin(1.0, iv.y * (1.0 + 1e-9));
    const int c = (lo >= 0.0 && hi <= 1.0 && lo <= hi) ? occupancyClass(lo, e, sm, coneCount) : -1;
    if (c >= 0 && c == occupancyClass(hi, e, sm, coneCount)) {
        prior[l] = 0.5 * (iv.x + iv.y);
        level[l] = s;
        stage[l] = PRIOR_DONE;
    } else {
        stage[l] = s + 1;
    }
}

__global__ void k_setLevel(const u32 *__restrict__ list, u32 n, uint8_t v, uint8_t *__restrict__ level) {
    const u32 i = blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) level[list[i]] = v;
}

__global__ void k_fillF(float *p, size_t n, float v) {
    const size_t i = (size_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void k_iotaU(u32 *p, u32 n, u32 base) {
    const u32 i = blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) p[i] = base + i;
}
__global__ void k_fillU(u32 *p, size_t n, u32 v) {
    const size_t i = (size_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) p[i] = v;
}


}  // namespace

__global__ void k_camKeys(const u32 *__restrict__ cams, u32 links, u32 camCount, u32 base, u32 *__restrict__ key,
                          u32 *__restrict__ idx) {
    const u32 l = blockIdx.x * BLOCK + threadIdx.x;
    if (l >= links) return;
    key[l] = min(cams[l], camCount);   // invalid links (>= camCount) last
    idx[l] = base + l;
}

constexpr u32 CONE_PT_BLOCKS = 2048;   // blocks of 256 threads (about two per SIMD pair at 4 waves/SIMD)

void priorRecords(tsr_context *ctx, const tsr_tree *tree, const tsr_field *field, PriorRecords &r) {
    hipStream_t st = ctx->stream;
    const u32 L = tree->leafCount, N = tree->nodeCount;
    r.sumsPad.alloc((size_t)N + NODE_PAD);
    r.linksPad.alloc((size_t)N + NODE_PAD);
    float4 *sums = r.sumsPad.p + NODE_PAD;
    DBuf<u32> aux(N);
    k_subtreeLeaves<<<gridFor(L, BLOCK), BLOCK, 0, st>>>(tree->leafNode.p, L, field->clen.p, tree->nodeDepth.p,
                                                         field->states.p, sums, aux.p);
    for (int lvl = (int)tree->maxDepth - 1; lvl >= 0; --lvl)
        k_subtreeLevel<<<gridFor(N, BLOCK), BLOCK, 0, st>>>(tree->child.p, tree->nodeDepth.p, N, (u32)lvl,
                                                            field->states.p, sums, aux.p);
    k_packPriorLinks<<<gridFor(N, BLOCK), BLOCK, 0, st>>>(tree->child.p, aux.p, N, r.linksPad.p + NODE_PAD, sums);
}

namespace {

__global__ void k_queryDepth(const u32 *__restrict__ q, u32 n, const u64 *__restrict__ leafLoc, u32 *__restrict__ key) {
    const u32 i = blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) key[i] = locDepth(leafLoc[q[i]]);
}

// The queries of a priors pass grouped by depth (stable: leaf order within a depth), so that a
// group of consecutive queries shares one sphere radius and their stack entries' query masks
// are denser (priors 531 -> 518 ms at 36.5M samples).  The prior of each leaf does not depend
// on the order.
void queriesByDepth(Scratch &sc, const tsr_tree *tree, u32 *queries, u32 n, hipStream_t st) {
    if (n < 2) return;
    DBuf<u32> keys(n), keysS(n), sorted(n);
    k_queryDepth<<<gridFor(n, BLOCK), BLOCK, 0, st>>>(queries, n, tree->leafLoc.p, keys.p);
    sortPairs(sc, keys.p, keysS.p, queries, sorted.p, n, 0, 5, st);
    HIPCHECK(hipMemcpyAsync(queries, sorted.p, (size_t)n * 4, hipMemcpyDeviceToDevice, st));
}

// leaves with flag 1, in leaf order, into list; returns their count
u32 flaggedLeaves(Scratch &sc, const u32 *flag, u32 L, u32 *list, hipStream_t st) {
    if (L == 0) return 0;
    DBuf<u32> pos(L);
    exclusiveScan(sc, flag, pos.p, L, st);
    k_priorQueryList<<<gridFor(L, BLOCK), BLOCK, 0, st>>>(flag, pos.p, L, list);
    return readScalar(pos.p + L - 1, st) + readScalar(flag + L - 1, st);
}

// One priors traversal over `queries` (nq leaves, grouped by depth): the short-stack launch,
// then the groups that overflowed it with the long stack.  capLevels 0: exact priors into
// prior[leaf]; otherwise depth-capped intervals into interval[query index].
void priorsPass(tsr_context *ctx, const tsr_tree *tree, const PriorRecords &rec, const u32 *queries, u32 nq,
                float relRadius, u32 capLevels, double *prior, double2 *interval, tsr_field *field) {
    if (nq == 0) return;
    hipStream_t st = ctx->stream;
    const u32 gq = capLevels ? PRIOR_GROUP_CAPPED : PRIOR_GROUP;
    const u32 groups = (nq + gq - 1) / gq;
    const u32 grid = 256u * 32u;   // multiple of the XCD count
    // counters: 8 XCD slices of the first launch, 1 for the second; overflow count; error flag
    DBuf<u32> counters(16), overflowList(groups);
    DBuf<unsigned long long> kst(4);
    HIPCHECK(hipMemsetAsync(counters.p, 0, 64, st));
    HIPCHECK(hipMemsetAsync(kst.p, 0, 32, st));
    u32 *err = counters.p + 10;
    const Frame f = tree->frame;
    const uint2 *links = rec.links();
    const double2 *sums = rec.sums();
    const u64 *leafLoc = tree->leafLoc.p;
    timedLaunch(ctx, KSTAT_PRIORS, [&] {
        if (capLevels) {
            k_priorsGroup<PRIOR_GROUP_CAPPED, PRIOR_STACK_SHORT, true><<<grid, 64, 0, st>>>(
                f, links, sums, leafLoc, queries, nq, relRadius, capLevels, prior, interval, err, kst.p, counters.p,
                nullptr, nullptr, overflowList.p, counters.p + 9);
            k_priorsGroup<PRIOR_GROUP_CAPPED, PRIOR_STACK, true><<<1024, 64, 0, st>>>(
                f, links, sums, leafLoc, queries, nq, relRadius, capLevels, prior, interval, err, kst.p, counters.p + 8,
                overflowList.p, counters.p + 9, nullptr, nullptr);
        } else {
            k_priorsGroup<PRIOR_GROUP, PRIOR_STACK_SHORT, false><<<grid, 64, 0, st>>>(
                f, links, sums, leafLoc, queries, nq, relRadius, 0, prior, interval, err, kst.p, counters.p, nullptr,
                nullptr, overflowList.p, counters.p + 9);
            k_priorsGroup<PRIOR_GROUP, PRIOR_STACK, false><<<1024, 64, 0, st>>>(
                f, links, sums, leafLoc, queries, nq, relRadius, 0, prior, interval, err, kst.p, counters.p + 8,
                overflowList.p, counters.p + 9, nullptr, nullptr);
        }
    });
    unsigned long long ps[3];
    HIPCHECK(hipMemcpyAsync(ps, kst.p, 24, hipMemcpyDeviceToHost, st));
    u32 cnt[16];
    HIPCHECK(hipMemcpyAsync(cnt, counters.p, 64, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    if (cnt[10]) throw tsr_error(TSR_ERR_HIP, "priors traversal stack overflow");
    field->priorsOverflowGroups += cnt[9];
    if (std::getenv("TSR_TRACE"))
        std::fprintf(stderr, "[tsr occupancy] priors (cap %u): %u queries, %llu pops, %llu tests, %llu summed, "
                             "%u of %u groups redone with the long stack\n",
                     capLevels, nq, ps[0], ps[1], ps[2], cnt[9], groups);
    tsr_kernel_stat &k = ctx->kstat[KSTAT_PRIORS];
    // per query: index 4 B + leaf code 8 B + result 8 B (16 B interval); per popped entry: the
    // 8 child links (8 B each); per child summed or bounded by some query of the group: its 16 B sums
    k.bytes += (u64)nq * (capLevels ? 28 : 20) + ps[0] * 64 + ps[2] * 16;
    k.units += nq;
}

}  // namespace

// Priors of the queried leaves (Occupancy.cpp:530-568).  Plain: one exact traversal.
// Certified (TSR_PRIORS_CERTIFIED, or TSR_CULL_LEARNED with the network choosing each
// query's first stage): stage 0 bounds every query's prior with the traversal capped 3
// levels above its depth, stage 1 with the cap 2 levels above, and k_certify keeps the
// leaves whose occupancy class the interval decides; stage 2 computes the exact prior of
// the rest.  The mesh is that of the plain field (equal classes), while the priors and
// occupancies of certified leaves are interval midpoints.
void priors(tsr_context *ctx, const tsr_tree *tree, const tsr_params &prm, int flags, const tsr_culler *culler,
            tsr_field *field) {
    hipStream_t st = ctx->stream;
    Scratch &sc = ctx->scratch;
    const u32 L = tree->leafCount;
    const bool certified = (flags & (TSR_PRIORS_CERTIFIED | TSR_CULL_LEARNED)) != 0 && field->coneCount > 0.0;
    PriorRecords rec;
    priorRecords(ctx, tree, field, rec);
    DBuf<u32> qflag(L), queries(L);
    field->priorsOverflowGroups = 0;
    for (auto &c : field->priorStages) c = 0;
    if (certified) field->levels.alloc(L);
    k_priorQueryFlags<<<gridFor(L, BLOCK), BLOCK, 0, st>>>(field->ksum.p, L, (double)prm.confidence_threshold,
                                                           (flags & TSR_PRIORS_ALL_LEAVES) ? 1 : 0, field->prior.p,
                                                           qflag.p, certified ? field->levels.p : nullptr);
    const u32 nq = flaggedLeaves(sc, qflag.p, L, queries.p, st);
    field->priorQueries = nq;
    field->priorStages[0] = nq;
    const float rr = prm.relative_radius_for_priors;
    if (!certified) {
        queriesByDepth(sc, tree, queries.p, nq, st);
        priorsPass(ctx, tree, rec, queries.p, nq, rr, 0, field->prior.p, nullptr, field);
        field->priorStages[3] = nq;
        return;
    }
    // stage per leaf (levels doubles as the stage while the cascade runs, then holds the
    // stage that decided each leaf)
    DBuf<uint8_t> stage(L);
    HIPCHECK(hipMemcpyAsync(stage.p, field->levels.p, L, hipMemcpyDeviceToDevice, st));
    if (culler && nq) {   // the network's first stage per query
        DBuf<u32> routed(2);
        HIPCHECK(hipMemsetAsync(routed.p, 0, 8, st));
        leafRoute(ctx, culler, tree, field, rec, queries.p, nq, stage.p, routed.p);
        u32 r[2];
        HIPCHECK(hipMemcpyAsync(r, routed.p, 8, hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
        field->priorStages[4] = r[0];
        field->priorStages[5] = r[1];
    }
    DBuf<double2> interval(std::max(nq, 1u));
    for (u32 s = 0; s < 3; ++s) {
        k_stageFlags<<<gridFor(L, BLOCK), BLOCK, 0, st>>>(stage.p, L, (uint8_t)s, qflag.p);
        const u32 n = flaggedLeaves(sc, qflag.p, L, queries.p, st);
        field->priorStages[1 + s] = n;
        queriesByDepth(sc, tree, queries.p, n, st);
        if (s == 2) {
            priorsPass(ctx, tree, rec, queries.p, n, rr, 0, field->prior.p, nullptr, field);
            k_setLevel<<<gridFor(std::max(n, 1u), BLOCK), BLOCK, 0, st>>>(queries.p, n, 2, field->levels.p);
            break;
        }
        priorsPass(ctx, tree, rec, queries.p, n, rr, s == 0 ? 3u : 2u, nullptr, interval.p, field);
        if (n)
            k_certify<<<gridFor(n, BLOCK), BLOCK, 0, st>>>(queries.p, n, interval.p, field->ksum.p, L,
                                                           field->coneCount, (uint8_t)s, stage.p, field->levels.p,
                                                           field->prior.p);
    }
}

// The record pipeline of one cone pass over links [0, links) of `in` (Occupancy.cpp:138-248
// updateKernels): per-leaf kernel sums and cone lengths into ksum / clen (Kahan sums over
// records sorted stably by leaf: deterministic), accepted hits per leaf into hitCount, and
// per link the confidence of its cone if not degenerate into confs (E pass, may be null).
void conePassRecords(tsr_context *ctx, const tsr_tree *tree, TraversalTree &tt,
                     const LinkInputs &in, u32 links, int sampleness, float *ksum, float *clen, u32 *hitCount,
                     double *confs, u64 &hitsOut, u32 &overOut) {
    hipStream_t st = ctx->stream;
    Scratch &sc = ctx->scratch;
    ConeArena &ar = ctx->cone;
    const u32 L = tree->leafCount;
    DBuf<u32> counts(links), offsets(links), segStart(L), segEnd(L);
    DBuf<unsigned long long> kst(4);
    auto readStats = [&](int n, unsigned long long *out) {
        HIPCHECK(hipMemcpyAsync(out, kst.p, 8 * (size_t)n, hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
    };
    // Links are processed in chunks of at most coneChunkLinks() (one chunk below that).
    // Per chunk, one traversal counts every link's accepted hits and stages its first
    // in.slots records; links with more are traversed again, writing at their final
    // offsets; the records are sorted by leaf (stable: link order within a leaf) and
    // reduced with Kahan sums.  Chunk results are folded in chunk order.
    // Algorithmic bytes per link: sample SoA 37 B (position, normal, scale, confidence,
    // camera link, depth) + 4 B count; per visited node 4 B child link (+4 B state in
    // the E pass); 12 B per staged record.
    const u64 nodeBytes = sampleness ? 4 : 8;
    const u32 chunkLinks = coneChunkLinks();
    DBuf<float> part, acc;
    DBuf<u32> partHits;
    bool merging = false;
    auto startMerging = [&] {
        merging = true;
        part.alloc(4 * (size_t)L);
        partHits.alloc(L);
        acc.alloc(4 * (size_t)L);
        HIPCHECK(hipMemsetAsync(acc.p, 0, 16 * (size_t)L, st));
        HIPCHECK(hipMemsetAsync(hitCount, 0, 4 * (size_t)L, st));
    };
    u64 totalHits = 0;
    u32 totalOver = 0;
    // link ranges in ascending order; the range length adapts to the records per link
    // seen so far, and a range whose records would exceed what one sort handles is
    // halved and traversed again
    const bool trace = std::getenv("TSR_TRACE") != nullptr;
    auto tmark = std::chrono::steady_clock::now();
    auto lap = [&](const char *what) {   // TSR_TRACE: synchronised sub-step times
        if (!trace) return;
        HIPCHECK(hipStreamSynchronize(st));
        const auto t = std::chrono::steady_clock::now();
        std::fprintf(stderr, "[tsr occupancy]   %-28s %8.2f ms\n", what,
                     std::chrono::duration<double, std::milli>(t - tmark).count());
        tmark = t;
    };
    DBuf<u32> linkCursor(8);
    double perLinkEstimate = 0.0;
    if ((u64)links * 64 > CONE_MAX_RECORDS) {
        // records per link estimated on every 64th link (counting traversal), so that the
        // first range already fits one sort
        LinkInputs ie = in;
        ie.linkStride = 64;
        const u32 sampled = (links + 63) / 64;
        HIPCHECK(hipMemsetAsync(kst.p, 0, 32, st));
        k_conePass<<<gridFor(sampled, BLOCK), BLOCK, 0, st>>>(tt, ie, links, sampleness, 0, counts.p, nullptr,
                                                              nullptr, nullptr, nullptr, kst.p);
        unsigned long long es[2];
        readStats(2, es);
        perLinkEstimate = std::max(1.0, 1.25 * (double)es[1] / sampled);
    }
    LinkRanges ranges(links, chunkLinks, CONE_MAX_RECORDS, perLinkEstimate);
    if (trace && perLinkEstimate > 0.0)
        std::fprintf(stderr, "[tsr occupancy] %s pass: ~%.1f records per link (1/64 sample), ranges of %u links\n",
                     sampleness ? "sampleness" : "emptiness", perLinkEstimate / 1.25, ranges.rangeLinks);
    while (!ranges.done()) {
        const u32 l0 = ranges.begin;
        const u32 n = ranges.next();
        if (ranges.merging && !merging) startMerging();
        lap("(chunk start)");
        LinkInputs ic = in;
        ic.linkBase = l0;
        ic.order = nullptr;
        DBuf<u32> chunkOrder;
        if (!sampleness) {   // the chunk's links grouped by camera
            DBuf<u32> camKey(n), camKeyS(n), iota(n);
            k_camKeys<<<gridFor(n, BLOCK), BLOCK, 0, st>>>(in.cams + l0, n, in.camCount, l0,
                                                           camKey.p, iota.p);
            chunkOrder.alloc(n);
            int bits = 1;
            while (bits < 32 && (1ull << bits) <= in.camCount) ++bits;
            sortPairs(sc, camKey.p, camKeyS.p, iota.p, chunkOrder.p, n, 0, bits, st);
            ic.order = chunkOrder.p;
        }
        ic.slots = (u32)std::min<size_t>(CONE_SLOTS_MAX, CONE_STAGING_BUDGET / (12 * (size_t)std::max(n, 1u)));
        const size_t stageCount = ((size_t)n + STAGE_GROUP - 1) / STAGE_GROUP * STAGE_GROUP * ic.slots;
        u32 *stLeaf = ar.stLeaf.ensure(stageCount);
        uint2 *stVal = ar.stVal.ensure(stageCount);
        lap("order + staging");
        HIPCHECK(hipMemsetAsync(kst.p, 0, 32, st));
        // persistent threads for both passes: a lane takes the next link as soon as its own
        // is done (the sampleness traversals are short, but at scale their lengths vary enough
        // to pay)
        HIPCHECK(hipMemsetAsync(linkCursor.p, 0, 32, st));
        timedLaunch(ctx, sampleness ? KSTAT_S_COUNT : KSTAT_E_COUNT, [&] {
            if (sampleness)
                k_conePassPT<true><<<std::min<u32>(gridFor(n, BLOCK), CONE_PT_BLOCKS), BLOCK, 0, st>>>(
                    tt, ic, n, counts.p, stLeaf, stVal, nullptr, linkCursor.p, XCD_COUNT, kst.p);
            else
                k_conePassPT<false><<<std::min<u32>(gridFor(n, BLOCK), CONE_PT_BLOCKS), BLOCK, 0, st>>>(
                    tt, ic, n, counts.p, stLeaf, stVal, confs, linkCursor.p, XCD_COUNT, kst.p);
        });
        unsigned long long cs[2];
        readStats(2, cs);
        {
            tsr_kernel_stat &k = ctx->kstat[sampleness ? KSTAT_S_COUNT : KSTAT_E_COUNT];
            k.bytes += (u64)n * 41 + cs[0] * nodeBytes + cs[1] * 12;
            k.units += n;
        }
        lap("traversal (record pass)");
        // cs[1] is the range's record total in 64 bits: a range over the cap is split
        // before the 32-bit offsets could wrap
        bool impossible = false;
        if (!ranges.accept(n, cs[1], impossible)) {
            if (impossible) throw tsr_error(TSR_ERR_OUT_OF_MEMORY, "too many kernel contributions for one link");
            continue;
        }
        exclusiveScan(sc, counts.p + l0, offsets.p + l0, n, st);   // chunk-relative offsets
        const u64 total = (u64)readScalar(offsets.p + l0 + n - 1, st) + readScalar(counts.p + l0 + n - 1, st);
        if (total != cs[1]) throw tsr_error(TSR_ERR_HIP, "cone pass record counts disagree with the hit total");
        totalHits += total;
        u32 *recLeaf = ar.recLeaf.ensure(total), *recLeafS = ar.recLeafS.ensure(total);
        uint2 *recVal = ar.recVal.ensure(total), *recValS = ar.recValS.ensure(total);
        k_gatherSlots<<<gridFor(n, BLOCK), BLOCK, 0, st>>>(counts.p, offsets.p, ic.order, l0, n, ic.slots,
                                                          stLeaf, stVal, recLeaf, recVal);
        lap("scan + gather");
        // links beyond their slots: list them and traverse those again
        DBuf<u32> ovFlag(n), ovList(n), ovCount(1);
        k_overflowFlags<<<gridFor(n, BLOCK), BLOCK, 0, st>>>(counts.p + l0, n, ic.slots, ovFlag.p);
        {
            DBuf<u32> iota(n);
            k_iotaU<<<gridFor(n, BLOCK), BLOCK, 0, st>>>(iota.p, n, l0);
            selectFlagged(sc, iota.p, ovFlag.p, ovList.p, ovCount.p, n, st);
        }
        const u32 nOver = readScalar(ovCount.p, st);
        totalOver += nOver;
        if (std::getenv("TSR_TRACE"))
            std::fprintf(stderr, "[tsr occupancy] %s pass, links %u..%u: %llu node visits, %llu hits, %u links over "
                         "%u slots\n", sampleness ? "sampleness" : "emptiness", l0, l0 + n - 1,
                         (unsigned long long)cs[0], (unsigned long long)total, nOver, ic.slots);
        if (nOver) {
            LinkInputs ino = ic;
            ino.order = ovList.p;
            HIPCHECK(hipMemsetAsync(kst.p, 0, 32, st));
            // few, long traversals: one wavefront per link
            DBuf<u32> werr(1);
            HIPCHECK(hipMemsetAsync(werr.p, 0, 4, st));
            timedLaunch(ctx, sampleness ? KSTAT_S_WRITE : KSTAT_E_WRITE, [&] {
                k_coneLinkWave<<<std::min<u32>(nOver, 65536), 64, 0, st>>>(
                    tt, ic, ovList.p, nOver, sampleness, tree->nodeLoc.p, offsets.p, counts.p, recLeaf,
                    recVal, werr.p, kst.p);
            });
            if (const u32 we = readScalar(werr.p, st))
                throw tsr_error(TSR_ERR_HIP, we & 1u ? "cone traversal stack overflow"
                                                     : "cone re-traversal disagrees with the counting pass");
            readStats(2, cs);
            tsr_kernel_stat &k = ctx->kstat[sampleness ? KSTAT_S_WRITE : KSTAT_E_WRITE];
            k.bytes += (u64)nOver * 45 + cs[0] * nodeBytes + cs[1] * 12;
            k.units += nOver;
        }
        lap("overflow list + re-traversal");
        int bits = 1;
        while (bits < 32 && (1ull << bits) < L) ++bits;
        sortPairs(sc, recLeaf, recLeafS, recVal, recValS, total, 0, bits, st);
        lap("sort by leaf");
        k_fillU<<<gridFor(L, BLOCK), BLOCK, 0, st>>>(segStart.p, L, INVALID);
        if (total)
            k_segmentStarts<<<gridFor(total, BLOCK), BLOCK, 0, st>>>(recLeafS, total, segStart.p, segEnd.p);
        if (!merging) {
            k_segmentReduce<<<gridFor(L, BLOCK), BLOCK, 0, st>>>(recValS, segStart.p, segEnd.p, L,
                                                                 ksum, clen, hitCount, nullptr, nullptr);
        } else {
            float *pk = part.p, *pl = part.p + L, *pkc = part.p + 2 * (size_t)L, *plc = part.p + 3 * (size_t)L;
            k_segmentReduce<<<gridFor(L, BLOCK), BLOCK, 0, st>>>(recValS, segStart.p, segEnd.p, L,
                                                                 pk, pl, partHits.p, pkc, plc);
            k_mergeChunk<<<gridFor(L, BLOCK), BLOCK, 0, st>>>(L, pk, pkc, pl, plc, partHits.p, acc.p,
                                                              acc.p + L, acc.p + 2 * (size_t)L,
                                                              acc.p + 3 * (size_t)L, hitCount);
        }
        ranges.finished(n, total);
    }
    if (merging)
        k_finishChunks<<<gridFor(L, BLOCK), BLOCK, 0, st>>>(L, acc.p, acc.p + L, acc.p + 2 * (size_t)L,
                                                            acc.p + 3 * (size_t)L, ksum, clen);
    hitsOut = totalHits;
    overOut = totalOver;
}

void computeOccupancy(tsr_context *ctx, const tsr_tree *tree, const tsr_scene *scene, const tsr_params &prm, int flags,
                      const tsr_culler *culler, tsr_field *field) {
    hipStream_t st = ctx->stream;
    Scratch &sc = ctx->scratch;
    const u32 L = tree->leafCount, N = tree->nodeCount;
    const u64 links64 = (u64)tree->sampleCount * tree->maxCams;
    if (links64 >= (u64)INVALID) throw tsr_error(TSR_ERR_LINK_OVERFLOW, "Number of traced rays exceeds maximum count. Supported maximum is 2^32-1.");
    const u32 links = (u32)links64;
    field->flags = flags;
    field->leafCount = L;
    field->nodeCount = N;
    field->ksum.alloc(2 * (size_t)L);
    field->clen.alloc(2 * (size_t)L);
    field->prior.alloc(L);
    field->states.alloc(N);
    field->occ.alloc(L);
    HIPCHECK(hipMemsetAsync(field->states.p, 0, (size_t)N * 4, st));

    TraversalTree tt{nullptr, nullptr, field->states.p, tree->frame, (float)(tree->frame.rootSize * 1e-5), nullptr};
    ConeArena &ar = ctx->cone;
    traversalLayout(st, tree, ar.childPad, ar.blockParent, tt);
    LinkInputs in{tree->px.p, tree->py.p, tree->pz.p, tree->nx.p, tree->ny.p, tree->nz.p, tree->scale.p,
                  tree->conf.p, tree->cams.p, scene->cx.p, scene->cy.p, scene->cz.p, tree->sampleDepth.p,
                  tree->maxCams, scene->camCount, prm.bandwidth_factor, prm.relative_sample_cone_length};
    in.linkStride = 1;
    in.order = nullptr;
    in.slots = 0;
    in.linkBase = 0;
    field->params = prm;
    // E-pass threads take links grouped by camera (per link chunk), samples in node
    // (Morton) order within a camera: neighbouring lanes trace nearly parallel cones
    // through the same nodes.  Records stay indexed by link, so results do not depend
    // on this order.
    const bool atomicMode = (flags & 0xff) == TSR_ACCUM_FP32_ATOMIC;

    DBuf<u32> hitCount(L);
    DBuf<double> confs(links);
    for (int pass = 0; pass < 2; ++pass) {
        const int sampleness = pass == 0 ? 1 : 0;
        const int type = sampleness;   // storage index: 0 = EMPTINESS, 1 = SAMPLENESS
        recordPhase(ctx, 1 + pass, true);
        float *ksum = field->ksum.p + (size_t)type * L, *clen = field->clen.p + (size_t)type * L;
        if (!sampleness) {   // E pass: S flags up the tree first
            k_setLeafFlags<<<gridFor(L, BLOCK), BLOCK, 0, st>>>(hitCount.p, tree->leafNode.p, L, FLAG_S, field->states.p);
            k_propagateFlag<<<gridFor(L, BLOCK), BLOCK, 0, st>>>(tree->leafNode.p, L, tree->parents.p, FLAG_S,
                                                                 field->states.p);
            // (child, state) pairs for the gated emptiness traversal, node-padded
            uint2 *packed = ar.packed.ensure((size_t)N + NODE_PAD) + NODE_PAD;
            k_packChildState<<<gridFor(N, BLOCK), BLOCK, 0, st>>>(tree->child.p, field->states.p, N, packed);
            tt.packed = packed;
        }
        // the gathered E pass marks entries in bit 31 of the link index: links < 2^31
        if (!sampleness && !atomicMode && !(flags & TSR_EMPTINESS_TRAVERSAL) && links < 0x80000000u) {
            emptinessByLeaves(ctx, tree, scene, prm, field->hitsS.p, ksum, clen, hitCount.p, confs.p,
                              field->hits[type]);
            field->overflowLinks[type] = 0;
        } else if (atomicMode) {
            DBuf<u32> counts(links), offsets(1);
            HIPCHECK(hipMemsetAsync(ksum, 0, (size_t)L * 4, st));
            HIPCHECK(hipMemsetAsync(clen, 0, (size_t)L * 4, st));
            HIPCHECK(hipMemsetAsync(hitCount.p, 0, (size_t)L * 4, st));
            k_conePass<<<xcdGrid(links, BLOCK), BLOCK, 0, st>>>(tt, in, links, sampleness, 0, counts.p, nullptr,
                                                                nullptr, nullptr, sampleness ? nullptr : confs.p,
                                                                nullptr);
            k_conePass<<<xcdGrid(links, BLOCK), BLOCK, 0, st>>>(tt, in, links, sampleness, 2, counts.p, ksum,
                                                                clen, hitCount.p, nullptr, nullptr);
            reduceSum(sc, counts.p, offsets.p, links, st);
            field->hits[type] = readScalar(offsets.p, st);
        } else {
            conePassRecords(ctx, tree, tt, in, links, sampleness, ksum, clen, hitCount.p,
                            sampleness ? nullptr : confs.p, field->hits[type], field->overflowLinks[type]);
        }
        if (!sampleness) {
            k_setLeafFlags<<<gridFor(L, BLOCK), BLOCK, 0, st>>>(hitCount.p, tree->leafNode.p, L, FLAG_E, field->states.p);
            // M: FP64 sum of confidences of non-degenerate emptiness cones (Occupancy.cpp:205-213,239-241)
            DBuf<double> m(1);
            reduceSum(sc, confs.p, m.p, links, st);
            field->coneCount = readScalar(m.p, st);
        }
        if (sampleness) {
            field->hitsS.alloc(L);
            HIPCHECK(hipMemcpyAsync(field->hitsS.p, hitCount.p, (size_t)L * 4, hipMemcpyDeviceToDevice, st));
        }
        recordPhase(ctx, 1 + pass, false);
    }
    recordPhase(ctx, 3, true);
    priors(ctx, tree, prm, flags, culler, field);
    recordPhase(ctx, 3, false);
    recordPhase(ctx, 4, true);
    k_occupancy<<<gridFor(L, BLOCK), BLOCK, 0, st>>>(field->ksum.p, field->prior.p, L, field->coneCount,
                                                     (double)prm.confidence_threshold, field->occ.p);
    recordPhase(ctx, 4, false);
    HIPCHECK(hipStreamSynchronize(st));
}

namespace {
__global__ void k_gatherSamples(const u32 *__restrict__ idx, u32 n, u32 maxCams, const float *__restrict__ px,
                                const float *__restrict__ py, const float *__restrict__ pz, const float *__restrict__ nx,
                                const float *__restrict__ ny, const float *__restrict__ nz, const float *__restrict__ sc,
                                const float *__restrict__ cf, const u32 *__restrict__ cams,
                                const uint8_t *__restrict__ depth, float *__restrict__ out, u32 *__restrict__ camsOut,
                                uint8_t *__restrict__ depthOut) {
    const u32 j = blockIdx.x * BLOCK + threadIdx.x;
    if (j >= n) return;
    const u32 i = idx[j];
    const float *src[8] = {px, py, pz, nx, ny, nz, sc, cf};
    for (int a = 0; a < 8; ++a) out[(size_t)a * n + j] = src[a][i];
    for (u32 k = 0; k < maxCams; ++k) camsOut[(size_t)j * maxCams + k] = cams[(size_t)i * maxCams + k];
    depthOut[j] = depth[i];
}
// target += sign * part (per leaf)
__global__ void k_signedAdd(float *__restrict__ ksum, float *__restrict__ clen, const float *__restrict__ pk,
                            const float *__restrict__ pl, u32 leafCount, float sign) {
    const u32 l = blockIdx.x * BLOCK + threadIdx.x;
    if (l >= leafCount) return;
    ksum[l] += sign * pk[l];
    clen[l] += sign * pl[l];
}
__global__ void k_checkSampleIndices(const u32 *__restrict__ idx, u32 n, u32 count, u32 *err) {
    const u32 j = blockIdx.x * BLOCK + threadIdx.x;
    if (j < n && idx[j] >= count) atomicOr(err, 1u);
}
}  // namespace

// Occupancy::addSamples / eraseSamples (Occupancy.h:167-175; Occupancy.cpp:129-136
// onSamplesChanged, 138-248 updateKernels over the chosen samples, Occupancy.h:222-227
// update): the S pass, then the S flags up the tree and the gated E pass, over the camera
// links of the chosen samples only; their per-leaf kernel sums, cone lengths and M are
// added to (or subtracted from) the field's.  Leaf flags are only ever set (additions), as
// in the reference.  Then priors (with the field's priors mode) and the occupancy ratio.
void updateOccupancy(tsr_context *ctx, const tsr_tree *tree, const tsr_scene *scene, const u32 *chosenHost, u32 n,
                     bool removal, tsr_field *field) {
    hipStream_t st = ctx->stream;
    Scratch &sc = ctx->scratch;
    const u32 L = tree->leafCount, N = tree->nodeCount;
    const tsr_params prm = field->params;
    const u32 maxCams = tree->maxCams;
    const u64 links64 = (u64)n * maxCams;
    if (links64 >= (u64)INVALID) throw tsr_error(TSR_ERR_LINK_OVERFLOW, "Number of traced rays exceeds maximum count. Supported maximum is 2^32-1.");
    const u32 links = (u32)links64;
    if (n) {
        DBuf<u32> idx(n), err(1);
        HIPCHECK(hipMemcpyAsync(idx.p, chosenHost, (size_t)n * 4, hipMemcpyHostToDevice, st));
        HIPCHECK(hipMemsetAsync(err.p, 0, 4, st));
        k_checkSampleIndices<<<gridFor(n, BLOCK), BLOCK, 0, st>>>(idx.p, n, tree->sampleCount, err.p);
        if (readScalar(err.p, st)) throw tsr_error(TSR_ERR_INVALID_ARGUMENT, "sample index out of range");
        DBuf<float> soa(8 * (size_t)n);
        DBuf<u32> cams(links);
        DBuf<uint8_t> depth(n);
        k_gatherSamples<<<gridFor(n, BLOCK), BLOCK, 0, st>>>(idx.p, n, maxCams, tree->px.p, tree->py.p, tree->pz.p,
                                                             tree->nx.p, tree->ny.p, tree->nz.p, tree->scale.p,
                                                             tree->conf.p, tree->cams.p, tree->sampleDepth.p, soa.p,
                                                             cams.p, depth.p);
        const float *g = soa.p;
        LinkInputs in{g, g + n, g + 2 * (size_t)n, g + 3 * (size_t)n, g + 4 * (size_t)n, g + 5 * (size_t)n,
                      g + 6 * (size_t)n, g + 7 * (size_t)n, cams.p, scene->cx.p, scene->cy.p, scene->cz.p, depth.p,
                      maxCams, scene->camCount, prm.bandwidth_factor, prm.relative_sample_cone_length};
        in.linkStride = 1;
        in.order = nullptr;
        in.slots = 0;
        in.linkBase = 0;
        TraversalTree tt{nullptr, nullptr, field->states.p, tree->frame, (float)(tree->frame.rootSize * 1e-5), nullptr};
        ConeArena &ar = ctx->cone;
        traversalLayout(st, tree, ar.childPad, ar.blockParent, tt);
        DBuf<float> pk(L), pl(L);
        DBuf<u32> hits(L);
        DBuf<double> confs(links);
        const float sign = removal ? -1.f : 1.f;
        u64 h = 0;
        u32 over = 0;
        // S pass
        conePassRecords(ctx, tree, tt, in, links, 1, pk.p, pl.p, hits.p, nullptr, h, over);
        k_signedAdd<<<gridFor(L, BLOCK), BLOCK, 0, st>>>(field->ksum.p + L, field->clen.p + L, pk.p, pl.p, L, sign);
        if (!removal) k_setLeafFlags<<<gridFor(L, BLOCK), BLOCK, 0, st>>>(hits.p, tree->leafNode.p, L, FLAG_S, field->states.p);
        field->hits[1] = removal ? field->hits[1] - std::min(field->hits[1], h) : field->hits[1] + h;
        // E pass: S flags up the tree first (Occupancy.cpp:141-142)
        k_propagateFlag<<<gridFor(L, BLOCK), BLOCK, 0, st>>>(tree->leafNode.p, L, tree->parents.p, FLAG_S,
                                                             field->states.p);
        uint2 *packed = ar.packed.ensure((size_t)N + NODE_PAD) + NODE_PAD;
        k_packChildState<<<gridFor(N, BLOCK), BLOCK, 0, st>>>(tree->child.p, field->states.p, N, packed);
        tt.packed = packed;
        conePassRecords(ctx, tree, tt, in, links, 0, pk.p, pl.p, hits.p, confs.p, h, over);
        k_signedAdd<<<gridFor(L, BLOCK), BLOCK, 0, st>>>(field->ksum.p, field->clen.p, pk.p, pl.p, L, sign);
        if (!removal) k_setLeafFlags<<<gridFor(L, BLOCK), BLOCK, 0, st>>>(hits.p, tree->leafNode.p, L, FLAG_E, field->states.p);
        field->hits[0] = removal ? field->hits[0] - std::min(field->hits[0], h) : field->hits[0] + h;
        DBuf<double> m(1);
        reduceSum(sc, confs.p, m.p, links, st);
        const double dm = readScalar(m.p, st);
        field->coneCount += removal ? -dm : dm;
    }
    // priors and occupancy of the changed field (a learned field continues without the network)
    int pf = field->flags & (TSR_PRIORS_ALL_LEAVES | TSR_PRIORS_CERTIFIED);
    if (field->flags & TSR_CULL_LEARNED) pf |= TSR_PRIORS_CERTIFIED;
    priors(ctx, tree, prm, pf, nullptr, field);
    k_occupancy<<<gridFor(L, BLOCK), BLOCK, 0, st>>>(field->ksum.p, field->prior.p, L, field->coneCount,
                                                     (double)prm.confidence_threshold, field->occ.p);
    HIPCHECK(hipStreamSynchronize(st));
}

void occupancyValues(tsr_context *ctx, const tsr_field *field, double confidenceThreshold) {
    const u32 L = field->leafCount;
    if (L == 0) return;
    tsr_field *fw = const_cast<tsr_field *>(field);
    if (fw->occ.n < L) fw->occ.alloc(L);
    k_occupancy<<<gridFor(L, BLOCK), BLOCK, 0, ctx->stream>>>(field->ksum.p, field->prior.p, L, field->coneCount,
                                                              confidenceThreshold, fw->occ.p);
    HIPCHECK(hipStreamSynchronize(ctx->stream));
}

__global__ void k_calibrationCopy(const float4 *__restrict__ src, float4 *__restrict__ dst, size_t n) {
    for (size_t i = (size_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (size_t)gridDim.x * BLOCK) dst[i] = src[i];
}

void calibrationCopy(tsr_context *ctx, u64 bytes) {
    const size_t n = (size_t)((bytes + 15) / 16);
    DBuf<float4> a(n), b(n);
    HIPCHECK(hipMemsetAsync(a.p, 0, n * 16, ctx->stream));
    k_calibrationCopy<<<std::min<size_t>(gridFor(n, BLOCK), 256 * 64), BLOCK, 0, ctx->stream>>>(a.p, b.p, n);
    HIPCHECK(hipStreamSynchronize(ctx->stream));
}

}  // namespace tsr
