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
#include <vector>

#include "handles.h"

namespace tsr {

namespace {

constexpr u32 BLOCK = 256;

struct TraversalTree {
    const u32 *child;      // first child id, or LEAF_BIT | leaf index
    const u32 *parents;
    const u32 *states;     // node state flags (E pass gate), may be null
    Frame f;
    float margin;          // absolute culling margin
};

struct LinkInputs {
    const float *px, *py, *pz, *nx, *ny, *nz, *scale, *conf;
    const u32 *cams;
    const float *cx, *cy, *cz;
    const uint8_t *sampleDepth;
    u32 maxCams, camCount;
    float bandwidth, relConeLength;
    const uint8_t *cullNode;   // learned culling mask (E pass), may be null
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
// calls onLeaf(leafIndex, centre) for every leaf that passes depth, flag and
// cull tests.  Inner nodes failing a test are skipped with their subtrees.
template <typename OnLeaf>
__device__ void traverseCone(const TraversalTree &t, const ConeCull &cull, u32 maxDepth, u32 required,
                             const uint8_t *cullNode, OnLeaf onLeaf) {
    u32 node = 0, d = 0, k[3] = {0, 0, 0};
    bool descend = true;   // whether to test `node` (true) or move on from it (false)
    for (;;) {
        if (descend) {
            bool accept = d <= maxDepth;
            if (accept && required) accept = (t.states[node] & required) == required;
            if (accept && cullNode) accept = cullNode[node] == 0;
            const u32 c = accept ? t.child[node] : 0u;
            if (accept) {
                if (c & LEAF_BIT) {
                    double x[3];
                    for (int a = 0; a < 3; ++a) x[a] = t.f.B(a, 2ull * k[a] + 1, d + 1);
                    onLeaf(c & ~LEAF_BIT, x);
                } else if (c != INVALID) {
                    const float h = (float)t.f.h(d);
                    const float lo0 = t.f.rootMinF[0] + (float)k[0] * h, lo1 = t.f.rootMinF[1] + (float)k[1] * h,
                                lo2 = t.f.rootMinF[2] + (float)k[2] * h;
                    if (!cullSeparated(cull, lo0, lo1, lo2, h, t.margin)) {
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
            node = t.parents[node];
            --d;
            k[0] >>= 1; k[1] >>= 1; k[2] >>= 1;
        }
    }
}

// mode 0: count accepted hits per link; 1: write records; 2: FP32 atomics (precision study)
__global__ void k_conePass(TraversalTree t, LinkInputs in, u32 linkCount, int sampleness, int mode,
                           u32 *__restrict__ counts, const u32 *__restrict__ offsets, u32 *__restrict__ recLeaf,
                           uint2 *__restrict__ recVal, float *__restrict__ atomicSum, float *__restrict__ atomicLen,
                           u32 *__restrict__ atomicStates, double *__restrict__ confOut) {
    const u32 link = blockIdx.x * BLOCK + threadIdx.x;
    if (link >= linkCount) return;
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
    u32 count = 0;
    const u32 base = mode == 1 ? offsets[link] : 0;
    traverseCone(t, cull, maxDepth, required, sampleness ? nullptr : in.cullNode, [&](u32 leaf, const double x[3]) {
        const double kv = coneKernel(cone, x, sampleness != 0, conf);
        if (kv == 0.0) return;
        if (mode == 1) {
            recLeaf[base + count] = leaf;
            recVal[base + count] = make_uint2(__float_as_uint((float)kv), __float_as_uint(weightedLength));
        } else if (mode == 2) {
            atomicAdd(&atomicSum[leaf], (float)kv);
            atomicAdd(&atomicLen[leaf], weightedLength);
            atomicOr(&atomicStates[leaf], 1u);
        }
        ++count;
    });
    if (mode == 0) counts[link] = count;
}

// first record of every leaf segment in the sorted record stream
__global__ void k_segmentStarts(const u32 *__restrict__ leafSorted, size_t n, u32 *__restrict__ segStart) {
    const size_t j = (size_t)blockIdx.x * BLOCK + threadIdx.x;
    if (j >= n) return;
    const u32 l = leafSorted[j];
    if (j == 0 || leafSorted[j - 1] != l) segStart[l] = (u32)j;
}

// Kahan sums of each leaf's records in record order (deterministic)
__global__ void k_segmentReduce(const u32 *__restrict__ leafSorted, const uint2 *__restrict__ valSorted, size_t n,
                                const u32 *__restrict__ segStart, u32 leafCount, float *__restrict__ ksum,
                                float *__restrict__ clen, u32 *__restrict__ hitCount) {
    const u32 l = blockIdx.x * BLOCK + threadIdx.x;
    if (l >= leafCount) return;
    const u32 s = segStart[l];
    float sk = 0.f, ck = 0.f, sl = 0.f, cl = 0.f;
    u32 cnt = 0;
    if (s != INVALID) {
        for (size_t j = s; j < n && leafSorted[j] == l; ++j, ++cnt) {
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
TSR_HD void twoSumAdd(float &hi, float &lo, float x) {   // hi + lo += x (Kahan-Babuska)
    const float s = hi + x;
    const float bp = s - hi;
    const float err = (hi - (s - bp)) + (x - bp);
    hi = s;
    lo += err;
}

// subtree sums (hi, lo) of E and S cone lengths and deepest leaf, one depth level at a time
__global__ void k_subtreeLeaves(const u32 *__restrict__ leafNode, u32 leafCount, const float *__restrict__ clen,
                                const uint8_t *__restrict__ depth, float4 *__restrict__ sums,
                                uint8_t *__restrict__ maxLeafDepth) {
    const u32 l = blockIdx.x * BLOCK + threadIdx.x;
    if (l >= leafCount) return;
    const u32 nd = leafNode[l];
    sums[nd] = make_float4(clen[l], 0.f, clen[leafCount + l], 0.f);
    maxLeafDepth[nd] = depth[nd];
}

__global__ void k_subtreeLevel(const u32 *__restrict__ child, const uint8_t *__restrict__ depth, u32 nodeCount,
                               u32 level, float4 *__restrict__ sums, uint8_t *__restrict__ maxLeafDepth) {
    const u32 nd = blockIdx.x * BLOCK + threadIdx.x;
    if (nd >= nodeCount || depth[nd] != level) return;
    const u32 c = child[nd];
    if (c & LEAF_BIT) return;
    float eh = 0.f, el = 0.f, sh = 0.f, sl = 0.f;
    uint8_t md = 0;
    for (u32 s = 0; s < 8; ++s) {
        const float4 v = sums[c + s];
        twoSumAdd(eh, el, v.x);
        el += v.y;
        twoSumAdd(sh, sl, v.z);
        sl += v.w;
        md = max(md, maxLeafDepth[c + s]);
    }
    sums[nd] = make_float4(eh, el, sh, sl);
    maxLeafDepth[nd] = md;
}

__global__ void k_priors(TraversalTree t, const u32 *__restrict__ leafNode, const u64 *__restrict__ leafLoc,
                         u32 leafCount, const uint8_t *__restrict__ depthArr, const float *__restrict__ clen,
                         const float4 *__restrict__ sums, const uint8_t *__restrict__ maxLeafDepth,
                         float relRadius, float *__restrict__ prior) {
    const u32 l = blockIdx.x * BLOCK + threadIdx.x;
    if (l >= leafCount) return;
    const u64 loc = leafLoc[l];
    u32 kq[3];
    locCoords(loc, kq[0], kq[1], kq[2]);
    const u32 dq = locDepth(loc);
    double c[3];
    for (int a = 0; a < 3; ++a) c[a] = t.f.centre(a, kq[a], dq);
    const double radius = t.f.h(dq) * (double)relRadius;
    const double r2 = radius * radius, rInner2 = r2 * (1.0 - 1e-9);
    const u32 maxDepth = dq + MAX_DEPTH_DIFFERENCE;
    float eh = 0.f, el = 0.f, sh = 0.f, sl = 0.f;
    u32 node = 0, d = 0, k[3] = {0, 0, 0};
    bool test = true;
    for (;;) {
        if (test && d <= maxDepth && (t.states[node] & FLAG_S)) {
            // SphereNodesChecker.cpp:29-40 (AABB/sphere, squared distance <= r^2)
            double dmin = 0.0, dmax = 0.0;
            for (int a = 0; a < 3; ++a) {
                const double lo = t.f.B(a, k[a], d), hi = t.f.B(a, (u64)k[a] + 1, d);
                if (c[a] < lo) dmin += (lo - c[a]) * (lo - c[a]);
                else if (c[a] > hi) dmin += (c[a] - hi) * (c[a] - hi);
                const double far = fmax(fabs(c[a] - lo), fabs(hi - c[a]));
                dmax += far * far;
            }
            if (dmin <= r2) {
                const u32 ch = t.child[node];
                if (ch & LEAF_BIT) {
                    const u32 j = ch & ~LEAF_BIT;
                    twoSumAdd(eh, el, clen[j]);
                    twoSumAdd(sh, sl, clen[leafCount + j]);
                } else if (dmax <= rInner2 && maxLeafDepth[node] <= maxDepth) {
                    const float4 v = sums[node];
                    twoSumAdd(eh, el, v.x);
                    el += v.y;
                    twoSumAdd(sh, sl, v.z);
                    sl += v.w;
                } else {
                    node = ch;
                    ++d;
                    k[0] *= 2; k[1] *= 2; k[2] *= 2;
                    continue;
                }
            }
        }
        test = true;
        for (;;) {
            if (node == 0) goto done;
            const u32 slot = (node - 1) & 7;
            if (slot < 7) {
                ++node;
                const u32 s1 = slot + 1;
                k[0] = (k[0] & ~1u) | (s1 & 1);
                k[1] = (k[1] & ~1u) | ((s1 >> 1) & 1);
                k[2] = (k[2] & ~1u) | ((s1 >> 2) & 1);
                break;
            }
            node = t.parents[node];
            --d;
            k[0] >>= 1; k[1] >>= 1; k[2] >>= 1;
        }
    }
done:
    const double E = (double)eh + (double)el, S = (double)sh + (double)sl;
    prior[l] = (float)(E / (E + S));
}

// Occupancy.cpp:570-587, FP64
__global__ void k_occupancy(const float *__restrict__ ksum, const float *__restrict__ prior, u32 leafCount,
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

__global__ void k_fillF(float *p, size_t n, float v) {
    const size_t i = (size_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void k_fillU(u32 *p, size_t n, u32 v) {
    const size_t i = (size_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) p[i] = v;
}


}  // namespace

void computeOccupancy(tsr_context *ctx, const tsr_tree *tree, const tsr_scene *scene, const tsr_params &prm, int flags,
                      const uint8_t *cullMask, tsr_field *field) {
    hipStream_t st = ctx->stream;
    Scratch &sc = ctx->scratch;
    const u32 L = tree->leafCount, N = tree->nodeCount;
    const u64 links64 = (u64)scene->n * scene->maxCams;
    if (links64 >= (u64)INVALID) throw tsr_error(TSR_ERR_LINK_OVERFLOW, "Number of traced rays exceeds maximum count. Supported maximum is 2^32-1.");
    const u32 links = (u32)links64;
    field->leafCount = L;
    field->nodeCount = N;
    field->ksum.alloc(2 * (size_t)L);
    field->clen.alloc(2 * (size_t)L);
    field->prior.alloc(L);
    field->states.alloc(N);
    field->occ.alloc(L);
    HIPCHECK(hipMemsetAsync(field->states.p, 0, (size_t)N * 4, st));

    TraversalTree tt{tree->child.p, tree->parents.p, field->states.p, tree->frame,
                     (float)(tree->frame.rootSize * 1e-5)};
    LinkInputs in{scene->px.p, scene->py.p, scene->pz.p, scene->nx.p, scene->ny.p, scene->nz.p, scene->scale.p,
                  scene->conf.p, scene->cams.p, scene->cx.p, scene->cy.p, scene->cz.p, tree->sampleDepth.p,
                  scene->maxCams, scene->camCount, prm.bandwidth_factor, prm.relative_sample_cone_length, nullptr};
    in.cullNode = cullMask;
    const bool atomicMode = (flags & 0xff) == TSR_ACCUM_FP32_ATOMIC;

    DBuf<u32> counts(links), offsets(links), hitCount(L), segStart(L);
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
        }
        if (atomicMode) {
            HIPCHECK(hipMemsetAsync(ksum, 0, (size_t)L * 4, st));
            HIPCHECK(hipMemsetAsync(clen, 0, (size_t)L * 4, st));
            HIPCHECK(hipMemsetAsync(hitCount.p, 0, (size_t)L * 4, st));
            k_conePass<<<gridFor(links, BLOCK), BLOCK, 0, st>>>(tt, in, links, sampleness, 0, counts.p, nullptr,
                                                                nullptr, nullptr, nullptr, nullptr, nullptr,
                                                                sampleness ? nullptr : confs.p);
            k_conePass<<<gridFor(links, BLOCK), BLOCK, 0, st>>>(tt, in, links, sampleness, 2, counts.p, nullptr,
                                                                nullptr, nullptr, ksum, clen, hitCount.p, nullptr);
            size_t tmp = 0;
            HIPCHECK(hipcub::DeviceReduce::Sum(nullptr, tmp, counts.p, offsets.p, (int)links, st));
            HIPCHECK(hipcub::DeviceReduce::Sum(sc.get(tmp), tmp, counts.p, offsets.p, (int)links, st));
            field->hits[type] = readScalar(offsets.p, st);
        } else {
            k_conePass<<<gridFor(links, BLOCK), BLOCK, 0, st>>>(tt, in, links, sampleness, 0, counts.p, nullptr,
                                                                nullptr, nullptr, nullptr, nullptr, nullptr,
                                                                sampleness ? nullptr : confs.p);
            exclusiveScan(sc, counts.p, offsets.p, links, st);
            const u64 total = (u64)readScalar(offsets.p + links - 1, st) + readScalar(counts.p + links - 1, st);
            if (total >= (u64)INVALID) throw tsr_error(TSR_ERR_OUT_OF_MEMORY, "too many kernel contributions in one pass");
            field->hits[type] = total;
            DBuf<u32> recLeaf(std::max<u64>(total, 1)), recLeafS(std::max<u64>(total, 1));
            DBuf<uint2> recVal(std::max<u64>(total, 1)), recValS(std::max<u64>(total, 1));
            k_conePass<<<gridFor(links, BLOCK), BLOCK, 0, st>>>(tt, in, links, sampleness, 1, counts.p, offsets.p,
                                                                recLeaf.p, recVal.p, nullptr, nullptr, nullptr,
                                                                nullptr);
            int bits = 1;
            while (bits < 32 && (1ull << bits) < L) ++bits;
            sortPairs(sc, recLeaf.p, recLeafS.p, recVal.p, recValS.p, total, 0, bits, st);
            k_fillU<<<gridFor(L, BLOCK), BLOCK, 0, st>>>(segStart.p, L, INVALID);
            k_segmentStarts<<<gridFor(total, BLOCK), BLOCK, 0, st>>>(recLeafS.p, total, segStart.p);
            k_segmentReduce<<<gridFor(L, BLOCK), BLOCK, 0, st>>>(recLeafS.p, recValS.p, total, segStart.p, L, ksum,
                                                                 clen, hitCount.p);
        }
        if (!sampleness) {
            k_setLeafFlags<<<gridFor(L, BLOCK), BLOCK, 0, st>>>(hitCount.p, tree->leafNode.p, L, FLAG_E, field->states.p);
            // M: FP64 sum of confidences of non-degenerate emptiness cones (Occupancy.cpp:205-213,239-241)
            DBuf<double> m(1);
            size_t tmp = 0;
            HIPCHECK(hipcub::DeviceReduce::Sum(nullptr, tmp, confs.p, m.p, (int)links, st));
            HIPCHECK(hipcub::DeviceReduce::Sum(sc.get(tmp), tmp, confs.p, m.p, (int)links, st));
            field->coneCount = readScalar(m.p, st);
        }
        recordPhase(ctx, 1 + pass, false);
    }
    recordPhase(ctx, 3, true);
    // priors
    {
        DBuf<float4> sums(N);
        DBuf<uint8_t> maxLeafDepth(N);
        k_subtreeLeaves<<<gridFor(L, BLOCK), BLOCK, 0, st>>>(tree->leafNode.p, L, field->clen.p, tree->nodeDepth.p,
                                                             sums.p, maxLeafDepth.p);
        for (int lvl = (int)tree->maxDepth - 1; lvl >= 0; --lvl)
            k_subtreeLevel<<<gridFor(N, BLOCK), BLOCK, 0, st>>>(tree->child.p, tree->nodeDepth.p, N, (u32)lvl, sums.p,
                                                                maxLeafDepth.p);
        k_priors<<<gridFor(L, BLOCK), BLOCK, 0, st>>>(tt, tree->leafNode.p, tree->leafLoc.p, L, tree->nodeDepth.p,
                                                      field->clen.p, sums.p, maxLeafDepth.p,
                                                      prm.relative_radius_for_priors, field->prior.p);
    }
    recordPhase(ctx, 3, false);
    recordPhase(ctx, 4, true);
    k_occupancy<<<gridFor(L, BLOCK), BLOCK, 0, st>>>(field->ksum.p, field->prior.p, L, field->coneCount,
                                                     (double)prm.confidence_threshold, field->occ.p);
    recordPhase(ctx, 4, false);
    HIPCHECK(hipStreamSynchronize(st));
}

}  // namespace tsr
