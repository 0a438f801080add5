// Emptiness pass gathered by leaves (Occupancy::updateKernels for EMPTINESS,
// SR/SurfaceExtraction/Occupancy.cpp:138-248, with addKernel :279-346).
//
// The reference walks the tree once per camera->sample link (LeavesIterator with a
// ConeNodesChecker gated by the sampleness flag and the link's depth limit) and adds the
// cone kernel at every leaf it meets.  Node flags are ORed up the tree
// (Occupancy.cpp:491-516) and the cone test only removes nodes that cannot hold a
// contributing leaf centre, so the set of (link, leaf) contributions is fixed by the leaf
// alone: a sampleness-flagged leaf, no deeper than the link's depth limit, whose centre
// gets a non-zero kernel.  This pass enumerates that set from the leaves' side.
//
// A centre x gets a non-zero emptiness kernel only inside the oblique cone from the camera
// centre a to the sample's end disc (centre e, radius R = bandwidth * scale): x = a + t (q - a)
// with q in the disc and t in (0, 1), so x - a lies within asin(R / |e - a|) of the cone axis
// and |x - a| < |e - a| + R.  Per camera the links are binned by the directions of their
// cones on the six faces of a cube map (gnomonic coordinates w_j / |w_i| on the face of the
// largest component i, a cell grid sized to the camera's mean cone angle).  A leaf looks up,
// per camera, the one cell its direction falls into, tests that cell's links with an FP32
// cone test whose margins exceed its rounding, and evaluates the reference's FP64 kernel
// (coneKernel) for those that pass.  Links whose cones cover too many cells (cameras very
// close to a sample) are listed per camera and tested by every leaf.
//
// Each leaf's contributions are summed by one wavefront in a fixed order (cameras in
// batches of 64, a cell's links in link order, then the camera's wide links; Kahan per
// lane, then a fixed two-sum tree over the lanes): deterministic, with no records and no
// sort by leaf.
#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "handles.h"

namespace tsr {

constexpr u32 FACES = 6;   // cube-map faces per camera

struct FaceGrid {
    int lo[2];       // first cell (u, v) of the bounding box of the face's links
    u32 dim[2];      // cells of the box (0: no links on the face)
    u32 base;        // index of the box's first cell in the cell arrays
};

struct CamGrid {
    float a[3];      // camera centre
    float inv;       // cells per unit of gnomonic coordinate (G / 2)
    u32 G;           // cells along a face side
    u32 wideBegin, wideEnd;
    FaceGrid face[FACES];
};

// The direction index of the gathered E pass over one set of links (emptinessByLeaves; built
// by buildEmptinessIndex): per-link FP64-stage records, per-camera cube-map grids, entries
// sorted by cell and by descending reach within a cell with their test records, the wide links
// per camera.
struct EmptinessIndex {
    u32 links = 0, camCount = 0, nWide = 0;
    float bandwidth = 0.f;
    unsigned long long sceneId = 0;
    u64 totalCells = 0, totalEntries = 0;
    DBuf<float4> eval;                 // EVAL_WORDS per link
    DBuf<CamGrid> grids;               // per camera
    DBuf<u32> cellStart, cellEnd;      // per cell: run of its entries
    DBuf<float> cellReach;             // per cell: largest reach
    DBuf<float> entryReach;            // per entry
    DBuf<float4> rec, wideRec;         // test records per entry / per wide link
};

namespace {

constexpr u32 BLOCK = 256;
constexpr u32 XCD_COUNT = 8;
constexpr u32 WIDE_CELLS = 4096;        // a link whose cone covers more cells is listed per camera
constexpr u32 MAX_FACE_CELLS = 1u << 15;  // cells along one side of a face grid
constexpr u32 LEAVES_PER_WAVE = 4;      // consecutive leaves per wavefront of the leaf pass
constexpr u32 CAND_UNROLL = 4;          // candidates per lane and step of the leaf pass
constexpr u32 GATHER_WAVES = 8;         // wavefronts per block of the leaf pass
constexpr u32 CAM_LDS = 256;            // cameras whose grids the leaf pass keeps in LDS
constexpr double REAL_MAX_D = 3.4028234663852886e38;

// link classes
constexpr uint8_t LINK_NONE = 0, LINK_GRID = 1, LINK_WIDE = 2;
constexpr u32 ENTRY_OWN_FACE = 0x80000000u;   // entry value bit: the cell is on the cone axis's face
constexpr float NO_DISC = 1e30f;               // disc radius of entries without a face-coordinate test
constexpr u32 EVAL_WORDS = 4;                  // float4s per link of LinkGeo::eval


// per-link data of the leaf pass, in link order
struct LinkGeo {
    float4 *axis;    // unit cone axis (e - a) / |e - a|, tan of the inflated cone angle (inf: wide)
    float4 *aux;     // reach (|e - a| + R, inflated), inflated cone angle (radians), depth limit, 0
    // the FP64 kernel's data, one 64-byte line per link (EVAL_WORDS float4s): sample position
    // and scale; sample normal and confidence; 1 / dot(e - a, n) and |e - a| (ConeD::invDenom,
    // ConeD::length) as two doubles; padding
    float4 *eval;
};


// confidence of each link's emptiness cone, 0 when it has none (invalid camera or degenerate
// cone): the terms of the cone count M (Occupancy.cpp:205-213,239-241)
__global__ void k_linkConfs(EmptinessLinks s, u32 links, double *__restrict__ confs) {
    const u32 l = blockIdx.x * BLOCK + threadIdx.x;
    if (l >= links) return;
    const u32 i = l / s.maxCams, cam = s.cams[l];
    double conf = 0.0;
    if (cam < s.camCount) {
        const double apex[3] = {s.cx[cam], s.cy[cam], s.cz[cam]};
        const double p[3] = {s.px[i], s.py[i], s.pz[i]};
        const double nrm[3] = {-(double)s.nx[i], -(double)s.ny[i], -(double)s.nz[i]};
        ConeD c;
        coneInit(c, apex, nrm, p, (double)s.bandwidth * (double)s.scale[i], REAL_MAX_D);
        if (!c.degenerate) conf = s.conf[i];
    }
    confs[l] = conf;
}

// Cone of link l as the traversal sets it up (Occupancy.cpp:174-201, getCameraToSampleLinkData
// :250-277): apex at the camera, end at the sample, end normal -n, R = bandwidth * scale,
// uncut.  Writes the link's leaf-pass data and its confidence (0: not an emptiness cone).
__global__ void k_gLinkSetup(EmptinessLinks s, u32 links, LinkGeo g, uint8_t *__restrict__ cls,
                             u32 *__restrict__ camKey, u32 *__restrict__ reachKey, u32 *__restrict__ iota) {
    const u32 l = blockIdx.x * BLOCK + threadIdx.x;
    if (l >= links) return;
    cls[l] = LINK_NONE;
    camKey[l] = s.camCount;   // sort keys: links without a cone last
    reachKey[l] = 0xffffffffu;
    iota[l] = l;
    const u32 i = l / s.maxCams, cam = s.cams[l];
    if (cam >= s.camCount) return;   // Cameras::isValid
    const double apex[3] = {s.cx[cam], s.cy[cam], s.cz[cam]};
    const double p[3] = {s.px[i], s.py[i], s.pz[i]};
    const double nrm[3] = {-(double)s.nx[i], -(double)s.ny[i], -(double)s.nz[i]};
    const double R = (double)s.bandwidth * (double)s.scale[i];
    ConeD c;
    coneInit(c, apex, nrm, p, R, REAL_MAX_D);
    if (c.degenerate) return;
    const double A = c.length;
    const bool wide = !(R < 0.5 * A);
    // tan of the cone angle asin(R / A), which also bounds the angle from above; inflated:
    // relative 1e-4 plus 1e-5 rad cover the FP32 rounding of the unit axis and of the face
    // coordinate ranges
    const double tanA = wide ? 0.0 : R / sqrt(A * A - R * R);
    const double thetaP = tanA * (1.0 + 1e-4) + 1e-5;
    const double tanT = wide ? (double)INFINITY : tanA * (1.0 + 1e-4);
    const double invA = 1.0 / A;
    g.axis[l] = make_float4((float)(c.v[0] * invA), (float)(c.v[1] * invA), (float)(c.v[2] * invA), (float)tanT);
    g.aux[l] = make_float4((float)((A + R) * (1.0 + 1e-6)), (float)thetaP,
                           (float)(s.sampleDepth[i] + MAX_DEPTH_DIFFERENCE), 0.f);
    float4 *ev = g.eval + EVAL_WORDS * (size_t)l;
    ev[0] = make_float4(s.px[i], s.py[i], s.pz[i], s.scale[i]);
    ev[1] = make_float4(s.nx[i], s.ny[i], s.nz[i], s.conf[i]);
    reinterpret_cast<double2 *>(ev)[2] = make_double2(c.invDenom, A);
    cls[l] = wide ? LINK_WIDE : LINK_GRID;
    camKey[l] = cam;
    reachKey[l] = ~__float_as_uint(g.aux[l].x);   // descending reach (positive floats order as bits)
}

// per camera, the sum of the grid links' inflated angles (in units of 2^-32 rad, integer:
// the sum does not depend on the order) and their count; links taken in camera order, one
// atomic per wavefront and camera
__global__ void k_camTheta(const u32 *__restrict__ byCam, u32 links, const uint8_t *__restrict__ cls,
                           const u32 *__restrict__ cams, const float4 *__restrict__ aux,
                           unsigned long long *__restrict__ camTheta) {
    const u32 t = blockIdx.x * BLOCK + threadIdx.x;
    const u32 l = t < links ? byCam[t] : 0u;
    bool has = t < links && cls[l] == LINK_GRID;
    const u32 cam = has ? cams[l] : 0u;
    const unsigned long long th = has ? (unsigned long long)((double)aux[l].y * 4294967296.0) : 0ull;
    for (u64 pending = __ballot(has); pending; pending = __ballot(has)) {
        const int leader = __ffsll((long long)pending) - 1;
        const u32 lc = __shfl(cam, leader);
        const bool mine = has && cam == lc;
        unsigned long long sum = mine ? th : 0ull, cnt = mine ? 1ull : 0ull;
        for (int off = 32; off > 0; off >>= 1) {
            sum += __shfl_xor(sum, off);
            cnt += __shfl_xor(cnt, off);
        }
        if ((int)(threadIdx.x & 63) == leader) {
            atomicAdd(&camTheta[2 * lc], sum);
            atomicAdd(&camTheta[2 * lc + 1], cnt);
        }
        if (mine) has = false;
    }
}

__device__ __forceinline__ int cellOf(float u, float inv, int G) {
    const int c = (int)floorf((u + 1.0f) * inv);
    return c < 0 ? 0 : (c >= G ? G - 1 : c);
}

// Cells of face F (axis F / 2, sign + for even F) that a cone of unit axis h and inflated
// angle tp can reach.  A direction w within tp of h differs from h by at most tp in every
// component; on face F its axis component is the largest and >= 1/sqrt(3), so its face
// coordinates w_j / w_i lie in the quotient intervals below (padded by 1e-5).
__device__ bool faceRange(const float h[3], float tp, u32 F, float inv, int G, int r[4]) {
    const u32 ax = F >> 1;
    const float sg = (F & 1) ? -1.f : 1.f;
    const u32 j = ax == 0 ? 1 : 0, k = ax == 2 ? 1 : 2;
    const float hi = sg * h[ax];
    const float m = fmaxf(fabsf(h[j]), fabsf(h[k]));
    if (hi + 2.f * tp + 1e-5f < m) return false;
    const float dlo = fmaxf(hi - tp, 0.57735027f - 1e-4f), dhi = hi + tp;
    if (dhi < dlo) return false;
    const float num[2] = {h[j], h[k]};
    for (int c = 0; c < 2; ++c) {
        const float nlo = num[c] - tp, nhi = num[c] + tp;
        const float lo = fmaxf(fminf(nlo / dlo, nlo / dhi) - 1e-5f, -1.f);
        const float up = fminf(fmaxf(nhi / dlo, nhi / dhi) + 1e-5f, 1.f);
        if (lo > up) return false;
        r[2 * c] = cellOf(lo, inv, G);
        r[2 * c + 1] = cellOf(up, inv, G);
    }
    return true;
}

// On the face of the cone axis h itself (the largest component of h, at angle phi0 from the
// face axis, cos phi0 = |h_i|), the face coordinates of a direction within tp of h lie within
// rho = tp / cos^2(phi0 + tp) of those of h: the gnomonic map stretches lengths by at most
// 1 / cos^2 of the angle from the face axis along a path inside the cap.  Cells of the box
// farther than rho from (u0, v0) are then skipped.  Valid while phi0 + tp stays below
// 0.95 rad (then the box alone is used).
struct CapDisc {
    u32 face;
    float u0, v0, rho;
    bool valid;
};
__device__ CapDisc capDisc(const float h[3], float tp) {
    CapDisc c;
    const float b0 = fabsf(h[0]), b1 = fabsf(h[1]), b2 = fabsf(h[2]);
    float m, u, v;
    if (b0 >= b1 && b0 >= b2) { c.face = h[0] >= 0.f ? 0 : 1; m = b0; u = h[1]; v = h[2]; }
    else if (b1 >= b2) { c.face = h[1] >= 0.f ? 2 : 3; m = b1; u = h[0]; v = h[2]; }
    else { c.face = h[2] >= 0.f ? 4 : 5; m = b2; u = h[0]; v = h[1]; }
    const float phi = acosf(fminf(m, 1.f)) + tp;
    c.valid = phi < 0.95f;
    const float cp = cosf(phi);
    c.u0 = u / m;
    c.v0 = v / m;
    c.rho = tp / (cp * cp) * (1.f + 1e-4f) + 1e-5f;
    return c;
}
// whether cell (cu, cv) of a grid with `inv` cells per unit may meet the disc
__device__ __forceinline__ bool cellMeetsDisc(const CapDisc &c, int cu, int cv, float inv) {
    const float cs = 1.f / inv;
    const float lo0 = (float)cu * cs - 1.f, lo1 = (float)cv * cs - 1.f;
    const float du = fmaxf(fmaxf(lo0 - c.u0, c.u0 - (lo0 + cs)), 0.f);
    const float dv = fmaxf(fmaxf(lo1 - c.v0, c.v0 - (lo1 + cs)), 0.f);
    return du * du + dv * dv <= c.rho * c.rho * (1.f + 1e-5f) + 1e-12f;
}

// wave minimum / maximum of v (every lane gets the result)
__device__ __forceinline__ int waveMin(int v) {
    for (int off = 32; off > 0; off >>= 1) v = min(v, __shfl_xor(v, off));
    return v;
}
__device__ __forceinline__ int waveMax(int v) {
    for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off));
    return v;
}

// Thread t takes link order[t].  mode 0 (links grouped by camera): cells per link (links over
// WIDE_CELLS become wide) and the per-(camera, face) bounding boxes of the cells, one atomic
// per wavefront and camera; mode 1 (links by descending reach): the (cell, link) entries at
// offset[t], so that a stable sort by cell keeps each cell's links by descending reach.
__global__ void k_gLinkCells(u32 links, int mode, const u32 *__restrict__ order, const u32 *__restrict__ cams,
                             const float4 *__restrict__ axis, const float4 *__restrict__ aux,
                             const CamGrid *__restrict__ grids, const u32 *__restrict__ camG,
                             uint8_t *__restrict__ cls, u32 *__restrict__ count, int *__restrict__ box,
                             const u32 *__restrict__ offset, u32 *__restrict__ keys, u32 *__restrict__ vals) {
    const u32 t = blockIdx.x * BLOCK + threadIdx.x;
    const u32 l = t < links ? order[t] : 0u;
    bool valid = t < links;
    if (valid && mode == 0) count[l] = 0;
    if (valid && cls[l] == LINK_NONE) valid = false;
    u32 cam = 0;
    float h[3] = {0.f, 0.f, 0.f};
    float tp = 0.f, inv = 1.f;
    int G = 1;
    bool isInfTan = false;
    if (valid) {
        cam = cams[l];
        const float4 ax = axis[l];
        h[0] = ax.x; h[1] = ax.y; h[2] = ax.z;
        isInfTan = isinf(ax.w);
        tp = aux[l].y;
        G = (int)camG[cam];
        inv = 0.5f * (float)G;
    }
    if (mode == 0) {
        if (valid && cls[l] == LINK_WIDE && !isInfTan) cls[l] = LINK_GRID;   // re-planned with other cells
        if (valid && cls[l] == LINK_WIDE) valid = false;
        int r[4];
        if (valid) {
            const CapDisc disc = capDisc(h, tp);
            u64 cells = 0;
            for (u32 F = 0; F < FACES; ++F) {
                if (!faceRange(h, tp, F, inv, G, r)) continue;
                const u64 box = (u64)(r[1] - r[0] + 1) * (u64)(r[3] - r[2] + 1);
                if (F == disc.face && disc.valid && box <= WIDE_CELLS) {
                    for (int v = r[2]; v <= r[3]; ++v)
                        for (int u = r[0]; u <= r[1]; ++u) cells += cellMeetsDisc(disc, u, v, inv) ? 1u : 0u;
                } else {
                    cells += box;
                }
            }
            if (cells > WIDE_CELLS) {
                cls[l] = LINK_WIDE;
                valid = false;
            } else {
                count[l] = (u32)cells;
            }
        }
        for (u32 F = 0; F < FACES; ++F) {
            bool has = valid && faceRange(h, tp, F, inv, G, r);
            // lanes of one camera at a time: their box reduced over the wavefront first
            for (u64 pending = __ballot(has); pending; pending = __ballot(has)) {
                const int leader = __ffsll((long long)pending) - 1;
                const u32 lc = __shfl(cam, leader);
                const bool mine = has && cam == lc;
                const int b0 = waveMin(mine ? r[0] : INT_MAX), b1 = waveMin(mine ? r[2] : INT_MAX);
                const int b2 = waveMax(mine ? r[1] : INT_MIN), b3 = waveMax(mine ? r[3] : INT_MIN);
                if ((int)(threadIdx.x & 63) == leader) {
                    int *b = box + 4 * ((size_t)lc * FACES + F);
                    atomicMin(&b[0], b0);
                    atomicMin(&b[1], b1);
                    atomicMax(&b[2], b2);
                    atomicMax(&b[3], b3);
                }
                if (mine) has = false;
            }
        }
        return;
    }
    if (!valid || cls[l] != LINK_GRID) return;
    const CamGrid &cg = grids[cam];
    const CapDisc disc = capDisc(h, tp);
    u32 o = offset[t];
    int r[4];
    for (u32 F = 0; F < FACES; ++F) {
        if (!faceRange(h, tp, F, inv, G, r)) continue;
        const FaceGrid &fg = cg.face[F];
        const bool useDisc = F == disc.face && disc.valid;   // the box has <= WIDE_CELLS cells here
        for (int v = r[2]; v <= r[3]; ++v)
            for (int u = r[0]; u <= r[1]; ++u) {
                if (useDisc && !cellMeetsDisc(disc, u, v, inv)) continue;
                keys[o] = fg.base + (u32)(v - fg.lo[1]) * fg.dim[0] + (u32)(u - fg.lo[0]);
                vals[o] = l | (useDisc ? ENTRY_OWN_FACE : 0u);   // links < 2^31
                ++o;
            }
    }
}

__global__ void k_gatherU32(const u32 *__restrict__ src, const u32 *__restrict__ order, u32 n,
                            u32 *__restrict__ dst) {
    const u32 t = blockIdx.x * BLOCK + threadIdx.x;
    if (t < n) dst[t] = src[order[t]];
}

// The leaf pass's test record of each entry (a cell's entries are read contiguously): on the
// face of the link's own cone axis, the axis's face coordinates (u0, v0) and the disc radius
// rho (capDisc); the leaf's face coordinates must lie within rho.  Elsewhere (a cone crossing
// to another face, or a wide link) rho is 1e30: no test.  The low 5 bits of rho hold the
// link's depth limit, after rho is rounded up to a multiple of 32 ulps.  Then the link.
__global__ void k_entryRecords(const u32 *__restrict__ val, size_t n, const float4 *__restrict__ axis,
                               const float4 *__restrict__ aux, float4 *__restrict__ rec, float *__restrict__ reach) {
    const size_t e = (size_t)blockIdx.x * BLOCK + threadIdx.x;
    if (e >= n) return;
    const u32 v = val[e];
    const u32 l = v & ~ENTRY_OWN_FACE;
    const float4 au = aux[l];
    float u0 = 0.f, v0 = 0.f, rho = NO_DISC;
    if (v & ENTRY_OWN_FACE) {
        const float4 ax = axis[l];
        const float h[3] = {ax.x, ax.y, ax.z};
        const CapDisc d = capDisc(h, au.y);
        u0 = d.u0;
        v0 = d.v0;
        rho = d.rho;
    }
    const u32 rb = ((__float_as_uint(rho) + 32u) & ~31u) | min((u32)au.z, 31u);
    rec[e] = make_float4(u0, v0, __uint_as_float(rb), __uint_as_float(l));
    if (reach) reach[e] = au.x;
}

__global__ void k_fillBox(int *box, size_t n) {
    const size_t i = (size_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) box[i] = (i & 2) ? INT_MIN : INT_MAX;
}

// sum of n values into *out: a fixed grid of SUM_BLOCKS blocks strides over the array, one
// atomic per wavefront
constexpr u32 SUM_BLOCKS = 1024;
__global__ void k_sumU32(const u32 *__restrict__ v, u32 n, unsigned long long *__restrict__ out) {
    unsigned long long x = 0;
    for (u32 i = blockIdx.x * BLOCK + threadIdx.x; i < n; i += SUM_BLOCKS * BLOCK) x += v[i];
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
    if ((threadIdx.x & 63) == 0 && x) atomicAdd(out, x);
}

// per cell: its run [start, end) in the sorted entries and the largest reach of its links
// (the first: entries of a cell are by descending reach)
__global__ void k_cellRuns(const u32 *__restrict__ keys, size_t n, const float *__restrict__ entryReach,
                           u32 *__restrict__ start, u32 *__restrict__ end, float *__restrict__ cellReach) {
    const size_t j = (size_t)blockIdx.x * BLOCK + threadIdx.x;
    if (j >= n) return;
    const u32 c = keys[j];
    if (j == 0 || keys[j - 1] != c) {
        start[c] = (u32)j;
        cellReach[c] = entryReach[j];
    }
    if (j + 1 == n || keys[j + 1] != c) end[c] = (u32)(j + 1);
}

__global__ void k_wideFlags(const uint8_t *__restrict__ cls, u32 links, u32 *__restrict__ flag) {
    const u32 l = blockIdx.x * BLOCK + threadIdx.x;
    if (l < links) flag[l] = cls[l] == LINK_WIDE ? 1u : 0u;
}

__global__ void k_iota(u32 *p, u32 n) {
    const u32 i = blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) p[i] = i;
}

__global__ void k_wideCams(const u32 *__restrict__ wide, u32 n, const u32 *__restrict__ cams,
                           u32 *__restrict__ key) {
    const u32 i = blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) key[i] = cams[wide[i]];
}

__global__ void k_wideRuns(const u32 *__restrict__ camSorted, u32 n, u32 *__restrict__ begin,
                           u32 *__restrict__ endOut) {
    const u32 j = blockIdx.x * BLOCK + threadIdx.x;
    if (j >= n) return;
    const u32 c = camSorted[j];
    if (j == 0 || camSorted[j - 1] != c) begin[c] = j;
    if (j + 1 == n || camSorted[j + 1] != c) endOut[c] = j + 1;
}

__global__ void k_flaggedLeafFlags(const u32 *__restrict__ hitsS, u32 L, u32 *__restrict__ flag) {
    const u32 l = blockIdx.x * BLOCK + threadIdx.x;
    if (l < L) flag[l] = hitsS[l] ? 1u : 0u;
}

__global__ void k_listFlagged(const u32 *__restrict__ flag, const u32 *__restrict__ pos, u32 n,
                              u32 *__restrict__ list) {
    const u32 l = blockIdx.x * BLOCK + threadIdx.x;
    if (l < n && flag[l]) list[pos[l]] = l;
}

struct GatherArgs {
    Frame f;
    const u64 *leafLoc;
    const CamGrid *grids;
    u32 camCount;
    const u32 *cellStart, *cellEnd;
    const float *cellReach;    // largest reach of the cell's links
    const float *entryReach;   // per entry, descending within a cell
    const float4 *rec;       // per entry (k_entryRecords), sorted by cell
    const float4 *wideRec;   // per wide link, sorted by camera (link order within a camera)
    LinkGeo g;
    double bandwidth;
};

TSR_HD void twoSum(float &hi, float &lo, float x) {   // hi + lo += x (Kahan-Babuska)
    const float s = hi + x;
    const float bp = s - hi;
    const float err = (hi - (s - bp)) + (x - bp);
    hi = s;
    lo += err;
}

// set bits of the wave-wide mask m below this lane
__device__ __forceinline__ u32 lanesBelow(u64 m) {
    return __builtin_amdgcn_mbcnt_hi((u32)(m >> 32), __builtin_amdgcn_mbcnt_lo((u32)m, 0u));
}

// The reference's FP64 kernel (coneKernel, Occupancy.cpp:279-346) of link `link` from camera
// `cam` at leaf centre x, added to this lane's Kahan sums of kernel values and weighted cone
// lengths when it is non-zero
__device__ __forceinline__ void evalQueued(const GatherArgs &a, const CamGrid *camL, const double x[3], u32 link,
                                           u32 cam, float &sk, float &ck, float &sl, float &cl, u32 &hits) {
    const CamGrid &cg = cam < CAM_LDS ? camL[cam] : a.grids[cam];
    const float4 *ev = a.g.eval + EVAL_WORDS * (size_t)link;
    const float4 ps = ev[0];
    const float4 nr = ev[1];
    const double2 cd = reinterpret_cast<const double2 *>(ev)[2];
    ConeD c;
    c.a[0] = cg.a[0]; c.a[1] = cg.a[1]; c.a[2] = cg.a[2];
    c.v[0] = (double)ps.x - c.a[0]; c.v[1] = (double)ps.y - c.a[1]; c.v[2] = (double)ps.z - c.a[2];
    c.n[0] = -(double)nr.x; c.n[1] = -(double)nr.y; c.n[2] = -(double)nr.z;
    c.R = a.bandwidth * (double)ps.w;
    c.invDenom = cd.x;
    c.tStart = 0.0;
    c.invCut = 1.0;
    c.length = cd.y;
    c.degenerate = false;
    const double conf = nr.w;
    const double kv = coneKernel(c, x, false, conf);
    if (kv == 0.0) return;
    const float va = (float)kv - ck;
    const float ta = sk + va;
    ck = (ta - sk) - va;
    sk = ta;
    const float vb = (float)(conf * cd.y) - cl;
    const float tb = sl + vb;
    cl = (tb - sl) - vb;
    sl = tb;
    ++hits;
}

// LDS written by some lanes of a wavefront and read by others after this point: the
// wavefront's LDS operations run in order, so a compiler fence at wavefront scope suffices
// (no s_barrier; each wavefront of a block works on its own LDS slice)
__device__ __forceinline__ void waveSync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}


// One wavefront per leaf (LEAVES_PER_WAVE consecutive leaves of the list in turn).  Per batch
// of 64 cameras, lane q looks up camera q's cell and keeps the cell's links whose reach is at
// least the leaf's distance from the camera (a prefix: a cell's links are by descending
// reach); the batch's candidates (each camera's prefix, then its wide links) are then spread
// over the 64 lanes in a fixed order.
__global__ __launch_bounds__(64 * GATHER_WAVES) void k_gLeaves(GatherArgs a, const u32 *__restrict__ leaves,
                                                               u32 nLeaves, u32 blocksPerXcd,
                                                               float *__restrict__ ksum, float *__restrict__ clen,
                                                               u32 *__restrict__ hitCount,
                                                               unsigned long long *__restrict__ stats) {
    __shared__ u32 segStartS[GATHER_WAVES][64], segBegS[GATHER_WAVES][64], segLenS[GATHER_WAVES][64],
        segWideS[GATHER_WAVES][64], segCamS[GATHER_WAVES][64];
    __shared__ float segUS[GATHER_WAVES][64], segVS[GATHER_WAVES][64];   // the leaf's face coordinates per segment
    __shared__ u32 qLinkS[GATHER_WAVES][64 * (CAND_UNROLL + 1)], qCamS[GATHER_WAVES][64 * (CAND_UNROLL + 1)];
    const u32 lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    u32 *segStart = segStartS[wv], *segBeg = segBegS[wv], *segLen = segLenS[wv], *segWide = segWideS[wv],
        *segCam = segCamS[wv];
    float *segU = segUS[wv], *segV = segVS[wv];
    u32 *qLink = qLinkS[wv], *qCam = qCamS[wv];   // candidates for the FP64 kernel
    u32 qn = 0;   // queued candidates (wave-uniform)
    // the first CAM_LDS cameras' grids, staged once per block (a lane reads its own camera's
    // grid: from LDS, not 64 scattered cache lines)
    __shared__ CamGrid camL[CAM_LDS];
    {
        const u32 words = (u32)(min(a.camCount, CAM_LDS) * (sizeof(CamGrid) / 4));
        const u32 *src = reinterpret_cast<const u32 *>(a.grids);
        u32 *dst = reinterpret_cast<u32 *>(camL);
        for (u32 i = threadIdx.x; i < words; i += blockDim.x) dst[i] = src[i];
        __syncthreads();
    }
    // blocks go round-robin over the XCDs: give each XCD a contiguous run of leaf groups, so
    // that one XCD's L2 holds the cells and links of one part of the scene
    const u32 group = ((blockIdx.x % XCD_COUNT) * blocksPerXcd + blockIdx.x / XCD_COUNT) * GATHER_WAVES + wv;
    u32 lookups = 0, cand = 0, evals = 0;
    const u32 K = a.camCount;
    for (u32 t = 0; t < LEAVES_PER_WAVE; ++t) {
        const u32 li = group * LEAVES_PER_WAVE + t;
        if (li >= nLeaves) break;   // wave-uniform
        const u32 leaf = leaves[li];
        const u64 loc = a.leafLoc[leaf];
        const u32 d = locDepth(loc);
        u32 k[3];
        locCoords(loc, k[0], k[1], k[2]);
        double x[3];
        for (int c = 0; c < 3; ++c) x[c] = a.f.B(c, 2u * k[c] + 1u, d + 1);   // the traversal's centre
        const float xf[3] = {(float)x[0], (float)x[1], (float)x[2]};
        const float xAbs = fabsf(xf[0]) + fabsf(xf[1]) + fabsf(xf[2]);
        const float depthF = (float)d;
        float sk = 0.f, ck = 0.f, sl = 0.f, cl = 0.f;
        u32 hits = 0;
        for (u32 cb = 0; cb < K; cb += 64) {
            const u32 cam = cb + lane;
            u32 beg = 0, len = 0, wb = 0, wl = 0;
            float uf = 0.f, vf = 0.f;
            if (cam < K) {
                const CamGrid &cg = cam < CAM_LDS ? camL[cam] : a.grids[cam];
                const double w0 = x[0] - (double)cg.a[0], w1 = x[1] - (double)cg.a[1], w2 = x[2] - (double)cg.a[2];
                const double b0 = fabs(w0), b1 = fabs(w1), b2 = fabs(w2);
                u32 F;
                double u, v, m;
                if (b0 >= b1 && b0 >= b2) { F = w0 >= 0.0 ? 0 : 1; m = b0; u = w1; v = w2; }
                else if (b1 >= b2) { F = w1 >= 0.0 ? 2 : 3; m = b1; u = w0; v = w2; }
                else { F = w2 >= 0.0 ? 4 : 5; m = b2; u = w0; v = w1; }
                const FaceGrid &fg = cg.face[F];
                if (m > 0.0 && fg.dim[0]) {
                    // face coordinates from the FP64 direction rounded to FP32 (relative error
                    // ~1e-7, far inside the 1e-5 padding of the links' cell ranges)
                    const int G = (int)cg.G;
                    const float mf = (float)m;
                    uf = (float)u / mf;
                    vf = (float)v / mf;
                    const int cu = cellOf(uf, cg.inv, G) - fg.lo[0];
                    const int cv = cellOf(vf, cg.inv, G) - fg.lo[1];
                    if (cu >= 0 && cv >= 0 && (u32)cu < fg.dim[0] && (u32)cv < fg.dim[1]) {
                        const u32 cell = fg.base + (u32)cv * fg.dim[0] + (u32)cu;
                        // a contributing link has |x - a| < |e - a| + R, and its reach is that
                        // inflated by 1e-6: dist below |x - a| by more than FP32 rounding
                        const float dist = sqrtf((float)(w0 * w0 + w1 * w1 + w2 * w2)) * (1.f - 4e-7f);
                        ++lookups;
                        if (a.cellReach[cell] >= dist) {
                            u32 lo = a.cellStart[cell], hi = a.cellEnd[cell];
                            beg = lo;
                            while (lo < hi) {   // first entry with reach < dist
                                const u32 mid = (lo + hi) >> 1;
                                if (a.entryReach[mid] >= dist) lo = mid + 1;
                                else hi = mid;
                            }
                            len = lo - beg;
                        }
                    }
                }
                wb = cg.wideBegin;
                wl = cg.wideEnd - cg.wideBegin;
            }
            const u32 n = len + wl;
            u32 incl = n;
            for (u32 off = 1; off < 64; off <<= 1) {
                const u32 y = __shfl_up(incl, off);
                if (lane >= off) incl += y;
            }
            const u32 total = __shfl(incl, 63);
            // the non-empty segments, compacted in camera order
            const u64 nonEmpty = __ballot(n > 0);
            const u32 nseg = (u32)__popcll(nonEmpty);
            if (n > 0) {
                const u32 at = lanesBelow(nonEmpty);
                segStart[at] = incl - n;
                segBeg[at] = beg;
                segLen[at] = len;
                segWide[at] = wb;
                segCam[at] = cam;
                segU[at] = uf;
                segV[at] = vf;
            }
            waveSync();
            // candidates passing the FP32 test are queued (link, camera) in LDS; the FP64 kernel
            // runs on 64 queued candidates at a time, one per lane.  A lane takes CAND_UNROLL
            // candidates per step (j, j + 64, ...) with their record loads issued together.
            for (u32 base = 0; base < total; base += 64 * CAND_UNROLL) {
                float4 rc[CAND_UNROLL];
                u32 qs[CAND_UNROLL];
#pragma unroll
                for (u32 uu = 0; uu < CAND_UNROLL; ++uu) {
                    const u32 j = base + 64 * uu + lane;
                    qs[uu] = 64;
                    if (j < total) {
                        u32 lo = 0, hi = nseg - 1;   // last segment starting at or before j
                        while (lo < hi) {
                            const u32 mid = (lo + hi + 1) >> 1;
                            if (segStart[mid] <= j) lo = mid;
                            else hi = mid - 1;
                        }
                        const u32 q = lo;
                        const u32 r = j - segStart[q];
                        const u32 sl0 = segLen[q];
                        rc[uu] = r < sl0 ? a.rec[segBeg[q] + r] : a.wideRec[segWide[q] + (r - sl0)];
                        qs[uu] = q;
                    }
                }
#pragma unroll
                for (u32 uu = 0; uu < CAND_UNROLL; ++uu) {
                    bool pass = false;
                    if (qs[uu] < 64) {
                        ++cand;
                        // not deeper than the link's depth limit, and the leaf's face coordinates
                        // within the cone's disc (every entry of the prefix reaches the leaf's
                        // distance from the camera)
                        const u32 rb = __float_as_uint(rc[uu].z);
                        const float du = segU[qs[uu]] - rc[uu].x, dv = segV[qs[uu]] - rc[uu].y;
                        pass = depthF <= (float)(rb & 31u) && du * du + dv * dv <= rc[uu].z * rc[uu].z;
                    }
                    const u64 m = __ballot(pass);
                    if (pass) {
                        ++evals;
                        const u32 at = qn + lanesBelow(m);
                        qLink[at] = __float_as_uint(rc[uu].w);
                        qCam[at] = segCam[qs[uu]];
                    }
                    qn += (u32)__popcll(m);
                }
                while (qn >= 64) {
                    waveSync();
                    evalQueued(a, camL, x, qLink[lane], qCam[lane], sk, ck, sl, cl, hits);
                    // the rest moves to the front (read all, then write)
                    const u32 rest = qn - 64;
                    u32 l2[CAND_UNROLL], c2[CAND_UNROLL];
#pragma unroll
                    for (u32 uu = 0; uu < CAND_UNROLL; ++uu) {
                        const u32 i = lane + 64 * uu;
                        l2[uu] = i < rest ? qLink[64 + i] : 0u;
                        c2[uu] = i < rest ? qCam[64 + i] : 0u;
                    }
                    waveSync();
#pragma unroll
                    for (u32 uu = 0; uu < CAND_UNROLL; ++uu) {
                        const u32 i = lane + 64 * uu;
                        if (i < rest) {
                            qLink[i] = l2[uu];
                            qCam[i] = c2[uu];
                        }
                    }
                    qn = rest;
                }
            }
            waveSync();
        }
        if (qn) {   // the queue's rest
            waveSync();
            if (lane < qn) evalQueued(a, camL, x, qLink[lane], qCam[lane], sk, ck, sl, cl, hits);
            qn = 0;
            waveSync();
        }
        // the lanes' Kahan sums (sk - ck, sl - cl) folded by a fixed two-sum tree into lane 0
        float hk = sk, lk = -ck, hl = sl, ll = -cl;
        for (int off = 32; off > 0; off >>= 1) {
            const float ohk = __shfl_down(hk, off), olk = __shfl_down(lk, off);
            const float ohl = __shfl_down(hl, off), oll = __shfl_down(ll, off);
            const u32 oh = __shfl_down(hits, off);
            twoSum(hk, lk, ohk);
            lk += olk;
            twoSum(hl, ll, ohl);
            ll += oll;
            hits += oh;
        }
        if (lane == 0) {
            ksum[leaf] = hk + lk;
            clen[leaf] = hl + ll;
            hitCount[leaf] = hits;
        }
    }
    if (stats) {
        unsigned long long v[3] = {lookups, cand, evals};
        for (int i = 0; i < 3; ++i) {
            unsigned long long y = v[i];
            for (int off = 32; off > 0; off >>= 1) y += __shfl_xor(y, off);
            if (lane == 0 && y) atomicAdd(&stats[i], y);
        }
    }
}

}  // namespace

std::shared_ptr<EmptinessIndex> buildEmptinessIndex(hipStream_t st, Scratch &sc, const EmptinessLinks &src) {
    const u32 K = src.camCount;
    const u32 links = src.samples * src.maxCams;
    auto ix = std::make_shared<EmptinessIndex>();
    ix->links = links;
    ix->camCount = K;
    ix->bandwidth = src.bandwidth;
    ix->sceneId = src.sceneId;
    const bool trace = std::getenv("TSR_TRACE") != nullptr;
    auto tmark = std::chrono::steady_clock::now();
    auto lap = [&](const char *what) {   // TSR_TRACE: synchronised sub-step times
        if (!trace) return;
        HIPCHECK(hipStreamSynchronize(st));
        const auto t = std::chrono::steady_clock::now();
        std::fprintf(stderr, "[tsr emptiness index] %-22s %8.2f ms\n", what,
                     std::chrono::duration<double, std::milli>(t - tmark).count());
        tmark = t;
    };
    ix->grids.alloc(std::max(K, 1u));
    if (links == 0) {
        ix->cellStart.alloc(1);
        ix->cellEnd.alloc(1);
        ix->cellReach.alloc(1);
        return ix;
    }

    // ---- per-link cones; links grouped by camera and sorted by descending reach
    DBuf<float4> axis(links), aux(links);
    ix->eval.alloc(EVAL_WORDS * (size_t)links);
    DBuf<uint8_t> cls(links);
    DBuf<unsigned long long> camTheta(2 * (size_t)std::max(K, 1u));
    HIPCHECK(hipMemsetAsync(camTheta.p, 0, camTheta.bytes(), st));
    const LinkGeo geo{axis.p, aux.p, ix->eval.p};
    DBuf<u32> byCam(links), byReach(links);
    {
        DBuf<u32> camKey(links), reachKey(links), iota(links), keyS(links);
        k_gLinkSetup<<<gridFor(links, BLOCK), BLOCK, 0, st>>>(src, links, geo, cls.p, camKey.p, reachKey.p, iota.p);
        int bits = 1;
        while (bits < 32 && (1ull << bits) <= K) ++bits;
        sortPairs(sc, camKey.p, keyS.p, iota.p, byCam.p, links, 0, bits, st);
        sortPairs(sc, reachKey.p, keyS.p, iota.p, byReach.p, links, 0, 32, st);
        k_camTheta<<<gridFor(links, BLOCK), BLOCK, 0, st>>>(byCam.p, links, cls.p, src.cams, aux.p, camTheta.p);
    }
    std::vector<unsigned long long> hTheta(2 * (size_t)K);
    std::vector<float> hcx(K), hcy(K), hcz(K);
    if (K) {
        HIPCHECK(hipMemcpyAsync(hTheta.data(), camTheta.p, 16 * (size_t)K, hipMemcpyDeviceToHost, st));
        HIPCHECK(hipMemcpyAsync(hcx.data(), src.cx, 4 * (size_t)K, hipMemcpyDeviceToHost, st));
        HIPCHECK(hipMemcpyAsync(hcy.data(), src.cy, 4 * (size_t)K, hipMemcpyDeviceToHost, st));
        HIPCHECK(hipMemcpyAsync(hcz.data(), src.cz, 4 * (size_t)K, hipMemcpyDeviceToHost, st));
    }
    HIPCHECK(hipStreamSynchronize(st));
    // cells along a face side per camera: a cell about the camera's mean cone angle
    std::vector<u32> camG(std::max(K, 1u), 1);
    for (u32 c = 0; c < K; ++c) {
        if (!hTheta[2 * c + 1]) continue;
        const double mean = (double)hTheta[2 * c] / 4294967296.0 / (double)hTheta[2 * c + 1];
        camG[c] = (u32)std::min<double>(MAX_FACE_CELLS, std::max(1.0, std::ceil(2.0 / mean)));
    }
    lap("link cones");

    // ---- cell grids: bounding boxes per (camera, face), coarsened until they fit the budget
    const u64 cellBudget = std::min<u64>(1ull << 31, std::max<u64>(1ull << 22, 8ull * links));
    const u64 entryBudget = 1ull << 31;
    DBuf<u32> camGd(std::max(K, 1u)), count(links);
    DBuf<int> box(4 * (size_t)FACES * std::max(K, 1u));
    DBuf<unsigned long long> sum1(1);
    std::vector<int> hbox(4 * (size_t)FACES * K);
    std::vector<CamGrid> grids(std::max(K, 1u));
    u64 totalCells = 0, totalEntries = 0;
    for (int attempt = 0;; ++attempt) {
        if (attempt > 40) return nullptr;
        HIPCHECK(hipMemcpyAsync(camGd.p, camG.data(), 4 * (size_t)std::max(K, 1u), hipMemcpyHostToDevice, st));
        k_fillBox<<<gridFor(box.n, BLOCK), BLOCK, 0, st>>>(box.p, box.n);
        k_gLinkCells<<<gridFor(links, BLOCK), BLOCK, 0, st>>>(links, 0, byCam.p, src.cams, axis.p, aux.p, nullptr,
                                                              camGd.p, cls.p, count.p, box.p, nullptr, nullptr,
                                                              nullptr);
        HIPCHECK(hipMemsetAsync(sum1.p, 0, 8, st));
        k_sumU32<<<SUM_BLOCKS, BLOCK, 0, st>>>(count.p, links, sum1.p);
        if (K) HIPCHECK(hipMemcpyAsync(hbox.data(), box.p, 4 * hbox.size(), hipMemcpyDeviceToHost, st));
        totalEntries = readScalar(sum1.p, st);
        totalCells = 0;
        for (u32 c = 0; c < K; ++c) {
            CamGrid &g = grids[c];
            g.a[0] = hcx[c]; g.a[1] = hcy[c]; g.a[2] = hcz[c];
            g.G = camG[c];
            g.inv = 0.5f * (float)camG[c];
            g.wideBegin = g.wideEnd = 0;
            for (u32 F = 0; F < FACES; ++F) {
                const int *b = &hbox[4 * ((size_t)c * FACES + F)];
                FaceGrid &fg = g.face[F];
                fg.base = (u32)std::min<u64>(totalCells, 0xffffffffull);
                if (b[0] > b[2]) {
                    fg.lo[0] = fg.lo[1] = 0;
                    fg.dim[0] = fg.dim[1] = 0;
                    continue;
                }
                fg.lo[0] = b[0];
                fg.lo[1] = b[1];
                fg.dim[0] = (u32)(b[2] - b[0] + 1);
                fg.dim[1] = (u32)(b[3] - b[1] + 1);
                totalCells += (u64)fg.dim[0] * fg.dim[1];
            }
        }
        if (totalCells <= cellBudget && totalEntries < entryBudget) break;
        bool coarser = false;
        for (u32 c = 0; c < K; ++c)
            if (camG[c] > 1) {
                camG[c] = (camG[c] + 1) / 2;
                coarser = true;
            }
        if (!coarser) return nullptr;
    }
    lap("cell grids");

    // ---- wide links per camera (link order within a camera) with their test records
    {
        DBuf<u32> flag(links), iota(links), cnt(1), wl(links);
        k_wideFlags<<<gridFor(links, BLOCK), BLOCK, 0, st>>>(cls.p, links, flag.p);
        k_iota<<<gridFor(links, BLOCK), BLOCK, 0, st>>>(iota.p, links);
        selectFlagged(sc, iota.p, flag.p, wl.p, cnt.p, links, st);
        ix->nWide = readScalar(cnt.p, st);
        const u32 nWide = ix->nWide;
        if (nWide) {
            DBuf<u32> key(nWide), keyS(nWide), wideSorted(nWide), wb(K), we(K);
            k_wideCams<<<gridFor(nWide, BLOCK), BLOCK, 0, st>>>(wl.p, nWide, src.cams, key.p);
            int bits = 1;
            while (bits < 32 && (1ull << bits) < K) ++bits;
            sortPairs(sc, key.p, keyS.p, wl.p, wideSorted.p, nWide, 0, bits, st);
            ix->wideRec.alloc(nWide);
            k_entryRecords<<<gridFor(nWide, BLOCK), BLOCK, 0, st>>>(wideSorted.p, nWide, axis.p, aux.p,
                                                                    ix->wideRec.p, nullptr);
            HIPCHECK(hipMemsetAsync(wb.p, 0, 4 * (size_t)K, st));
            HIPCHECK(hipMemsetAsync(we.p, 0, 4 * (size_t)K, st));
            k_wideRuns<<<gridFor(nWide, BLOCK), BLOCK, 0, st>>>(keyS.p, nWide, wb.p, we.p);
            std::vector<u32> hb(K), he(K);
            HIPCHECK(hipMemcpyAsync(hb.data(), wb.p, 4 * (size_t)K, hipMemcpyDeviceToHost, st));
            HIPCHECK(hipMemcpyAsync(he.data(), we.p, 4 * (size_t)K, hipMemcpyDeviceToHost, st));
            HIPCHECK(hipStreamSynchronize(st));
            for (u32 c = 0; c < K; ++c) {
                grids[c].wideBegin = hb[c];
                grids[c].wideEnd = he[c];
            }
        }
    }
    HIPCHECK(hipMemcpyAsync(ix->grids.p, grids.data(), sizeof(CamGrid) * std::max(K, 1u), hipMemcpyHostToDevice,
                            st));

    // ---- (cell, link) entries sorted by cell, descending reach within a cell, with the
    // links' test records alongside
    const size_t nE = (size_t)totalEntries;
    ix->totalCells = totalCells;
    ix->totalEntries = totalEntries;
    ix->rec.alloc(std::max<size_t>(nE, 1));
    ix->entryReach.alloc(std::max<size_t>(nE, 1));
    DBuf<u32> keysS(std::max<size_t>(nE, 1));
    {
        DBuf<u32> keys(std::max<size_t>(nE, 1)), vals(std::max<size_t>(nE, 1)), entries(std::max<size_t>(nE, 1));
        {
            DBuf<u32> countR(links), offR(links);
            k_gatherU32<<<gridFor(links, BLOCK), BLOCK, 0, st>>>(count.p, byReach.p, links, countR.p);
            exclusiveScan(sc, countR.p, offR.p, links, st);
            k_gLinkCells<<<gridFor(links, BLOCK), BLOCK, 0, st>>>(links, 1, byReach.p, src.cams, axis.p, aux.p,
                                                                  ix->grids.p, camGd.p, cls.p, nullptr, nullptr,
                                                                  offR.p, keys.p, vals.p);
        }
        int bits = 1;
        while (bits < 32 && (1ull << bits) < totalCells) ++bits;
        sortPairs(sc, keys.p, keysS.p, vals.p, entries.p, nE, 0, bits, st);
        if (nE) k_entryRecords<<<gridFor(nE, BLOCK), BLOCK, 0, st>>>(entries.p, nE, axis.p, aux.p, ix->rec.p,
                                                                     ix->entryReach.p);
    }
    const size_t nC = std::max<u64>(totalCells, 1);
    ix->cellStart.alloc(nC);
    ix->cellEnd.alloc(nC);
    ix->cellReach.alloc(nC);
    HIPCHECK(hipMemsetAsync(ix->cellStart.p, 0, 4 * nC, st));
    HIPCHECK(hipMemsetAsync(ix->cellEnd.p, 0, 4 * nC, st));
    HIPCHECK(hipMemsetAsync(ix->cellReach.p, 0, 4 * nC, st));
    if (nE) k_cellRuns<<<gridFor(nE, BLOCK), BLOCK, 0, st>>>(keysS.p, nE, ix->entryReach.p, ix->cellStart.p,
                                                             ix->cellEnd.p, ix->cellReach.p);
    lap("cell index");
    if (trace)
        std::fprintf(stderr, "[tsr emptiness index] %u links, %llu cells, %llu entries, %u wide links\n", links,
                     (unsigned long long)totalCells, (unsigned long long)totalEntries, ix->nWide);
    return ix;
}

void emptinessByLeaves(tsr_context *ctx, const tsr_tree *tree, const tsr_scene *scene, const tsr_params &prm,
                       const u32 *hitsS, float *ksum, float *clen, u32 *hitCount, double *confs, u64 &hitsOut) {
    hipStream_t st = ctx->stream;
    Scratch &sc = ctx->scratch;
    const u32 L = tree->leafCount, K = scene->camCount;
    const u32 links = tree->sampleCount * tree->maxCams;
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
    lap("(emptiness start)");
    HIPCHECK(hipMemsetAsync(ksum, 0, (size_t)L * 4, st));
    HIPCHECK(hipMemsetAsync(clen, 0, (size_t)L * 4, st));
    HIPCHECK(hipMemsetAsync(hitCount, 0, (size_t)L * 4, st));
    hitsOut = 0;
    // the tree's links (node order): the cone count's terms, and the index when the tree has none
    const EmptinessLinks treeLinks{tree->px.p, tree->py.p, tree->pz.p, tree->nx.p, tree->ny.p, tree->nz.p,
                                   tree->scale.p, tree->conf.p, tree->cams.p, scene->cx.p, scene->cy.p,
                                   scene->cz.p, tree->sampleDepth.p, tree->sampleCount, tree->maxCams, K,
                                   prm.bandwidth_factor, scene->id};
    if (links) k_linkConfs<<<gridFor(links, BLOCK), BLOCK, 0, st>>>(treeLinks, links, confs);
    if (links == 0 || L == 0) return;
    std::shared_ptr<EmptinessIndex> ix = tree->eIndex;
    if (!ix || ix->sceneId != scene->id || ix->bandwidth != prm.bandwidth_factor || ix->links != links ||
        ix->camCount != K) {
        ix = buildEmptinessIndex(st, sc, treeLinks);
        if (!ix) throw tsr_error(TSR_ERR_OUT_OF_MEMORY, "emptiness cell grid does not fit its budget");
    }
    lap("index");

    // ---- the sampleness-flagged leaves, in leaf order
    DBuf<u32> lflag(L), lpos(L), leafList(L);
    k_flaggedLeafFlags<<<gridFor(L, BLOCK), BLOCK, 0, st>>>(hitsS, L, lflag.p);
    exclusiveScan(sc, lflag.p, lpos.p, L, st);
    k_listFlagged<<<gridFor(L, BLOCK), BLOCK, 0, st>>>(lflag.p, lpos.p, L, leafList.p);
    const u32 nLeaves = readScalar(lpos.p + L - 1, st) + readScalar(lflag.p + L - 1, st);
    lflag.release();
    lpos.release();

    // ---- the leaf pass
    const LinkGeo geo{nullptr, nullptr, ix->eval.p};
    GatherArgs ga{tree->frame, tree->leafLoc.p, ix->grids.p, K, ix->cellStart.p, ix->cellEnd.p, ix->cellReach.p,
                  ix->entryReach.p, ix->rec.p, ix->nWide ? ix->wideRec.p : ix->rec.p, geo,
                  (double)prm.bandwidth_factor};
    DBuf<unsigned long long> kst(4);
    HIPCHECK(hipMemsetAsync(kst.p, 0, 32, st));
    const u32 groups = (nLeaves + LEAVES_PER_WAVE - 1) / LEAVES_PER_WAVE;
    const u32 blocks = (groups + GATHER_WAVES - 1) / GATHER_WAVES;
    const u32 perXcd = (blocks + XCD_COUNT - 1) / XCD_COUNT;
    if (nLeaves) {
        timedLaunch(ctx, KSTAT_E_GATHER, [&] {
            k_gLeaves<<<perXcd * XCD_COUNT, 64 * GATHER_WAVES, 0, st>>>(ga, leafList.p, nLeaves, perXcd, ksum, clen,
                                                                         hitCount, kst.p);
        });
    }
    unsigned long long ks[3] = {0, 0, 0};
    HIPCHECK(hipMemcpyAsync(ks, kst.p, 24, hipMemcpyDeviceToHost, st));
    DBuf<unsigned long long> sum1(1);
    HIPCHECK(hipMemsetAsync(sum1.p, 0, 8, st));
    k_sumU32<<<SUM_BLOCKS, BLOCK, 0, st>>>(hitCount, L, sum1.p);
    hitsOut = readScalar(sum1.p, st);
    {
        // algorithmic bytes: per leaf its list entry, code and the three outputs (24 B); per
        // (leaf, camera) cell lookup the cell's run and reach (12 B); per candidate its test
        // record (16 B); per FP64 evaluation the link's 64-byte record
        tsr_kernel_stat &k = ctx->kstat[KSTAT_E_GATHER];
        k.bytes += (u64)nLeaves * 24 + ks[0] * 12 + ks[1] * 16 + ks[2] * 64;
        k.units += nLeaves;
    }
    lap("leaf pass");
    if (trace)
        std::fprintf(stderr, "[tsr occupancy] emptiness by leaves: %u flagged leaves, %llu cell lookups, "
                     "%llu candidates, %llu kernel evaluations, %llu hits\n", nLeaves,
                     (unsigned long long)ks[0], (unsigned long long)ks[1], (unsigned long long)ks[2],
                     (unsigned long long)hitsOut);
}

}  // namespace tsr
