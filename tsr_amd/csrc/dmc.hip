// Dual marching cells on the GPU (DualMarchingCells::extractMesh,
// SR/SurfaceExtraction/DualMarchingCells.cpp:41-169).
//
//   classify   per cell: corner occupancies (MC_CORNER_REORDERING), configuration, count of
//              non-degenerate triangles (DualMarchingCell.cpp:34-43,73-101)
//   emit       scan of counts -> triangles in cell order; corner keys (min leaf, max leaf)
//              (EdgeVertexIndex.cpp:15-19)
//   dedup      stable radix sort of corner keys: the first occurrence of a key defines the
//              vertex; vertices are numbered in first-appearance order (the std::map +
//              addVertex sequence of DualMarchingCells.cpp:153-169)
//   position   FP64 interpolation at the first occurrence (DualMarchingCell.cpp:51-71,
//              DualMarchingCells.cpp:132-151)
//   manifold   undirected edges used by more than two triangles -> TSR_ERR_NON_MANIFOLD_EDGE
//              (FlexibleMesh.cpp:1184-1260)
#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "handles.h"

namespace tsr {

namespace {

constexpr u32 BLOCK = 256;

// Paul Bourke's public-domain marching-cubes triangulation (the table the
// reference uses, DualMarchingCellTables.cpp:57-314), one word per cube
// configuration: low nibble = number of edge entries (3 per triangle), then one
// nibble per cube edge index, first entry in bits 4..7.
__constant__ u64 TRI_TABLE[256] = {
    0x0000000000000000ull, 0x0000000000003803ull, 0x0000000000009103ull, 0x0000000001893816ull,
    0x000000000000a213ull, 0x000000000a213806ull, 0x000000000920a296ull, 0x00000089a8a23829ull,
    0x0000000000002b33ull, 0x0000000000b82b06ull, 0x000000000b320916ull, 0x000000b89b912b19ull,
    0x0000000003ab1a36ull, 0x000000ab8a801a09ull, 0x0000009ab9b30939ull, 0x000000000b8aa896ull,
    0x0000000000008743ull, 0x0000000004370346ull, 0x0000000007489106ull, 0x0000001371749149ull,
    0x000000000748a216ull, 0x000000a214037439ull, 0x000000748209a299ull, 0x0004973727929a2cull,
    0x0000000002b37486ull, 0x00000040242b74b9ull, 0x000000b327481099ull, 0x0001292b9b49b74cull,
    0x000000487ab31a39ull, 0x0004b7401b41ab1cull, 0x00030bab9b09874cull, 0x000000ab99b4b749ull,
    0x0000000000004593ull, 0x0000000003804596ull, 0x0000000000514506ull, 0x0000005135384589ull,
    0x000000000459a216ull, 0x000000594a218039ull, 0x000000204245a259ull, 0x0008434535235a2cull,
    0x000000000b324596ull, 0x000000594b802b09ull, 0x000000b325104509ull, 0x000584b82852512cull,
    0x00000045931ab3a9ull, 0x000ab81a8180594cull, 0x00030bab5b05045cull, 0x000000b8aa858459ull,
    0x0000000009758796ull, 0x0000003753590399ull, 0x0000007517108709ull, 0x0000000007533516ull,
    0x00000021a7598799ull, 0x00037503505921acull, 0x00025a758528208cull, 0x0000007533525a29ull,
    0x0000002b39875979ull, 0x000b72029279759cull, 0x000751871810b32cull, 0x00000051771b12b9ull,
    0x000b3a31a758859cull, 0x0aba010b7905075full, 0x07570805a30b0abfull, 0x0000000005b75ab6ull,
    0x00000000000056a3ull, 0x0000000006a53806ull, 0x0000000006a51096ull, 0x0000006a58913819ull,
    0x0000000001625616ull, 0x0000008036215619ull, 0x0000006206095699ull, 0x000823625285895cull,
    0x00000000056ab326ull, 0x00000056a02b80b9ull, 0x0000006a5b329109ull, 0x000b892b92916a5cull,
    0x000000315356b369ull, 0x0006b51505b0b80cull, 0x0009505606306b3cull, 0x00000089bb969569ull,
    0x0000000008746a56ull, 0x000000a563740349ull, 0x0000007486a50919ull, 0x00049737179156acull,
    0x0000008741562169ull, 0x000743403625521cull, 0x000620560509748cull, 0x962695923497937full,
    0x00000056a4872b39ull, 0x000b720242746a5cull, 0x0006a5b32874910cull, 0x6a54b7b492b9129full,
    0x0006b51535b3748cull, 0xb404b7b016b5b15full, 0x74836b630560950full, 0x0009b7974b96956cull,
    0x000000000a4694a6ull, 0x000000380a946a49ull, 0x00000004606a10a9ull, 0x000a16468618138cull,
    0x0000004624219419ull, 0x000462942921803cull, 0x0000000006244206ull, 0x0000006244282389ull,
    0x00000032b46a94a9ull, 0x0006a4a94b82280cull, 0x000a164606102b3cull, 0x1b8b12184a16146full,
    0x00036b319639469cull, 0x14641916b0181b8full, 0x0000004600636b39ull, 0x00000000086b8466ull,
    0x000000a98a876a79ull, 0x000a76a907a0370cull, 0x0000818717a176acull, 0x00000037117a76a9ull,
    0x000768981861621cull, 0x937390976192962full, 0x0000002066070879ull, 0x0000000002762376ull,
    0x00076898a86ab32cull, 0x7a9a76790b72702full, 0xb32a767a1871081full, 0x00017616a71b12bcull,
    0x63136b619768698full, 0x00000000076b1906ull, 0x00006b0b3607087cull, 0x0000000000006b73ull,
    0x000000000000b673ull, 0x00000000067b8036ull, 0x00000000067b9106ull, 0x00000067b1389189ull,
    0x0000000007b621a6ull, 0x0000007b6803a219ull, 0x0000007b69a20929ull, 0x00089a38a3a27b6cull,
    0x0000000007263276ull, 0x0000000260678079ull, 0x0000009107326729ull, 0x000678891681261cull,
    0x00000073171a67a9ull, 0x000801781a7167acull, 0x0007a69a0a70730cull, 0x0000009a88a7a679ull,
    0x00000000068b4866ull, 0x000000640603b639ull, 0x000000109648b689ull, 0x00063b139369649cull,
    0x0000001a28b64869ull, 0x000640b60b03a21cull, 0x0009a2920b648b4cull, 0x36463b34923a39afull,
    0x0000002642483289ull, 0x0000000002642406ull, 0x000834642432091cull, 0x0000006422414919ull,
    0x0001a6648168318cull, 0x00000040660a01a9ull, 0x39a9303a6834364full, 0x0000000004a649a6ull,
    0x000000000b675946ull, 0x00000067b5943809ull, 0x000000b670451059ull, 0x00051345343867bcull,
    0x000000b6721a4599ull, 0x000594380a217b6cull, 0x000204a24a45b67cull, 0x67b25a523453843full,
    0x0000009452673279ull, 0x000786260680459cull, 0x000045051673263cull, 0x851584812786826full,
    0x00073167161a459cull, 0x459078701671a61full, 0xa737a6a305a4a04full, 0x000a84a458a7a67cull,
    0x00000098b9b65969ull, 0x000590650360b63cull, 0x000b65510b508b0cull, 0x0000001355363b69ull,
    0x00065b8b9b59a21cull, 0xa21965690b603b0full, 0x52025a50865b58bfull, 0x00035a3a25363b6cull,
    0x000283265825985cull, 0x0000002600696599ull, 0x826283865081851full, 0x0000000006126516ull,
    0x698965683a61631full, 0x00006505960a01acull, 0x000000000a658306ull, 0x00000000000065a3ull,
    0x000000000b57a5b6ull, 0x00000003857ba5b9ull, 0x000000091ba57b59ull, 0x0001381897ba57acull,
    0x00000015717b21b9ull, 0x000b27571721380cull, 0x0007b2209729579cull, 0x289823295b27257full,
    0x000000573532a529ull, 0x00052a578258028cull, 0x0002a37353a5109cull, 0x25752a278129289full,
    0x0000000005735316ull, 0x0000005711707809ull, 0x0000007355393099ull, 0x0000000007957896ull,
    0x0000008ba8a54859ull, 0x00003bba50b5405cull, 0x00054aba8a48910cull, 0x41314943b54a4bafull,
    0x0008548b2582152cull, 0xb151b2b543b0b40full, 0x58b8545b2950520full, 0x0000000003b25496ull,
    0x000483543253a52cull, 0x0000000244252a59ull, 0x910854583a532a3full, 0x0002492914252a5cull,
    0x0000001533585489ull, 0x0000000005015406ull, 0x000530509358548cull, 0x0000000000005493ull,
    0x000000ba9b947b49ull, 0x000ba97b9794380cull, 0x000b470414b1ba1cull, 0x4bab474a1843413full,
    0x000219b294b97b4cull, 0x3801b2b197b9479full, 0x00000004224b47b9ull, 0x00042343824b47bcull,
    0x000947732972a92cull, 0x70207872a4797a9full, 0xa040a1a472a3a73full, 0x0000000004782a16ull,
    0x0000003177141949ull, 0x000178180714194cull, 0x0000000003473046ull, 0x0000000000007843ull,
    0x0000000008ba8a96ull, 0x000000a9bb939039ull, 0x000000ba88a0a109ull, 0x000000000a3ba136ull,
    0x0000008b99b1b219ull, 0x0009b2921b93903cull, 0x000000000b08b206ull, 0x000000000000b233ull,
    0x00000098aa828329ull, 0x0000000002902a96ull, 0x0008a1810a82832cull, 0x0000000000002a13ull,
    0x0000000008198316ull, 0x0000000000001903ull, 0x0000000000008303ull, 0x0000000000000000ull,
};
// corner order of the tables -> child slot order of the dual cell (DualMarchingCellTables.cpp:17-20)
__constant__ u32 CORNER_REORDER[8] = {0, 1, 5, 4, 2, 3, 7, 6};
// cube edge -> its two corners (DualMarchingCellTables.cpp:48-53)
__constant__ u32 EDGE_ENDS[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {4, 5}, {5, 6},
                                     {6, 7}, {7, 4}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};

// edges whose corners disagree in configuration `cfg` (MC_EDGE_CUTTINGS_TABLE)
__device__ u32 cutEdges(u32 cfg) {
    u32 m = 0;
    for (u32 e = 0; e < 12; ++e)
        if (((cfg >> EDGE_ENDS[e][0]) & 1) != ((cfg >> EDGE_ENDS[e][1]) & 1)) m |= 1u << e;
    return m;
}

struct CellCorners {
    u32 leaf[8];
    double dist[8];
    u32 cfg;
};

__device__ void loadCell(const u32 *__restrict__ cells, const double *__restrict__ occ, u32 c, CellCorners &cc) {
    cc.cfg = 0;
    for (u32 i = 0; i < 8; ++i) {
        const u32 l = cells[8 * (size_t)c + CORNER_REORDER[i]];
        cc.leaf[i] = l;
        cc.dist[i] = occ[l];
        if (cc.dist[i] > 0.0) cc.cfg |= 1u << i;   // DualMarchingCell.cpp:39
    }
}

// DualMarchingCell.cpp:73-101 isDegeneratedTriangle
__device__ bool degenerate(const CellCorners &cc, const u32 te[3]) {
    for (int k = 0; k < 3; ++k)
        if (cc.leaf[EDGE_ENDS[te[k]][0]] == cc.leaf[EDGE_ENDS[te[k]][1]]) return true;
    for (int k = 0; k < 3; ++k) {
        const u32 a0 = cc.leaf[EDGE_ENDS[te[k]][0]], a1 = cc.leaf[EDGE_ENDS[te[k]][1]];
        const u32 b0 = cc.leaf[EDGE_ENDS[te[(k + 1) % 3]][0]], b1 = cc.leaf[EDGE_ENDS[te[(k + 1) % 3]][1]];
        if (a0 == b0 && a1 == b1) return true;
        if (a0 == b1 && a1 == b0) return true;
    }
    return false;
}

template <typename F>
__device__ void forTriangles(const CellCorners &cc, F f) {
    if (cutEdges(cc.cfg) == 0) return;
    const u64 w = TRI_TABLE[cc.cfg];
    const u32 entries = (u32)(w & 0xf);
    for (u32 q = 0; q < entries; q += 3) {
        const u32 te[3] = {(u32)((w >> (4 * (q + 1))) & 0xf), (u32)((w >> (4 * (q + 2))) & 0xf),
                           (u32)((w >> (4 * (q + 3))) & 0xf)};
        if (degenerate(cc, te)) continue;
        f(te);
    }
}

__global__ void k_countTriangles(const u32 *__restrict__ cells, const double *__restrict__ occ, u32 cellCount,
                                 u32 *__restrict__ counts) {
    const u32 c = blockIdx.x * BLOCK + threadIdx.x;
    if (c >= cellCount) return;
    CellCorners cc;
    loadCell(cells, occ, c, cc);
    u32 n = 0;
    forTriangles(cc, [&](const u32 *) { ++n; });
    counts[c] = n;
}

// corner keys in occurrence order; triangle record = cell index + packed edges
__global__ void k_emitTriangles(const u32 *__restrict__ cells, const double *__restrict__ occ, u32 cellCount,
                                const u32 *__restrict__ offsets, u64 *__restrict__ cornerKey,
                                u32 *__restrict__ triCell, u32 *__restrict__ triEdges) {
    const u32 c = blockIdx.x * BLOCK + threadIdx.x;
    if (c >= cellCount) return;
    CellCorners cc;
    loadCell(cells, occ, c, cc);
    u32 t = offsets[c];
    forTriangles(cc, [&](const u32 *te) {
        for (int k = 0; k < 3; ++k) {
            const u32 l0 = cc.leaf[EDGE_ENDS[te[k]][0]], l1 = cc.leaf[EDGE_ENDS[te[k]][1]];
            const u32 lo = min(l0, l1), hi = max(l0, l1);
            cornerKey[3 * (size_t)t + k] = ((u64)hi << 32) | lo;   // EdgeVertexIndex order: max first
        }
        triCell[t] = c;
        triEdges[t] = te[0] | (te[1] << 4) | (te[2] << 8);
        ++t;
    });
}

__global__ void k_minusOne(u32 *p, u32 n) {
    const u32 i = blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) p[i] -= 1;
}

__global__ void k_iota32(u32 *p, u32 n) {
    const u32 i = blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) p[i] = i;
}

// head flags of runs of equal keys
__global__ void k_runHeads(const u64 *__restrict__ keys, u32 n, u32 *__restrict__ head) {
    const u32 j = blockIdx.x * BLOCK + threadIdx.x;
    if (j < n) head[j] = (j == 0 || keys[j] != keys[j - 1]) ? 1u : 0u;
}

// for each run r: first occurrence (the run head carries the smallest occurrence
// since the sort is stable)
__global__ void k_runFirst(const u32 *__restrict__ head, const u32 *__restrict__ runIdx,
                           const u32 *__restrict__ occSorted, u32 n, u32 *__restrict__ runFirstOcc) {
    const u32 j = blockIdx.x * BLOCK + threadIdx.x;
    if (j < n && head[j]) runFirstOcc[runIdx[j]] = occSorted[j];
}

__global__ void k_scatterRank(const u32 *__restrict__ runSortedByOcc, u32 runs, u32 *__restrict__ vidOfRun) {
    const u32 v = blockIdx.x * BLOCK + threadIdx.x;
    if (v < runs) vidOfRun[runSortedByOcc[v]] = v;
}

__global__ void k_triIndices(const u32 *__restrict__ runIdx, const u32 *__restrict__ occSorted,
                             const u32 *__restrict__ vidOfRun, u32 n, u32 *__restrict__ tris) {
    const u32 j = blockIdx.x * BLOCK + threadIdx.x;
    if (j < n) tris[occSorted[j]] = vidOfRun[runIdx[j]];
}

// DualMarchingCell.cpp:51-71 getInterpolationFactor (FP64)
__device__ double interpolationFactor(double d0, double d1) {
    if (EPS_D >= fabs(d1 - d0)) return 0.5;
    if (EPS_D >= fabs(d0)) return 0.0;
    if (EPS_D >= fabs(d1)) return 1.0;
    return (0.0 - d0) / (d1 - d0);
}

__global__ void k_vertexData(Frame f, const u32 *__restrict__ cells, const double *__restrict__ occ,
                             const u64 *__restrict__ leafLoc, const u32 *__restrict__ firstOccByVertex,
                             const u32 *__restrict__ triCell, const u32 *__restrict__ triEdges, u32 V,
                             double *__restrict__ pos, double *__restrict__ scale, u32 *__restrict__ keys) {
    const u32 v = blockIdx.x * BLOCK + threadIdx.x;
    if (v >= V) return;
    const u32 occIdx = firstOccByVertex[v];
    const u32 t = occIdx / 3, k = occIdx % 3;
    CellCorners cc;
    loadCell(cells, occ, triCell[t], cc);
    const u32 e = (triEdges[t] >> (4 * k)) & 0xf;
    const u32 a = EDGE_ENDS[e][0], b = EDGE_ENDS[e][1];
    const double fct = interpolationFactor(cc.dist[a], cc.dist[b]);
    u32 ka[3], kb[3];
    const u64 la = leafLoc[cc.leaf[a]], lb = leafLoc[cc.leaf[b]];
    locCoords(la, ka[0], ka[1], ka[2]);
    locCoords(lb, kb[0], kb[1], kb[2]);
    const u32 da = locDepth(la), db = locDepth(lb);
    for (int q = 0; q < 3; ++q) {
        const double p0 = f.centre(q, ka[q], da), p1 = f.centre(q, kb[q], db);
        pos[3 * (size_t)v + q] = p0 + (p1 - p0) * fct;
    }
    const double s0 = f.h(da), s1 = f.h(db);
    scale[v] = s0 + (s1 - s0) * fct;
    keys[2 * (size_t)v] = min(cc.leaf[a], cc.leaf[b]);
    keys[2 * (size_t)v + 1] = max(cc.leaf[a], cc.leaf[b]);
}

__global__ void k_edgeKeys(const u32 *__restrict__ tris, u32 T, u64 *__restrict__ ek) {
    const u32 t = blockIdx.x * BLOCK + threadIdx.x;
    if (t >= T) return;
    for (int k = 0; k < 3; ++k) {
        const u32 a = tris[3 * (size_t)t + k], b = tris[3 * (size_t)t + (k + 1) % 3];
        ek[3 * (size_t)t + k] = ((u64)min(a, b) << 32) | max(a, b);
    }
}

__global__ void k_edgeConflicts(const u64 *__restrict__ sorted, u32 n, u32 *__restrict__ conflict) {
    const u32 j = blockIdx.x * BLOCK + threadIdx.x;
    if (j + 2 < n && sorted[j] == sorted[j + 2]) atomicOr(conflict, 1u);
}

}  // namespace

namespace {

__global__ void k_zeroU32(u32 *p, u32 n) {
    const u32 i = blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// Device temporaries of one extraction: either blocks from the context's pool (kept in
// `bufs` until the extraction ends) or consecutive 256-byte aligned pieces of a fixed
// region [base, base + cap) -- the no-allocation entry point's scratch array.
struct DmcArena {
    std::vector<DBuf<unsigned char>> *bufs = nullptr;
    unsigned char *base = nullptr;
    size_t cap = 0, used = 0;
    template <typename T>
    T *get(size_t n) {
        const size_t bytes = std::max<size_t>(n, 1) * sizeof(T);
        if (bufs) {
            bufs->emplace_back(bytes);
            return (T *)bufs->back().p;
        }
        const size_t at = (used + 255) & ~size_t(255);
        if (at + bytes > cap)
            throw tsr_error(TSR_ERR_INVALID_ARGUMENT,
                            "the mesh needs more device scratch than the no-allocation maximum of " +
                                std::to_string(cap >> 20) + " MiB (use tsr_dmc_extract)");
        used = at + bytes;
        return (T *)(base + at);
    }
};

template <typename K, typename V>
void sortPairsIn(DmcArena &ar, const K *kin, K *kout, const V *vin, V *vout, u32 n, int endBit, hipStream_t st) {
    if (n == 0) return;
    size_t tmp = 0;
    HIPCHECK(rocprim::radix_sort_pairs(nullptr, tmp, kin, kout, vin, vout, n, 0u, (unsigned)endBit, st));
    void *d = ar.get<unsigned char>(tmp);
    HIPCHECK(rocprim::radix_sort_pairs(d, tmp, kin, kout, vin, vout, n, 0u, (unsigned)endBit, st));
}

template <typename K>
void sortKeysIn(DmcArena &ar, const K *kin, K *kout, u32 n, int endBit, hipStream_t st) {
    if (n == 0) return;
    size_t tmp = 0;
    HIPCHECK(rocprim::radix_sort_keys(nullptr, tmp, kin, kout, n, 0u, (unsigned)endBit, st));
    void *d = ar.get<unsigned char>(tmp);
    HIPCHECK(rocprim::radix_sort_keys(d, tmp, kin, kout, n, 0u, (unsigned)endBit, st));
}

template <bool Inclusive>
void scanIn(DmcArena &ar, const u32 *in, u32 *out, u32 n, hipStream_t st) {
    if (n == 0) return;
    size_t tmp = 0;
    if (Inclusive) HIPCHECK(rocprim::inclusive_scan(nullptr, tmp, in, out, (size_t)n, rocprim::plus<u32>(), st));
    else HIPCHECK(rocprim::exclusive_scan(nullptr, tmp, in, out, 0u, (size_t)n, rocprim::plus<u32>(), st));
    void *d = ar.get<unsigned char>(tmp);
    if (Inclusive) HIPCHECK(rocprim::inclusive_scan(d, tmp, in, out, (size_t)n, rocprim::plus<u32>(), st));
    else HIPCHECK(rocprim::exclusive_scan(d, tmp, in, out, 0u, (size_t)n, rocprim::plus<u32>(), st));
}

// Where the mesh goes: the mesh handle's own buffers (allocated once the counts are
// known), or caller-provided device buffers of stated capacities.
struct DmcOutput {
    tsr_mesh *mesh = nullptr;
    double *pos = nullptr, *scale = nullptr;
    u32 *keys = nullptr, *tris = nullptr;
    u32 maxVertices = 0, maxTriangles = 0;
    u32 vertexCount = 0, triangleCount = 0;
    void reserveTriangles(u32 T) {
        triangleCount = T;
        if (mesh) {
            mesh->triangleCount = T;
            mesh->tris.alloc(3 * (size_t)T);
            tris = mesh->tris.p;
        } else if (T > maxTriangles) {
            throw tsr_error(TSR_ERR_INVALID_ARGUMENT, "the mesh has " + std::to_string(T) +
                                                          " triangles, more than the given capacity " +
                                                          std::to_string(maxTriangles));
        }
    }
    void reserveVertices(u32 V) {
        vertexCount = V;
        if (mesh) {
            mesh->vertexCount = V;
            mesh->pos.alloc(3 * (size_t)V);
            mesh->scale.alloc(V);
            mesh->keys.alloc(2 * (size_t)V);
            pos = mesh->pos.p;
            scale = mesh->scale.p;
            keys = mesh->keys.p;
        } else if (V > maxVertices) {
            throw tsr_error(TSR_ERR_INVALID_ARGUMENT, "the mesh has " + std::to_string(V) +
                                                          " vertices, more than the given capacity " +
                                                          std::to_string(maxVertices));
        }
    }
};

}  // namespace

// DualMarchingCells::extractMesh over the per-leaf occupancy `occ` (device, FP64)
static void dmcCore(tsr_context *ctx, const tsr_tree *tree, const double *occ, DmcArena &ar, DmcOutput &out) {
    hipStream_t st = ctx->stream;
    const u32 C = tree->cellCount;
    out.reserveTriangles(0);
    out.reserveVertices(0);
    if (C == 0) return;
    u32 *counts = ar.get<u32>(C), *offsets = ar.get<u32>(C);
    k_countTriangles<<<gridFor(C, BLOCK), BLOCK, 0, st>>>(tree->cells.p, occ, C, counts);
    scanIn<false>(ar, counts, offsets, C, st);
    const u64 T64 = (u64)readScalar(offsets + C - 1, st) + readScalar(counts + C - 1, st);
    if (3 * T64 >= (u64)INVALID) throw tsr_error(TSR_ERR_OUT_OF_MEMORY, "too many triangles");
    const u32 T = (u32)T64;
    out.reserveTriangles(T);
    if (T == 0) return;
    const u32 n3 = 3 * T;
    u64 *cornerKey = ar.get<u64>(n3), *keySorted = ar.get<u64>(n3);
    u32 *triCell = ar.get<u32>(T), *triEdges = ar.get<u32>(T), *occIdx = ar.get<u32>(n3);
    u32 *occSorted = ar.get<u32>(n3), *head = ar.get<u32>(n3), *runIdx = ar.get<u32>(n3);
    k_emitTriangles<<<gridFor(C, BLOCK), BLOCK, 0, st>>>(tree->cells.p, occ, C, offsets, cornerKey, triCell,
                                                         triEdges);
    k_iota32<<<gridFor(n3, BLOCK), BLOCK, 0, st>>>(occIdx, n3);
    sortPairsIn(ar, cornerKey, keySorted, occIdx, occSorted, n3, 64, st);
    k_runHeads<<<gridFor(n3, BLOCK), BLOCK, 0, st>>>(keySorted, n3, head);
    scanIn<true>(ar, head, runIdx, n3, st);
    const u32 V = readScalar(runIdx + n3 - 1, st);
    k_minusOne<<<gridFor(n3, BLOCK), BLOCK, 0, st>>>(runIdx, n3);
    out.reserveVertices(V);
    u32 *runFirstOcc = ar.get<u32>(V), *runOrder = ar.get<u32>(V), *firstOccSorted = ar.get<u32>(V);
    u32 *runSortedByOcc = ar.get<u32>(V), *vidOfRun = ar.get<u32>(V);
    k_runFirst<<<gridFor(n3, BLOCK), BLOCK, 0, st>>>(head, runIdx, occSorted, n3, runFirstOcc);
    k_iota32<<<gridFor(V, BLOCK), BLOCK, 0, st>>>(runOrder, V);
    sortPairsIn(ar, runFirstOcc, firstOccSorted, runOrder, runSortedByOcc, V, 32, st);
    k_scatterRank<<<gridFor(V, BLOCK), BLOCK, 0, st>>>(runSortedByOcc, V, vidOfRun);
    k_triIndices<<<gridFor(n3, BLOCK), BLOCK, 0, st>>>(runIdx, occSorted, vidOfRun, n3, out.tris);
    k_vertexData<<<gridFor(V, BLOCK), BLOCK, 0, st>>>(tree->frame, tree->cells.p, occ, tree->leafLoc.p,
                                                      firstOccSorted, triCell, triEdges, V, out.pos, out.scale,
                                                      out.keys);
    // FlexibleMesh::addEdge: an undirected edge may border at most two triangles
    u64 *ek = cornerKey, *ekSorted = keySorted;   // the corner keys are no longer needed
    u32 *conflict = ar.get<u32>(1);
    k_zeroU32<<<1, BLOCK, 0, st>>>(conflict, 1);
    k_edgeKeys<<<gridFor(T, BLOCK), BLOCK, 0, st>>>(out.tris, T, ek);
    sortKeysIn(ar, ek, ekSorted, n3, 64, st);
    k_edgeConflicts<<<gridFor(n3, BLOCK), BLOCK, 0, st>>>(ekSorted, n3, conflict);
    if (readScalar(conflict, st))
        throw tsr_error(TSR_ERR_NON_MANIFOLD_EDGE, "BOOM: a DMC edge is shared by more than two triangles");
    HIPCHECK(hipStreamSynchronize(st));
}

// the no-allocation entry point's temporaries: one array per GPU, alive until the process ends
constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(256) << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

void extractDMC(tsr_context *ctx, const tsr_tree *tree, const double *occ, tsr_mesh *mesh) {
    std::vector<DBuf<unsigned char>> bufs;
    bufs.reserve(32);
    DmcArena ar;
    ar.bufs = &bufs;
    DmcOutput out;
    out.mesh = mesh;
    dmcCore(ctx, tree, occ, ar, out);
}

// The device address of the scratch array on the current device, looked up once per device
// (each GPU holds its own copy of the array).
void *noAllocScratch(int device) {
    static std::mutex m;
    static std::vector<void *> ptrs;
    std::lock_guard<std::mutex> g(m);
    if (device < 0) throw tsr_error(TSR_ERR_INVALID_ARGUMENT, "bad device");
    if ((size_t)device >= ptrs.size()) ptrs.resize(device + 1, nullptr);
    if (!ptrs[device]) {
        void *p = nullptr;
        HIPCHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        ptrs[device] = p;
    }
    return ptrs[device];
}

void extractDMCNoAlloc(tsr_context *ctx, const tsr_tree *tree, const double *occ, u32 maxVertices, u32 maxTriangles,
                       double *pos, double *scale, u32 *keys, u32 *tris, u32 *vertexCount, u32 *triangleCount) {
    DmcArena ar;
    ar.base = (unsigned char *)noAllocScratch(ctx->device);
    ar.cap = NOALLOC_SCRATCH_BYTES;
    DmcOutput out;
    out.pos = pos;
    out.scale = scale;
    out.keys = keys;
    out.tris = tris;
    out.maxVertices = maxVertices;
    out.maxTriangles = maxTriangles;
    try {
        dmcCore(ctx, tree, occ, ar, out);
    } catch (...) {
        *vertexCount = out.vertexCount;
        *triangleCount = out.triangleCount;
        throw;
    }
    *vertexCount = out.vertexCount;
    *triangleCount = out.triangleCount;
}

}  // namespace tsr
