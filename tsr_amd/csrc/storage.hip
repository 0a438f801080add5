// Leaf face-neighbour graph on the GPU and the on-disk formats of the tree, the field and
// the samples (paths relative to SurfaceReconstruction/SurfaceReconstruction/ in TSR):
//
//   neighbours  Leaves::createNeighborsEdges (Scene/Tree/Leaves.cpp:82-180): one thread per
//               (leaf, side) finds the deepest node, no deeper than the leaf, holding the
//               same-size neighbour cell (a location-code hash lookup per level), keeps the
//               pair when that node is a leaf that is coarser, or of the same depth with a
//               larger node index; each kept pair becomes two records keyed by (owner leaf,
//               leaf * 6 + side), and a radix sort by that key gives every leaf's list in the
//               reference's insertion order.
//   files       Nodes::saveToFile/loadFromFile (Nodes.cpp:770-807), Leaves (Leaves.cpp:230-292),
//               DualCells (DualCells.cpp:359-385), Occupancy (SurfaceExtraction/Occupancy.cpp:
//               663-766) and Samples (Scene/Samples.cpp:693-765): the reference's payloads in
//               its order, behind a 16-byte header of this library (the reference's BaseProject
//               File header is not available): "TSRF", kind, version 0, bytes per Real.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "handles.h"

namespace tsr {

namespace {

constexpr u32 BLOCK = 256;

// one slot per (leaf, side): the neighbour leaf of an edge this leaf is responsible for,
// or INVALID (getUnresponsibleEdgeNeighbor, Leaves.cpp:145-163; sides as in
// Nodes::getSideCenterPositionWS, Nodes.cpp:60-84: -x, +x, -y, +y, -z, +z)
__global__ void k_neighbourSlots(HashView nodes, const u32 *__restrict__ child, const u64 *__restrict__ leafLoc,
                                 const u32 *__restrict__ leafNode, u32 L, u32 *__restrict__ other) {
    const u64 i = (u64)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= 6ull * L) return;
    const u32 l = (u32)(i / 6), side = (u32)(i % 6);
    const u64 loc = leafLoc[l];
    const u32 d = locDepth(loc);
    u32 q[3];
    locCoords(loc, q[0], q[1], q[2]);
    const u32 axis = side >> 1;
    if (side & 1) {
        if (q[axis] + 1 >= (1u << d)) { other[i] = INVALID; return; }
        ++q[axis];
    } else {
        if (q[axis] == 0) { other[i] = INVALID; return; }
        --q[axis];
    }
    // deepest existing node holding the cell (the root, code 1, always exists)
    u32 dd = d, n = INVALID;
    for (;;) {
        const u32 sh = d - dd;
        n = nodes.find(makeLoc(q[0] >> sh, q[1] >> sh, q[2] >> sh, dd));
        if (n != INVALID || dd == 0) break;
        --dd;
    }
    if (n == INVALID) { other[i] = INVALID; return; }
    const u32 c = child[n];
    if (!(c & LEAF_BIT)) { other[i] = INVALID; return; }          // finer leaves count this edge
    if (dd == d && n < leafNode[l]) { other[i] = INVALID; return; }   // counted by the other leaf
    other[i] = c & ~LEAF_BIT;
}

__global__ void k_slotValid(const u32 *__restrict__ other, u64 n, u32 *__restrict__ valid) {
    const u64 i = (u64)blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) valid[i] = other[i] != INVALID ? 1u : 0u;
}

// two records per edge: (owner, step) -> the other leaf; step = leaf * 6 + side < 2^32
__global__ void k_neighbourRecords(const u32 *__restrict__ other, const u32 *__restrict__ pos, u64 n,
                                   u64 *__restrict__ key, u32 *__restrict__ val) {
    const u64 i = (u64)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n || other[i] == INVALID) return;
    const u32 l = (u32)(i / 6), o = other[i], p = 2 * pos[i];
    key[p] = ((u64)l << 32) | (u32)i;
    val[p] = o;
    key[p + 1] = ((u64)o << 32) | (u32)i;
    val[p + 1] = l;
}

// offsets[l] = first record of owner l (lower bound in the sorted keys)
__global__ void k_ownerOffsets(const u64 *__restrict__ keys, u64 m, u32 L, u32 *__restrict__ offsets) {
    const u32 l = blockIdx.x * BLOCK + threadIdx.x;
    if (l > L) return;
    const u64 want = (u64)l << 32;
    u64 lo = 0, hi = m;
    while (lo < hi) {
        const u64 mid = (lo + hi) / 2;
        if (keys[mid] < want) lo = mid + 1;
        else hi = mid;
    }
    offsets[l] = (u32)lo;
}

__global__ void k_insertNodes(HashView h, const u64 *__restrict__ loc, u32 n, u32 *err) {
    const u32 i = blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) h.insert(loc[i], i, err);
}

// ---- files -------------------------------------------------------------------------
enum : u32 { KIND_NODES = 1, KIND_LEAVES = 2, KIND_DUAL_CELLS = 3, KIND_OCCUPANCY = 4, KIND_SAMPLES = 5, KIND_CAMERAS = 6 };
constexpr u32 FILE_VERSION = 0;   // Nodes/Leaves/DualCells/Occupancy/Samples::FILE_VERSION

[[noreturn]] void fileError(const std::string &path, const std::string &what) {
    throw tsr_error(TSR_ERR_FILE, what + ": " + path);
}

struct File {
    std::FILE *f = nullptr;
    std::string path;
    u32 realBytes = 4;
    File(const char *p, bool write, u32 kind, u32 realBytesWrite = 4) : path(p ? p : "") {
        if (!p) throw tsr_error(TSR_ERR_INVALID_ARGUMENT, "null path");
        f = std::fopen(p, write ? "wb" : "rb");
        if (!f) fileError(path, write ? "cannot create the file" : "cannot open the file");
        u32 h[4];
        if (write) {
            std::memcpy(&h[0], "TSRF", 4);
            h[1] = kind;
            h[2] = FILE_VERSION;
            h[3] = realBytes = realBytesWrite;
            put(h, 4, 4);
        } else {
            get(h, 4, 4);
            if (std::memcmp(&h[0], "TSRF", 4) != 0 || h[1] != kind) fileError(path, "not a file of this kind");
            if (h[2] != FILE_VERSION) fileError(path, "unsupported file version");
            if (h[3] != 4 && h[3] != 8) fileError(path, "invalid Real size");
            realBytes = h[3];
        }
    }
    ~File() {
        if (f) std::fclose(f);
    }
    void put(const void *p, size_t elemBytes, size_t n) {
        if (n && std::fwrite(p, elemBytes, n, f) != n) fileError(path, "write failed");
    }
    void get(void *p, size_t elemBytes, size_t n) {
        if (n && std::fread(p, elemBytes, n, f) != n) fileError(path, "file too short");
    }
    void end() {
        if (std::fgetc(f) != EOF) fileError(path, "trailing data");
    }
    // Real arrays: float in a Real = float file, double otherwise
    void putReals(const std::vector<double> &v) {
        if (realBytes == 8) return put(v.data(), 8, v.size());
        std::vector<float> fl(v.begin(), v.end());
        put(fl.data(), 4, fl.size());
    }
    std::vector<double> getReals(size_t n) {
        std::vector<double> v(n);
        if (realBytes == 8) {
            get(v.data(), 8, n);
        } else {
            std::vector<float> fl(n);
            get(fl.data(), 4, n);
            for (size_t i = 0; i < n; ++i) v[i] = fl[i];
        }
        return v;
    }
};

template <typename T>
std::vector<T> toHost(const DBuf<T> &d, size_t n, hipStream_t st) {
    std::vector<T> h(n);
    if (n) HIPCHECK(hipMemcpyAsync(h.data(), d.p, n * sizeof(T), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    return h;
}

template <typename T>
void toDevice(DBuf<T> &d, const std::vector<T> &h, hipStream_t st) {
    d.alloc(h.size());
    if (!h.empty()) HIPCHECK(hipMemcpyAsync(d.p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, st));
    HIPCHECK(hipStreamSynchronize(st));
}

}  // namespace

void leafNeighbours(tsr_context *ctx, tsr_tree *tree) {
    if (tree->hasNeighbours) return;
    hipStream_t st = ctx->stream;
    Scratch &sc = ctx->scratch;
    const u32 L = tree->leafCount;
    tree->nbrOffsets.alloc((size_t)L + 1);
    if (L == 0) {
        HIPCHECK(hipMemsetAsync(tree->nbrOffsets.p, 0, 4, st));
        tree->nbrCount = 0;
        tree->hasNeighbours = true;
        return;
    }
    const u64 slots = 6ull * L;
    if (slots >= 0xffffffffull) throw tsr_error(TSR_ERR_NODE_OVERFLOW, "too many leaves for the neighbour graph");
    DBuf<u32> other(slots), valid(slots), pos(slots);
    if (tree->nodeHash.capacity == 0) {   // location code -> node id, built on first need
        DBuf<u32> err(1);
        HIPCHECK(hipMemsetAsync(err.p, 0, 4, st));
        tree->nodeHash.alloc(tree->nodeCount, st);
        k_insertNodes<<<gridFor(tree->nodeCount, BLOCK), BLOCK, 0, st>>>(tree->nodeHash.view(), tree->nodeLoc.p,
                                                                       tree->nodeCount, err.p);
        if (readScalar(err.p, st)) throw tsr_error(TSR_ERR_HIP, "node hash full");
    }
    k_neighbourSlots<<<gridFor(slots, BLOCK), BLOCK, 0, st>>>(tree->nodeHash.view(), tree->child.p, tree->leafLoc.p,
                                                              tree->leafNode.p, L, other.p);
    k_slotValid<<<gridFor(slots, BLOCK), BLOCK, 0, st>>>(other.p, slots, valid.p);
    exclusiveScan(sc, valid.p, pos.p, slots, st);
    const u64 edges = (u64)readScalar(pos.p + slots - 1, st) + readScalar(valid.p + slots - 1, st);
    const u64 m = 2 * edges;
    if (m >= 0xffffffffull) throw tsr_error(TSR_ERR_NODE_OVERFLOW, "too many neighbour edges");
    DBuf<u64> key(m), keySorted(m);
    DBuf<u32> val(m);
    tree->nbrs.alloc(std::max<u64>(m, 1));
    k_neighbourRecords<<<gridFor(slots, BLOCK), BLOCK, 0, st>>>(other.p, pos.p, slots, key.p, val.p);
    int bits = 32;
    while (bits < 64 && ((u64)L >> (bits - 32)) != 0) ++bits;
    sortPairs(sc, key.p, keySorted.p, val.p, tree->nbrs.p, m, 0, bits, st);
    k_ownerOffsets<<<gridFor((size_t)L + 1, BLOCK), BLOCK, 0, st>>>(keySorted.p, m, L, tree->nbrOffsets.p);
    HIPCHECK(hipStreamSynchronize(st));
    tree->nbrCount = m;
    tree->hasNeighbours = true;
}

// ---- tree files ----------------------------------------------------------------------
void saveTree(tsr_context *ctx, tsr_tree *tree, const char *nodesPath, const char *leavesPath,
              const char *cellsPath) {
    hipStream_t st = ctx->stream;
    const u32 N = tree->nodeCount, L = tree->leafCount;
    if (nodesPath) {   // Nodes.cpp:789-807
        std::vector<u32> c = toHost(tree->child, N, st);
        for (u32 &x : c) x = (x & LEAF_BIT) ? INVALID : x;
        File f(nodesPath, true, KIND_NODES);
        f.put(&N, 4, 1);
        f.put(c.data(), 4, N);
        f.put(toHost(tree->parents, N, st).data(), 4, N);
        f.put(toHost(tree->sampleCounts, N, st).data(), 4, N);
        f.put(toHost(tree->sampleStarts, N, st).data(), 4, N);
    }
    if (leavesPath) {   // Leaves.cpp:262-292
        leafNeighbours(ctx, tree);
        const std::vector<u64> loc = toHost(tree->leafLoc, L, st);
        const std::vector<u32> ln = toHost(tree->leafNode, L, st);
        // Scope (Scope.h:36-38): min corner (Vector3), size (Real), node index
        std::vector<unsigned char> scopes((size_t)L * 20);
        for (u32 l = 0; l < L; ++l) {
            u32 k[3];
            locCoords(loc[l], k[0], k[1], k[2]);
            const u32 d = locDepth(loc[l]);
            float v[4];
            for (int a = 0; a < 3; ++a) v[a] = (float)tree->frame.B(a, k[a], d);
            v[3] = (float)tree->frame.h(d);
            std::memcpy(&scopes[20 * (size_t)l], v, 16);
            std::memcpy(&scopes[20 * (size_t)l + 16], &ln[l], 4);
        }
        const u32 counts[3] = {N, L, (u32)tree->nbrCount};
        File f(leavesPath, true, KIND_LEAVES);
        f.put(counts, 4, 3);
        f.put(toHost(tree->nodeToLeaf, N, st).data(), 4, N);
        f.put(scopes.data(), 20, L);
        f.put(toHost(tree->nbrOffsets, (size_t)L + 1, st).data(), 4, (size_t)L + 1);
        f.put(toHost(tree->nbrs, tree->nbrCount, st).data(), 4, tree->nbrCount);
    }
    if (cellsPath) {   // DualCells.cpp:378-385
        const u32 indexCount = 8 * tree->cellCount;
        File f(cellsPath, true, KIND_DUAL_CELLS);
        f.put(&indexCount, 4, 1);
        f.put(toHost(tree->cells, indexCount, st).data(), 4, indexCount);
    }
}

// Tree::Tree(const Path&): the node, leaf and dual-cell files of a tree whose samples are the
// scene's, in node order (as Samples files hold them).  The root scope comes from the
// samples (Tree::getRootScope); node depths, location codes and the leaf list are derived
// from the child links.
void loadTree(tsr_context *ctx, const tsr_scene *scene, const tsr_params &prm, const char *nodesPath,
              const char *leavesPath, const char *cellsPath, tsr_tree *tree) {
    hipStream_t st = ctx->stream;
    u32 N = 0;
    std::vector<u32> children, parents, counts, starts;
    {
        File f(nodesPath, false, KIND_NODES);
        f.get(&N, 4, 1);
        if (N == 0 || N % 8 != 1) fileError(f.path, "invalid node count");
        children.resize(N); parents.resize(N); counts.resize(N); starts.resize(N);
        f.get(children.data(), 4, N);
        f.get(parents.data(), 4, N);
        f.get(counts.data(), 4, N);
        f.get(starts.data(), 4, N);
        f.end();
    }
    // depths, location codes, leaves in depth-first order; child blocks must lie inside the table
    std::vector<u64> loc(N, 0);
    std::vector<uint8_t> depth(N, 0);
    std::vector<u32> leafNode, nodeToLeaf(N, INVALID);
    {
        std::vector<u32> stack{0};
        loc[0] = 1;
        std::vector<uint8_t> seen(N, 0);
        seen[0] = 1;
        u32 rank = 0;   // pre-order rank of the inner nodes: their child blocks are 1 + 8 rank
        while (!stack.empty()) {
            const u32 n = stack.back();
            stack.pop_back();
            const u32 c = children[n];
            if (c == INVALID) {
                nodeToLeaf[n] = (u32)leafNode.size();
                leafNode.push_back(n);
                continue;
            }
            if (c == 0 || (u64)c + 8 > N || depth[n] + 1u > MAX_DEPTH) fileError(nodesPath, "invalid child link");
            if (c != 1 + 8 * rank++) fileError(nodesPath, "nodes are not in depth-first child-block order");
            for (int s = 7; s >= 0; --s) {
                const u32 cn = c + (u32)s;
                if (seen[cn] || parents[cn] != n) fileError(nodesPath, "inconsistent child/parent links");
                seen[cn] = 1;
                loc[cn] = (loc[n] << 3) | (u64)s;
                depth[cn] = depth[n] + 1;
                stack.push_back(cn);
            }
        }
        for (u32 n = 0; n < N; ++n)
            if (!seen[n]) fileError(nodesPath, "unreachable node");
    }
    const u32 L = (u32)leafNode.size();
    u64 total = 0;
    for (u32 n = 0; n < N; ++n) {
        if (starts[n] != total && counts[n]) fileError(nodesPath, "sample ranges are not contiguous");
        total += counts[n];
    }
    if (total != scene->n) fileError(nodesPath, "node sample counts do not match the scene's samples");
    std::vector<u32> nbrOffsets, nbrs;
    bool haveNbrs = false;
    if (leavesPath) {   // Leaves.cpp:230-260
        File f(leavesPath, false, KIND_LEAVES);
        u32 c3[3];
        f.get(c3, 4, 3);
        if (c3[0] != N || c3[1] != L) fileError(f.path, "leaf file does not match the node file");
        std::vector<u32> n2l(N);
        f.get(n2l.data(), 4, N);
        if (n2l != nodeToLeaf) fileError(f.path, "node-to-leaf links do not match the node file");
        std::vector<unsigned char> scopes((size_t)L * 20);
        f.get(scopes.data(), 20, L);
        nbrOffsets.resize((size_t)L + 1);
        f.get(nbrOffsets.data(), 4, (size_t)L + 1);
        if (nbrOffsets[0] != 0 || nbrOffsets[L] != c3[2]) fileError(f.path, "invalid neighbour offsets");
        nbrs.resize(c3[2]);
        f.get(nbrs.data(), 4, c3[2]);
        f.end();
        for (u32 x : nbrs)
            if (x >= L) fileError(f.path, "invalid neighbour index");
        haveNbrs = true;
    }
    std::vector<u32> cells;
    if (cellsPath) {   // DualCells.cpp:359-376
        File f(cellsPath, false, KIND_DUAL_CELLS);
        u32 indexCount = 0;
        f.get(&indexCount, 4, 1);
        if (indexCount % 8) fileError(f.path, "invalid dual-cell index count");
        cells.resize(indexCount);
        f.get(cells.data(), 4, indexCount);
        f.end();
        for (u32 x : cells)
            if (x >= L) fileError(f.path, "invalid leaf index in a dual cell");
    }

    const u32 n = scene->n;
    const float maxRel = prm.max_relative_sampling_distance;
    tree->maxRel = maxRel;
    tree->sampleCount = n;
    tree->frame = rootFrame(ctx, scene->px.p, scene->py.p, scene->pz.p, scene->scale.p, n, maxRel);
    tree->nodeCount = N;
    tree->innerCount = (N - 1) / 8;
    tree->leafCount = L;
    tree->cellCount = (u32)(cells.size() / 8);
    u32 md = 0;
    for (uint8_t d : depth) md = std::max<u32>(md, d);
    tree->maxDepth = md;
    std::vector<u32> child(children);
    for (u32 i = 0; i < N; ++i)
        if (child[i] == INVALID) child[i] = LEAF_BIT | nodeToLeaf[i];
    toDevice(tree->child, child, st);
    toDevice(tree->parents, parents, st);
    toDevice(tree->sampleCounts, counts, st);
    toDevice(tree->sampleStarts, starts, st);
    toDevice(tree->nodeToLeaf, nodeToLeaf, st);
    toDevice(tree->nodeDepth, depth, st);
    toDevice(tree->nodeLoc, loc, st);
    std::vector<u64> leafLoc(L);
    for (u32 l = 0; l < L; ++l) leafLoc[l] = loc[leafNode[l]];
    toDevice(tree->leafLoc, leafLoc, st);
    toDevice(tree->leafNode, leafNode, st);
    if (!cells.empty()) toDevice(tree->cells, cells, st);
    {
        DBuf<u32> err(1);
        HIPCHECK(hipMemsetAsync(err.p, 0, 4, st));
        tree->nodeHash.alloc(N, st);
        k_insertNodes<<<gridFor(N, BLOCK), BLOCK, 0, st>>>(tree->nodeHash.view(), tree->nodeLoc.p, N, err.p);
        if (readScalar(err.p, st)) throw tsr_error(TSR_ERR_HIP, "node hash full");
    }
    // samples: already in node order
    std::vector<u32> ident(n);
    for (u32 i = 0; i < n; ++i) ident[i] = i;
    toDevice(tree->sampleTarget, ident, st);
    auto copyF = [&](const DBuf<float> &src, DBuf<float> &dst) {
        dst.alloc(n);
        HIPCHECK(hipMemcpyAsync(dst.p, src.p, (size_t)n * 4, hipMemcpyDeviceToDevice, st));
    };
    copyF(scene->px, tree->px); copyF(scene->py, tree->py); copyF(scene->pz, tree->pz);
    copyF(scene->nx, tree->nx); copyF(scene->ny, tree->ny); copyF(scene->nz, tree->nz);
    copyF(scene->scale, tree->scale); copyF(scene->conf, tree->conf);
    tree->maxCams = scene->maxCams;
    tree->cams.alloc((size_t)n * scene->maxCams);
    HIPCHECK(hipMemcpyAsync(tree->cams.p, scene->cams.p, (size_t)n * scene->maxCams * 4, hipMemcpyDeviceToDevice, st));
    {
        DBuf<u64> cloc(n);
        tree->sampleDepth.alloc(n);
        sampleContainers(ctx, tree->frame, tree->px.p, tree->py.p, tree->pz.p, tree->scale.p, n, maxRel, cloc.p,
                         tree->sampleDepth.p);
    }
    if (haveNbrs) {
        toDevice(tree->nbrOffsets, nbrOffsets, st);
        tree->nbrs.alloc(std::max<size_t>(nbrs.size(), 1));
        if (!nbrs.empty())
            HIPCHECK(hipMemcpyAsync(tree->nbrs.p, nbrs.data(), nbrs.size() * 4, hipMemcpyHostToDevice, st));
        tree->nbrCount = nbrs.size();
        tree->hasNeighbours = true;
    }
    HIPCHECK(hipStreamSynchronize(st));
}

// ---- occupancy files (Occupancy.cpp:663-766) ------------------------------------------
void saveField(tsr_context *ctx, const tsr_field *field, const char *path, int realBytes) {
    hipStream_t st = ctx->stream;
    const u32 L = field->leafCount, N = field->nodeCount;
    const std::vector<float> ks = toHost(field->ksum, 2 * (size_t)L, st), cl = toHost(field->clen, 2 * (size_t)L, st);
    const std::vector<double> pr = toHost(field->prior, L, st);
    File f(path, true, KIND_OCCUPANCY, (u32)realBytes);
    const u32 counts[2] = {L, N};
    f.put(counts, 4, 2);
    for (int type = 0; type < 2; ++type) {   // EMPTINESS, SAMPLENESS: cone lengths, kernel sums
        f.putReals(std::vector<double>(cl.begin() + (size_t)type * L, cl.begin() + (size_t)(type + 1) * L));
        f.putReals(std::vector<double>(ks.begin() + (size_t)type * L, ks.begin() + (size_t)(type + 1) * L));
    }
    f.putReals(pr);
    f.put(toHost(field->states, N, st).data(), 4, N);
    f.putReals(std::vector<double>{field->coneCount});
}

void loadField(tsr_context *ctx, const tsr_tree *tree, const tsr_params &prm, const char *path, tsr_field *field) {
    hipStream_t st = ctx->stream;
    const u32 L = tree->leafCount, N = tree->nodeCount;
    File f(path, false, KIND_OCCUPANCY);
    u32 counts[2];
    f.get(counts, 4, 2);
    if (counts[0] != L || counts[1] != N) fileError(f.path, "Invalid occupancy file format.");
    std::vector<float> ks(2 * (size_t)L), cl(2 * (size_t)L);
    for (int type = 0; type < 2; ++type) {
        const std::vector<double> c = f.getReals(L), k = f.getReals(L);
        for (u32 l = 0; l < L; ++l) {
            cl[(size_t)type * L + l] = (float)c[l];
            ks[(size_t)type * L + l] = (float)k[l];
        }
    }
    const std::vector<double> pr = f.getReals(L);
    std::vector<u32> states(N);
    f.get(states.data(), 4, N);
    field->coneCount = f.getReals(1)[0];
    f.end();
    field->leafCount = L;
    field->nodeCount = N;
    field->params = prm;
    toDevice(field->ksum, ks, st);
    toDevice(field->clen, cl, st);
    toDevice(field->prior, pr, st);
    toDevice(field->states, states, st);
    field->occ.alloc(L);
    occupancyValues(ctx, field, (double)prm.confidence_threshold);
}

// ---- samples files (Samples.cpp:693-765): the tree's samples in node order -------------
void saveTreeSamples(tsr_context *ctx, const tsr_tree *tree, const char *path) {
    hipStream_t st = ctx->stream;
    const u32 n = tree->sampleCount, mc = tree->maxCams;
    auto v3 = [&](const DBuf<float> &x, const DBuf<float> &y, const DBuf<float> &z) {
        const std::vector<float> a = toHost(x, n, st), b = toHost(y, n, st), c = toHost(z, n, st);
        std::vector<float> out(3 * (size_t)n);
        for (u32 i = 0; i < n; ++i) {
            out[3 * (size_t)i] = a[i];
            out[3 * (size_t)i + 1] = b[i];
            out[3 * (size_t)i + 2] = c[i];
        }
        return out;
    };
    const std::vector<float> pos = v3(tree->px, tree->py, tree->pz), nrm = v3(tree->nx, tree->ny, tree->nz);
    float aabb[6] = {3.4028235e38f, 3.4028235e38f, 3.4028235e38f, -3.4028235e38f, -3.4028235e38f, -3.4028235e38f};
    for (u32 i = 0; i < n; ++i)
        for (int a = 0; a < 3; ++a) {
            aabb[a] = std::min(aabb[a], pos[3 * (size_t)i + a]);
            aabb[3 + a] = std::max(aabb[3 + a], pos[3 * (size_t)i + a]);
        }
    File f(path, true, KIND_SAMPLES);
    f.put(&n, 4, 1);
    f.put(&mc, 4, 1);
    f.put(aabb, 4, 6);
    const std::vector<float> colours(3 * (size_t)n, 0.f);   // the scene carries no colours
    f.put(colours.data(), 12, n);
    f.put(nrm.data(), 12, n);
    f.put(pos.data(), 12, n);
    f.put(toHost(tree->conf, n, st).data(), 4, n);
    f.put(toHost(tree->scale, n, st).data(), 4, n);
    f.put(toHost(tree->cams, (size_t)n * mc, st).data(), 4, (size_t)n * mc);
}

// Cameras::saveToFile / loadFromFile (Scene/Camera/Cameras.cpp:116-177): the camera count,
// then one CameraData record per camera (CameraData.h:21-36) as the struct lays it out: 13
// Reals (orientation quaternion, position, principal point, 2 distortion coefficients, focal
// length, image aspect ratio) and the view ID, padded to the Real size.
void camerasFileWrite(const char *path, u32 count, const double *data13, const u32 *viewIds, int realBytes) {
    File f(path, true, KIND_CAMERAS, (u32)realBytes);
    f.put(&count, 4, 1);
    const u32 pad = realBytes == 8 ? 4 : 0;
    const u32 zero = 0;
    for (u32 c = 0; c < count; ++c) {
        f.putReals(std::vector<double>(data13 + 13 * (size_t)c, data13 + 13 * (size_t)c + 13));
        f.put(&viewIds[c], 4, 1);
        if (pad) f.put(&zero, 4, 1);
    }
}

void camerasFileInfo(const char *path, u32 &count) {
    File f(path, false, KIND_CAMERAS);
    f.get(&count, 4, 1);
}

void camerasFileRead(const char *path, double *data13, u32 *viewIds) {
    File f(path, false, KIND_CAMERAS);
    u32 count = 0;
    f.get(&count, 4, 1);
    for (u32 c = 0; c < count; ++c) {
        const std::vector<double> r = f.getReals(13);
        if (data13) std::copy(r.begin(), r.end(), data13 + 13 * (size_t)c);
        u32 id = 0;
        f.get(&id, 4, 1);
        if (viewIds) viewIds[c] = id;
        if (f.realBytes == 8) f.get(&id, 4, 1);
    }
    f.end();
}

// header of a samples file: count and cameras per sample
void samplesFileInfo(const char *path, u32 &count, u32 &maxCams) {
    File f(path, false, KIND_SAMPLES);
    f.get(&count, 4, 1);
    f.get(&maxCams, 4, 1);
}

void samplesFileRead(const char *path, float *px, float *py, float *pz, float *nx, float *ny, float *nz,
                     float *scale, float *conf, u32 *cams) {
    File f(path, false, KIND_SAMPLES);
    u32 n = 0, mc = 0;
    f.get(&n, 4, 1);
    f.get(&mc, 4, 1);
    float aabb[6];
    f.get(aabb, 4, 6);
    std::vector<float> tmp(3 * (size_t)n);
    f.get(tmp.data(), 12, n);   // colours (unused)
    auto split = [&](float *x, float *y, float *z) {
        f.get(tmp.data(), 12, n);
        for (u32 i = 0; i < n; ++i) {
            if (x) x[i] = tmp[3 * (size_t)i];
            if (y) y[i] = tmp[3 * (size_t)i + 1];
            if (z) z[i] = tmp[3 * (size_t)i + 2];
        }
    };
    split(nx, ny, nz);
    split(px, py, pz);
    std::vector<float> one(n);
    f.get(one.data(), 4, n);
    if (conf) std::memcpy(conf, one.data(), (size_t)n * 4);
    f.get(one.data(), 4, n);
    if (scale) std::memcpy(scale, one.data(), (size_t)n * 4);
    std::vector<u32> c((size_t)n * mc);
    f.get(c.data(), 4, (size_t)n * mc);
    if (cams) std::memcpy(cams, c.data(), c.size() * 4);
    f.end();
}

namespace {
constexpr u32 ERASE_BLOCK = 256;
// erased[i] = 1 for the samples to drop
__global__ void k_erasedFlags(const uint8_t *__restrict__ inliers, u32 n, u32 *__restrict__ erased) {
    const u32 i = blockIdx.x * ERASE_BLOCK + threadIdx.x;
    if (i < n) erased[i] = inliers[i] ? 0u : 1u;
}
// Nodes::eraseSamples (Nodes.cpp:644-671): offsets[i] = erased samples before i (n + 1 entries)
__global__ void k_eraseInNodes(u32 *__restrict__ counts, u32 *__restrict__ starts, u32 nodeCount,
                               const u32 *__restrict__ offsets) {
    const u32 nd = blockIdx.x * ERASE_BLOCK + threadIdx.x;
    if (nd >= nodeCount) return;
    const u32 start = starts[nd], end = start + counts[nd];
    counts[nd] -= offsets[end] - offsets[start];
    starts[nd] = start - offsets[start];
}
// Samples::compact: every kept sample moves offsets[i] places towards the front
__global__ void k_compactSamples(const uint8_t *__restrict__ inliers, u32 n, const u32 *__restrict__ offsets,
                                 u32 maxCams, const float *const *__restrict__ src, float *const *__restrict__ dst,
                                 const u32 *__restrict__ cams, u32 *__restrict__ camsOut,
                                 const uint8_t *__restrict__ depth, uint8_t *__restrict__ depthOut) {
    const u32 i = blockIdx.x * ERASE_BLOCK + threadIdx.x;
    if (i >= n || !inliers[i]) return;
    const u32 j = i - offsets[i];
    for (int a = 0; a < 8; ++a) dst[a][j] = src[a][i];
    for (u32 k = 0; k < maxCams; ++k) camsOut[(size_t)j * maxCams + k] = cams[(size_t)i * maxCams + k];
    depthOut[j] = depth[i];
}
// input index -> new node-order index (INVALID for erased samples)
__global__ void k_remapTargets(u32 *__restrict__ target, u32 m, const uint8_t *__restrict__ inliers,
                               const u32 *__restrict__ offsets) {
    const u32 k = blockIdx.x * ERASE_BLOCK + threadIdx.x;
    if (k >= m) return;
    const u32 t = target[k];
    if (t == INVALID) return;
    target[k] = inliers[t] ? t - offsets[t] : INVALID;
}
}  // namespace

// Tree::eraseSamples (Tree.cpp:50-53, Nodes.cpp:644-671) and Samples::compact, as called by
// Scene::eraseSamples (Scene.cpp:270-296): the node sample counts and starts follow the
// erased samples, and the tree's node-ordered sample arrays are compacted.
void eraseTreeSamples(tsr_context *ctx, tsr_tree *tree, const uint8_t *inliersHost) {
    hipStream_t st = ctx->stream;
    tree->eIndex.reset();   // built over the samples before the erase
    const u32 n = tree->sampleCount, N = tree->nodeCount, mc = tree->maxCams;
    if (n == 0) return;
    DBuf<uint8_t> inl(n);
    HIPCHECK(hipMemcpyAsync(inl.p, inliersHost, n, hipMemcpyHostToDevice, st));
    DBuf<u32> erased(n), offsets((size_t)n + 1);
    k_erasedFlags<<<gridFor(n, ERASE_BLOCK), ERASE_BLOCK, 0, st>>>(inl.p, n, erased.p);
    exclusiveScan(ctx->scratch, erased.p, offsets.p, n, st);
    const u32 removed = readScalar(offsets.p + n - 1, st) + readScalar(erased.p + n - 1, st);
    HIPCHECK(hipMemcpyAsync(offsets.p + n, &removed, 4, hipMemcpyHostToDevice, st));
    k_eraseInNodes<<<gridFor(N, ERASE_BLOCK), ERASE_BLOCK, 0, st>>>(tree->sampleCounts.p, tree->sampleStarts.p, N,
                                                                    offsets.p);
    const u32 m = n - removed;
    DBuf<float> px(m), py(m), pz(m), nx(m), ny(m), nz(m), sc(m), cf(m);
    DBuf<u32> cams((size_t)m * mc);
    DBuf<uint8_t> depth(m);
    const float *src[8] = {tree->px.p, tree->py.p, tree->pz.p, tree->nx.p, tree->ny.p, tree->nz.p, tree->scale.p,
                           tree->conf.p};
    float *dst[8] = {px.p, py.p, pz.p, nx.p, ny.p, nz.p, sc.p, cf.p};
    DBuf<const float *> srcD(8);
    DBuf<float *> dstD(8);
    HIPCHECK(hipMemcpyAsync(srcD.p, src, sizeof(src), hipMemcpyHostToDevice, st));
    HIPCHECK(hipMemcpyAsync(dstD.p, dst, sizeof(dst), hipMemcpyHostToDevice, st));
    k_compactSamples<<<gridFor(n, ERASE_BLOCK), ERASE_BLOCK, 0, st>>>(inl.p, n, offsets.p, mc, srcD.p, dstD.p,
                                                                      tree->cams.p, cams.p, tree->sampleDepth.p,
                                                                      depth.p);
    const u32 targets = (u32)tree->sampleTarget.n;
    if (targets)
        k_remapTargets<<<gridFor(targets, ERASE_BLOCK), ERASE_BLOCK, 0, st>>>(tree->sampleTarget.p, targets, inl.p,
                                                                              offsets.p);
    HIPCHECK(hipStreamSynchronize(st));
    std::swap(tree->px, px); std::swap(tree->py, py); std::swap(tree->pz, pz);
    std::swap(tree->nx, nx); std::swap(tree->ny, ny); std::swap(tree->nz, nz);
    std::swap(tree->scale, sc); std::swap(tree->conf, cf);
    std::swap(tree->cams, cams);
    std::swap(tree->sampleDepth, depth);
    tree->sampleCount = m;
}

}  // namespace tsr
