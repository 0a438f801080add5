// Handle structs behind the opaque C ABI types.
#pragma once

#include <functional>
#include <vector>

#include "device_util.h"

// live timing + algorithmic traffic of selected kernels (TSR_KSTAT_*)
struct tsr_kernel_stat {
    double ms = 0;
    uint64_t launches = 0, bytes = 0, units = 0;
};

namespace tsr {
struct EmptinessIndex;   // direction index of the gathered emptiness pass (emptiness.hip)
// buffers of the cone passes, kept by the context between passes and calls (grow-only):
// repeated allocation and release of tens of GB costs more than the passes themselves
struct ConeArena {
    DBuf<u32> stLeaf, recLeaf, recLeafS;
    DBuf<uint2> stVal, recVal, recValS;
    DBuf<u32> childPad, blockParent;   // traversal copies of the node arrays
    DBuf<uint2> packed;
    void release() {
        stLeaf.release(); recLeaf.release(); recLeafS.release();
        stVal.release(); recVal.release(); recValS.release();
        childPad.release(); blockParent.release(); packed.release();
    }
};
}  // namespace tsr

struct tsr_context {
    int device = 0;
    hipStream_t stream = nullptr;
    // second stream (default priority): the emptiness index is built on it, from a worker
    // thread, while the tree build replays the sampling refinement on the host
    hipStream_t stream2 = nullptr;
    hipEvent_t ev[12] = {};   // start/end pair per phase
    hipEvent_t kev[2] = {};   // start/end of an instrumented kernel
    float phaseMs[6] = {0, 0, 0, 0, 0, 0};
    tsr_kernel_stat kstat[8];
    std::shared_ptr<tsr::DevicePool> pool;   // device blocks reused across the stages' temporaries
    // blocks of the work on stream2 (the emptiness index built during the tree build), kept
    // for the next build instead of being mapped and unmapped each time
    std::shared_ptr<tsr::DevicePool> pool2;
    tsr::Scratch scratch;
    tsr::ConeArena cone;
    // pinned host staging for large device->host transfers, one grow-only buffer per use
    // (slot 0: the sampling visit index, slot 1: the containment tree's inner codes, slot 2:
    // the sampling event flags as bits), freed with the context
    void *pinned[3] = {nullptr, nullptr, nullptr};
    size_t pinnedBytes[3] = {0, 0, 0};
};

struct tsr_scene {
    tsr_context *ctx = nullptr;
    int device = 0;   // destroy needs only the device, not the context
    unsigned long long id = 0;   // unique per scene of the process (indexes built from it name it)
    tsr::u32 n = 0, maxCams = 1, camCount = 0;
    tsr::DBuf<float> px, py, pz, nx, ny, nz, scale, conf;
    tsr::DBuf<tsr::u32> cams;
    tsr::DBuf<float> cx, cy, cz;
    // host copies of the geometry (input order), read by the order-dependent
    // replay of the sampling refinement
    std::vector<float> hpx, hpy, hpz, hnx, hny, hnz, hscale;
};

struct tsr_tree {
    tsr_context *ctx = nullptr;
    int device = 0;   // destroy needs only the device, not the context
    tsr::Frame frame{};
    float maxRel = 2.0f;
    tsr::u32 nodeCount = 0, innerCount = 0, leafCount = 0, cellCount = 0, maxDepth = 0, sampleCount = 0;
    tsr::DBuf<tsr::u32> child;         // per node: first child id, or LEAF_BIT | leaf index
    tsr::DBuf<tsr::u32> parents;       // Nodes::mParents
    tsr::DBuf<tsr::u32> sampleCounts;  // Nodes::mSamplesPerNodes
    tsr::DBuf<tsr::u32> sampleStarts;  // Nodes::mSampleStartIndices
    tsr::DBuf<tsr::u32> nodeToLeaf;    // Leaves::mNodeToLeafLinks
    tsr::DBuf<uint8_t> nodeDepth;
    tsr::DBuf<tsr::u64> nodeLoc;       // location code per node id
    tsr::DBuf<tsr::u64> leafLoc;       // location code per leaf (DFS order)
    tsr::DBuf<tsr::u32> leafNode;      // node id per leaf
    tsr::DBuf<tsr::u32> cells;         // 8 leaf indices per dual cell (DualCells::mIndices)
    tsr::DBuf<tsr::u32> sampleTarget;  // new index of each original sample
    tsr::DBuf<uint8_t> sampleDepth;    // container depth per sample (tree order)
    // the scene's samples permuted into node order (Samples::reorder, Samples.cpp:825-835)
    tsr::u32 maxCams = 1;
    tsr::DBuf<float> px, py, pz, nx, ny, nz, scale, conf;
    tsr::DBuf<tsr::u32> cams;
    tsr::DHash nodeHash;               // location code -> node id (all nodes), built when first needed
    tsr::u64 samplingEvents = 0, samplingCreated = 0, samplingUnits = 0;
    tsr::u64 samplingMisses = 0;   // visit-index lookups answered by the device (leaves outside the hot part)
    // the emptiness pass's direction index over the scene's links, built during the tree build
    // (null after an erase, for a loaded tree, or when it did not fit)
    std::shared_ptr<tsr::EmptinessIndex> eIndex;
    // leaf face-neighbour graph (Leaves.cpp:82-180), CSR built on request or loaded
    bool hasNeighbours = false;
    tsr::DBuf<tsr::u32> nbrOffsets, nbrs;
    tsr::u64 nbrCount = 0;
};

struct tsr_field {
    tsr_context *ctx = nullptr;
    int device = 0;   // destroy needs only the device, not the context
    tsr::u32 leafCount = 0, nodeCount = 0;
    tsr::DBuf<float> ksum;     // [2][L] kernel sums, 0 = EMPTINESS, 1 = SAMPLENESS
    tsr::DBuf<float> clen;     // [2][L] cone length sums
    tsr::DBuf<double> prior;   // [L] prior for emptiness, E / (E + S) in FP64
    tsr::DBuf<tsr::u32> states;
    tsr::DBuf<double> occ;     // [L] occupancy (FP64)
    double coneCount = 0;
    tsr::u64 hits[2] = {0, 0};
    tsr::u32 priorQueries = 0;
    tsr::u32 priorsOverflowGroups = 0;   // query groups redone with the long traversal stack
    tsr::DBuf<tsr::u32> hitsS;   // sampleness hits per leaf
    tsr::u32 overflowLinks[2] = {0, 0};   // links traversed twice (more hits than staging slots), E / S
    // certified priors: per leaf the stage that decided it (0, 1: capped bounds, 2: exact,
    // 255: not queried); queries, entering stage 0 / 1 / 2, sent by the network to stage 1 / 2
    tsr::DBuf<uint8_t> levels;
    tsr::u32 priorStages[6] = {0, 0, 0, 0, 0, 0};
    tsr_params params{};
    int flags = 0;   // tsr_occupancy_compute flags of the field
};

struct tsr_mesh {
    tsr_context *ctx = nullptr;
    int device = 0;   // destroy needs only the device, not the context
    tsr::u32 vertexCount = 0, triangleCount = 0;
    tsr::DBuf<double> pos;      // 3 per vertex
    tsr::DBuf<double> scale;    // 1 per vertex
    tsr::DBuf<tsr::u32> keys;   // 2 per vertex
    tsr::DBuf<tsr::u32> tris;   // 3 per triangle
    tsr::DBuf<double> normals;  // 3 per vertex once post-processed
};

namespace tsr {
// pinned host memory of the context's staging slot, at least `bytes` (contents not
// preserved on growth)
void *pinnedHost(tsr_context *ctx, size_t bytes, int slot = 0);
// runs fn(begin, end) over [0, n) on up to 16 host threads (OMP_NUM_THREADS caps it),
// at least `grain` items per thread
void parallelFor(size_t n, const std::function<void(size_t, size_t)> &fn, size_t grain = 65536);
enum { KSTAT_PRIORS = 0, KSTAT_E_COUNT = 1, KSTAT_E_WRITE = 2, KSTAT_S_COUNT = 3, KSTAT_S_WRITE = 4, KSTAT_E_GATHER = 5 };
// time one kernel launch on the context stream with HIP events (synchronises)
template <typename Launch>
void timedLaunch(tsr_context *ctx, int stat, Launch launch) {
    HIPCHECK(hipEventRecord(ctx->kev[0], ctx->stream));
    launch();
    HIPCHECK(hipEventRecord(ctx->kev[1], ctx->stream));
    HIPCHECK(hipEventSynchronize(ctx->kev[1]));
    float ms = 0.f;
    HIPCHECK(hipEventElapsedTime(&ms, ctx->kev[0], ctx->kev[1]));
    ctx->kstat[stat].ms += ms;
    ctx->kstat[stat].launches += 1;
}
// phase: 0 tree, 1 S pass, 2 E pass, 3 priors, 4 occupancy, 5 DMC
inline void recordPhase(tsr_context *ctx, int phase, bool start) {
    HIPCHECK(hipEventRecord(ctx->ev[2 * phase + (start ? 0 : 1)], ctx->stream));
}
void buildTree(tsr_context *ctx, const tsr_scene *scene, const tsr_params &prm, tsr_tree *tree);
// root scope of a sample set (Tree.cpp:70-101); checks the scales (Samples.cpp:280-288)
Frame rootFrame(tsr_context *ctx, const float *px, const float *py, const float *pz, const float *scale, u32 n,
                float maxRel);
// container location code and depth per sample (device arrays of n)
void sampleContainers(tsr_context *ctx, const Frame &f, const float *px, const float *py, const float *pz,
                      const float *scale, u32 n, float maxRel, u64 *loc, uint8_t *depth);
// FP64 occupancy ratio of every leaf from the field's sums and priors (Occupancy.cpp:570-587)
void occupancyValues(tsr_context *ctx, const tsr_field *field, double confidenceThreshold);
// Leaves::createNeighborsEdges on the GPU, kept on the tree (no-op when present)
void leafNeighbours(tsr_context *ctx, tsr_tree *tree);
// on-disk formats (storage.hip)
void saveTree(tsr_context *ctx, tsr_tree *tree, const char *nodesPath, const char *leavesPath, const char *cellsPath);
void loadTree(tsr_context *ctx, const tsr_scene *scene, const tsr_params &prm, const char *nodesPath,
              const char *leavesPath, const char *cellsPath, tsr_tree *tree);
void saveField(tsr_context *ctx, const tsr_field *field, const char *path, int realBytes);
void loadField(tsr_context *ctx, const tsr_tree *tree, const tsr_params &prm, const char *path, tsr_field *field);
void saveTreeSamples(tsr_context *ctx, const tsr_tree *tree, const char *path);
void samplesFileInfo(const char *path, u32 &count, u32 &maxCams);
void camerasFileWrite(const char *path, u32 count, const double *data13, const u32 *viewIds, int realBytes);
void camerasFileInfo(const char *path, u32 &count);
void camerasFileRead(const char *path, double *data13, u32 *viewIds);
void samplesFileRead(const char *path, float *px, float *py, float *pz, float *nx, float *ny, float *nz,
                     float *scale, float *conf, u32 *cams);
void computeOccupancy(tsr_context *ctx, const tsr_tree *tree, const tsr_scene *scene, const tsr_params &prm,
                      int flags, const tsr_culler *culler, tsr_field *field);
// E pass gathered by the sampleness-flagged leaves (hitsS[leaf] > 0): per-leaf kernel sums,
// cone length sums and hit counts (L each, device), the confidence of every link in the
// tree's order (0: no emptiness cone; device, links) and the total hit count.  Uses the
// tree's direction index when it was built from this scene with this bandwidth.
void emptinessByLeaves(tsr_context *ctx, const tsr_tree *tree, const tsr_scene *scene, const tsr_params &prm,
                       const u32 *hitsS, float *ksum, float *clen, u32 *hitCount, double *confs, u64 &hitsOut);
// the links an emptiness index is built over: samples (SoA, device) with their camera links
// and container depths, in one order; cameras; the kernel bandwidth factor
struct EmptinessLinks {
    const float *px, *py, *pz, *nx, *ny, *nz, *scale, *conf;
    const u32 *cams;
    const float *cx, *cy, *cz;
    const uint8_t *sampleDepth;
    u32 samples, maxCams, camCount;
    float bandwidth;
    unsigned long long sceneId;
};
// direction index of the gathered E pass, built on stream st with scratch sc (the caller's
// current block pool); null when its cell grid cannot fit its budget
std::shared_ptr<EmptinessIndex> buildEmptinessIndex(hipStream_t st, Scratch &sc, const EmptinessLinks &src);
// Occupancy::addSamples / eraseSamples: the field follows a change of samples (indices into
// the tree's node-ordered samples, host array of n)
void updateOccupancy(tsr_context *ctx, const tsr_tree *tree, const tsr_scene *scene, const u32 *chosenHost, u32 n,
                     bool removal, tsr_field *field);
// Tree::eraseSamples + Samples::compact: drops the samples with inliers[i] == 0 (node order,
// host array of the tree's sample count)
void eraseTreeSamples(tsr_context *ctx, tsr_tree *tree, const uint8_t *inliersHost);
// Traversal records of the priors, node-padded (node i at index i + NODE_PAD = 7, so that
// the 8 children of a node are one aligned run): (child link, state | deepest leaf depth
// << 8) and the FP64 subtree sums (E, S) of the cone lengths
struct PriorRecords {
    DBuf<float4> sumsPad;   // FP64 pairs after packing
    DBuf<uint2> linksPad;
    const uint2 *links() const { return linksPad.p + 7; }
    const double2 *sums() const { return reinterpret_cast<const double2 *>(sumsPad.p + 7); }
};
void priorRecords(tsr_context *ctx, const tsr_tree *tree, const tsr_field *field, PriorRecords &r);
// priors of the queried leaves, plain or certified (TSR_PRIORS_CERTIFIED / TSR_CULL_LEARNED)
void priors(tsr_context *ctx, const tsr_tree *tree, const tsr_params &prm, int flags, const tsr_culler *culler,
            tsr_field *field);
// TSR_LEAF_FEATURES floats per listed leaf (all leaves when list is null) into X (device)
void leafFeatures(tsr_context *ctx, const tsr_tree *tree, const tsr_field *field, const PriorRecords &rec,
                  const u32 *list, u32 n, float *X);
// first stage of the certified cascade per listed query from the network (stage[leaf]);
// routed[0] / [1] count the queries sent to stage 1 / 2
void leafRoute(tsr_context *ctx, const tsr_culler *culler, const tsr_tree *tree, const tsr_field *field,
               const PriorRecords &rec, const u32 *queries, u32 nq, uint8_t *stage, u32 *routed);
// DualMarchingCells::extractMesh over the per-leaf occupancy `occ` (device, FP64)
void extractDMC(tsr_context *ctx, const tsr_tree *tree, const double *occ, tsr_mesh *mesh);
// the same into caller-provided device buffers, temporaries in the no-allocation scratch
void extractDMCNoAlloc(tsr_context *ctx, const tsr_tree *tree, const double *occ, u32 maxVertices, u32 maxTriangles,
                       double *pos, double *scale, u32 *keys, u32 *tris, u32 *vertexCount, u32 *triangleCount);
// device address of the no-allocation scratch array on `device` (looked up once per device;
// that device must be current)
void *noAllocScratch(int device);
void calibrationCopy(tsr_context *ctx, u64 bytes);
// DualMarchingCells.cpp:26-39 after extractMesh: isle removal, hole filling, normals
void postprocessCrust(tsr_context *ctx, tsr_mesh *mesh, u32 minIsleSize, bool fill, bool normals);
}  // namespace tsr


// set the thread-local message returned by tsr_last_error
void tsrSetLastError(const char *msg);
