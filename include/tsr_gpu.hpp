// tsr_gpu.hpp - C++ face of the GPU occupancy + DMC library, shaped like the
// reference's hot-path classes so that Scene::reconstruct-style code reads the
// same (paths relative to SurfaceReconstruction/SurfaceReconstruction/):
//
//   tsrgpu::Tree              Scene/Tree/Tree.h:26-75       (getNodes, getLeaves, getDualCells,
//                                                            getRootScope, getRootNodeSize, getCenter,
//                                                            getNodeDepth, getSampleDepth, getLeafIndex)
//   tsrgpu::Nodes             Scene/Tree/Nodes.h:66-89      (getChildBlock, getChildScope, getCount,
//                                                            getSampleCount, isLeaf)
//   tsrgpu::Leaves            Scene/Tree/Leaves.h:36-67     (getCount, getLeafIndex, getNodeToLeafLinks,
//                                                            getScope, getScopes)
//   tsrgpu::DualCells         Scene/Tree/DualCells.h:33-35  (getCellCount, getIndexCount, getIndices)
//   tsrgpu::Scope             Scene/Tree/Scope.h:20-81
//   tsrgpu::Occupancy         SurfaceExtraction/Occupancy.h:49-90 (getOccupancy, getEmptinessCPD,
//                                                            getSamplenessCPD, getObservationPD,
//                                                            getPriorForEmptiness/Sampleness,
//                                                            getNodeStates, extractCrust, getCrust)
//   tsrgpu::DualMarchingCells SurfaceExtraction/DualMarchingCells.h:34-37 (getSurface)
//
// The reference reads its inputs from the Scene singleton; here a tsrgpu::Scene
// (device copy of the samples and cameras) is passed explicitly.  Errors throw
// tsrgpu::Exception carrying the C ABI status (the reference throws
// FailureHandling exceptions for the same conditions).  Host tables are
// downloaded once, when a Tree / Occupancy is constructed.
#ifndef TSR_GPU_HPP
#define TSR_GPU_HPP

#include <cstdint>
#include <cmath>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "tsr_gpu.h"

namespace tsrgpu {

typedef uint32_t uint32;
typedef float Real;   // the reference's default Real (float unless DOUBLE_PRECISION)

constexpr uint32 INVALID_INDEX = 0xFFFFFFFFu;

class Exception : public std::runtime_error {
public:
    Exception(tsr_status status, const std::string &message)
        : std::runtime_error(message), mStatus(status) {}
    tsr_status getStatus() const { return mStatus; }

private:
    tsr_status mStatus;
};

inline void check(tsr_status s) {
    if (s != TSR_OK) throw Exception(s, tsr_last_error());
}

struct Vector3 {
    Real x = 0, y = 0, z = 0;
    Vector3() = default;
    Vector3(Real x_, Real y_, Real z_) : x(x_), y(y_), z(z_) {}
};

// Scope.h:20-81: cube minimum corner, side length and node index
class Scope {
public:
    Scope() = default;
    Scope(const Vector3 &positionWS, Real size, uint32 idx) : mPositionWS(positionWS), mSize(size), mIdx(idx) {}
    Vector3 getCenterPosition() const {
        const Real h = 0.5f * mSize;
        return Vector3(mPositionWS.x + h, mPositionWS.y + h, mPositionWS.z + h);
    }
    const Vector3 &getMinimumCoordinates() const { return mPositionWS; }
    Vector3 getMaximumCoordinates() const {
        return Vector3(mPositionWS.x + mSize, mPositionWS.y + mSize, mPositionWS.z + mSize);
    }
    uint32 getNodeIndex() const { return mIdx; }
    Real getSize() const { return mSize; }

private:
    Vector3 mPositionWS;
    Real mSize = 0;
    uint32 mIdx = INVALID_INDEX;
};

// One HIP device and stream.
class Context {
public:
    explicit Context(int device = 0) { check(tsr_context_create(device, &mCtx)); }
    ~Context() { tsr_context_destroy(mCtx); }
    Context(const Context &) = delete;
    Context &operator=(const Context &) = delete;
    tsr_context *get() const { return mCtx; }
    void synchronize() const { check(tsr_context_synchronize(mCtx)); }

private:
    tsr_context *mCtx = nullptr;
};

// Device copy of the Scene singleton's samples and cameras.
class Scene {
public:
    Scene(const Context &ctx, const tsr_samples_view &samples, const tsr_cameras_view &cameras)
        : mCtx(&ctx), mSampleCount(samples.count), mMaxCams(samples.max_cameras_per_sample) {
        check(tsr_scene_upload(ctx.get(), &samples, &cameras, &mScene));
    }
    ~Scene() { tsr_scene_destroy(mScene); }
    Scene(const Scene &) = delete;
    Scene &operator=(const Scene &) = delete;
    const Context &getContext() const { return *mCtx; }
    tsr_scene *get() const { return mScene; }
    uint32 getSampleCount() const { return mSampleCount; }
    uint32 getMaxCamerasPerSample() const { return mMaxCams; }

private:
    const Context *mCtx;
    tsr_scene *mScene = nullptr;
    uint32 mSampleCount, mMaxCams;
};

class Tree;

// Nodes.h:66-89 (host mirror of mChildren / mParents / mSamplesPerNodes / mSampleStartIndices)
class Nodes {
public:
    uint32 getChildBlock(uint32 nodeIdx) const { return mChildren[nodeIdx]; }
    Scope getChildScope(const Scope &scope, uint32 childIdx) const {
        return getChildScope(scope, childIdx, 0.5f * scope.getSize());
    }
    Scope getChildScope(const Scope &scope, uint32 childIdx, Real childSize) const {
        const Vector3 &m = scope.getMinimumCoordinates();
        const Vector3 c(m.x + ((childIdx & 1) ? childSize : 0), m.y + ((childIdx & 2) ? childSize : 0),
                        m.z + ((childIdx & 4) ? childSize : 0));
        return Scope(c, childSize, mChildren[scope.getNodeIndex()] + childIdx);
    }
    uint32 getCount() const { return (uint32)mChildren.size(); }
    uint32 getSampleCount(uint32 nodeIdx) const { return mSampleCounts[nodeIdx]; }
    uint32 getSampleStartIndex(uint32 nodeIdx) const { return mSampleStarts[nodeIdx]; }
    uint32 getParent(uint32 nodeIdx) const { return mParents[nodeIdx]; }
    bool isLeaf(uint32 nodeIdx) const { return mChildren[nodeIdx] == INVALID_INDEX; }

private:
    friend class Tree;
    std::vector<uint32> mChildren, mParents, mSampleCounts, mSampleStarts;
};

// Leaves.h:36-67 (DFS-ordered leaf scopes and the node -> leaf map)
class Leaves {
public:
    uint32 getCount() const { return (uint32)mScopes.size(); }
    uint32 getLeafIndex(uint32 nodeIndex) const { return mNodeToLeaf[nodeIndex]; }
    const std::vector<uint32> &getNodeToLeafLinks() const { return mNodeToLeaf; }
    const Scope &getScope(uint32 leafIdx) const { return mScopes[leafIdx]; }
    const std::vector<Scope> &getScopes() const { return mScopes; }
    // Leaves.h:49-55,132-140: the face-neighbour graph (Leaves.cpp:82-180) as CSR
    const uint32 *getNeighbors(uint32 &size, uint32 leafIdx) const {
        const uint32 start = mNeighborsOffsets[leafIdx];
        size = mNeighborsOffsets[leafIdx + 1] - start;
        return mNeighbors.data() + start;
    }
    const uint32 *getNeighborsOffsets() const { return mNeighborsOffsets.data(); }
    const uint32 *getNeighbors() const { return mNeighbors.data(); }

private:
    friend class Tree;
    std::vector<Scope> mScopes;
    std::vector<uint32> mNodeToLeaf;
    std::vector<uint32> mNeighborsOffsets, mNeighbors;
};

// DualCells.h:33-35: 8 leaf indices per interior dual cell, in child-slot order
class DualCells {
public:
    uint32 getCellCount() const { return (uint32)(mIndices.size() / 8); }
    uint32 getIndexCount() const { return (uint32)mIndices.size(); }
    const uint32 *getIndices() const { return mIndices.data(); }

private:
    friend class Tree;
    std::vector<uint32> mIndices;
};

// Tree.h:26-75: Tree::Tree() = octree nodes, leaves and dual cells, built on the GPU;
// Tree(const Path&) loads them from <filesBeginning>.nodes / .leaves / .dualCells
class Tree {
public:
    explicit Tree(const Scene &scene, const tsr_params *params = nullptr) : mScene(&scene) {
        setParams(params);
        check(tsr_tree_build(scene.getContext().get(), scene.get(), &mParams, &mTree));
        mirror();
    }
    // the scene must hold the tree's samples in node order (a samples file of the tree)
    Tree(const Scene &scene, const std::string &filesBeginning, const tsr_params *params = nullptr) : mScene(&scene) {
        setParams(params);
        const std::string n = filesBeginning + ".nodes", l = filesBeginning + ".leaves",
                          c = filesBeginning + ".dualCells";
        check(tsr_tree_load(scene.getContext().get(), scene.get(), &mParams, n.c_str(), l.c_str(), c.c_str(), &mTree));
        mirror();
    }
    ~Tree() { tsr_tree_destroy(mTree); }
    Tree(const Tree &) = delete;
    Tree &operator=(const Tree &) = delete;

    // Tree.cpp:129-134 (the samples in node order go to <filesBeginning>.samples)
    void saveToFiles(const std::string &filesBeginning) const {
        const std::string n = filesBeginning + ".nodes", l = filesBeginning + ".leaves",
                          c = filesBeginning + ".dualCells", s = filesBeginning + ".samples";
        tsr_context *ctx = mScene->getContext().get();
        check(tsr_tree_save(ctx, mTree, n.c_str(), l.c_str(), c.c_str()));
        check(tsr_tree_samples_save(ctx, mTree, s.c_str()));
    }

    // Tree.cpp:50-53 / Nodes.cpp:644-671 with Samples::compact: sampleOffsets[i] = samples
    // erased before sample i (node order, sample count + 1 entries, as Scene.cpp:270-289 forms them)
    void eraseSamples(const uint32 *sampleOffsets) {
        uint32 S = 0;
        for (uint32 c : mNodes.mSampleCounts) S += c;
        std::vector<uint8_t> inliers(S);
        for (uint32 i = 0; i < S; ++i) inliers[i] = sampleOffsets[i + 1] == sampleOffsets[i] ? 1 : 0;
        check(tsr_tree_erase_samples(mScene->getContext().get(), mTree, inliers.data()));
        mirror();
    }

private:
    void setParams(const tsr_params *params) {
        tsr_params p;
        tsr_params_default(&p);
        if (params) p = *params;
        mParams = p;
    }
    void mirror() {
        tsr_context *ctx = mScene->getContext().get();
        const Scene &scene = *mScene;
        tsr_tree_info info;
        check(tsr_tree_get_info(mTree, &info));
        mInfo = info;
        const uint32 N = info.node_count, L = info.leaf_count, C = info.cell_count, S = scene.getSampleCount();
        mNodes.mChildren.resize(N);
        mNodes.mParents.resize(N);
        mNodes.mSampleCounts.resize(N);
        mNodes.mSampleStarts.resize(N);
        mLeaves.mNodeToLeaf.resize(N);
        mLeafNodes.resize(L);
        std::vector<float> scopes((size_t)4 * L);
        mDualCells.mIndices.resize((size_t)8 * C);
        mSampleTarget.resize(S);
        mNodeDepth.resize(N);
        check(tsr_tree_download(ctx, mTree, mNodes.mChildren.data(), mNodes.mParents.data(),
                                mNodes.mSampleCounts.data(), mNodes.mSampleStarts.data(),
                                mLeaves.mNodeToLeaf.data(), mLeafNodes.data(), scopes.data(),
                                mDualCells.mIndices.data(), mSampleTarget.data(), mNodeDepth.data()));
        mSampleDepth.resize(S);
        for (uint32 nd = 0; nd < N; ++nd)
            for (uint32 j = 0; j < mNodes.mSampleCounts[nd]; ++j)
                mSampleDepth[mNodes.mSampleStarts[nd] + j] = (uint8_t)mNodeDepth[nd];
        mLeaves.mScopes.resize(L);
        for (uint32 l = 0; l < L; ++l)
            mLeaves.mScopes[l] = Scope(Vector3(scopes[4 * l], scopes[4 * l + 1], scopes[4 * l + 2]), scopes[4 * l + 3],
                                       mLeafNodes[l]);
        uint64_t edges = 0;   // Leaves.cpp:22-28: the neighbour graph comes with the leaves
        check(tsr_tree_neighbor_count(ctx, mTree, &edges));
        mLeaves.mNeighborsOffsets.resize((size_t)L + 1);
        mLeaves.mNeighbors.resize(edges);
        check(tsr_tree_neighbors_download(ctx, mTree, mLeaves.mNeighborsOffsets.data(), mLeaves.mNeighbors.data()));
    }

public:
    const Nodes &getNodes() const { return mNodes; }
    const Leaves &getLeaves() const { return mLeaves; }
    const DualCells &getDualCells() const { return mDualCells; }
    Real getRootNodeSize() const { return mInfo.root_size; }
    Scope getRootScope() const {
        return Scope(Vector3(mInfo.root_min[0], mInfo.root_min[1], mInfo.root_min[2]), mInfo.root_size, 0);
    }
    Vector3 getCenter() const { return getRootScope().getCenterPosition(); }
    uint32 getNodeDepth(uint32 nodeIdx) const { return mNodeDepth[nodeIdx]; }
    uint32 getMaxDepth() const { return mInfo.max_depth; }
    // Tree.cpp:93-101: level of the sample's container node (original sample index)
    uint32 getSampleDepth(uint32 sampleIdx) const { return mSampleDepth[mSampleTarget[sampleIdx]]; }
    // Leaves.cpp getLeafIndex(rootScope, pos): descend by child slot (x >= c) | (y >= c) << 1 | (z >= c) << 2;
    // INVALID_INDEX outside the root cube
    uint32 getLeafIndex(const Vector3 &queryPosWS) const {
        const Scope root = getRootScope();
        const Vector3 &m = root.getMinimumCoordinates();
        const double size = root.getSize();
        if (queryPosWS.x < m.x || queryPosWS.y < m.y || queryPosWS.z < m.z || queryPosWS.x >= m.x + size ||
            queryPosWS.y >= m.y + size || queryPosWS.z >= m.z + size)
            return INVALID_INDEX;
        uint32 node = 0;
        double lo[3] = {m.x, m.y, m.z}, h = size;
        const double q[3] = {queryPosWS.x, queryPosWS.y, queryPosWS.z};
        while (!mNodes.isLeaf(node)) {
            h *= 0.5;
            uint32 slot = 0;
            for (int a = 0; a < 3; ++a)
                if (q[a] >= lo[a] + h) {
                    slot |= 1u << a;
                    lo[a] += h;
                }
            node = mNodes.getChildBlock(node) + slot;
        }
        return mLeaves.getLeafIndex(node);
    }
    const Scene &getScene() const { return *mScene; }
    const tsr_params &getParams() const { return mParams; }
    tsr_tree *get() const { return mTree; }

private:
    const Scene *mScene;
    tsr_tree *mTree = nullptr;
    tsr_tree_info mInfo;
    tsr_params mParams;
    Nodes mNodes;
    Leaves mLeaves;
    DualCells mDualCells;
    std::vector<uint32> mLeafNodes, mSampleTarget, mNodeDepth;
    std::vector<uint8_t> mSampleDepth;   // per sample in node order
};

// FlexibleMesh subset produced by DMC: vertex positions and scales (FP64), the
// (min leaf, max leaf) key of every vertex's dual edge and the triangle indices
class FlexibleMesh {
public:
    uint32 getVertexCount() const { return (uint32)mScales.size(); }
    uint32 getTriangleCount() const { return (uint32)(mIndices.size() / 3); }
    uint32 getIndexCount() const { return (uint32)mIndices.size(); }
    const double *getPositionsWS() const { return mPositions.data(); }
    const double *getScales() const { return mScales.data(); }
    const uint32 *getIndices() const { return mIndices.data(); }
    const uint32 *getEdgeKeys() const { return mEdgeKeys.data(); }
    const double *getNormals() const { return mNormals.data(); }

private:
    friend class DualMarchingCells;
    std::vector<double> mPositions, mScales, mNormals;
    std::vector<uint32> mIndices, mEdgeKeys;
};

class Occupancy;

// DualMarchingCells.h:34-37 / DualMarchingCells.cpp:26-39: extractMesh, then isle removal
// (vertex components below minTriangleIsleSize), hole filling and vertex normals
class DualMarchingCells {
public:
    DualMarchingCells(const Occupancy &occupancy, uint32 minTriangleIsleSize);
    const FlexibleMesh &getSurface() const { return mSurface; }
    uint32 getMinTriangleIsleSize() const { return mMinIsle; }

private:
    FlexibleMesh mSurface;
    uint32 mMinIsle;
};

// Occupancy.h:49-90: Occupancy::Occupancy(const Tree*) on the GPU
class Occupancy {
public:
    enum DataType { EMPTINESS, SAMPLENESS, DATA_TYPE_COUNT };
    enum NodeStateFlag { NODE_FLAG_SAMPLENESS = 0x1 << 0, NODE_FLAG_EMPTINESS = 0x1 << 1 };

    // Priors for every leaf, as Occupancy.cpp:530-568 computes them; TSR_ACCUM_KAHAN_SEGMENTED
    // alone is the faster form whose priors are NaN where the occupancy ratio does not read them.
    explicit Occupancy(const Tree *tree, int flags = TSR_ACCUM_KAHAN_SEGMENTED | TSR_PRIORS_ALL_LEAVES,
                       const tsr_culler *culler = nullptr)
        : mTree(tree) {
        tsr_context *ctx = tree->getScene().getContext().get();
        check(tsr_occupancy_compute(ctx, tree->get(), tree->getScene().get(), &tree->getParams(), flags, culler,
                                    &mField));
        mirror();
    }
    // Occupancy(const Path&): Occupancy.cpp:663-720
    Occupancy(const Tree *tree, const std::string &fileName) : mTree(tree) {
        check(tsr_field_load(tree->getScene().getContext().get(), tree->get(), &tree->getParams(), fileName.c_str(),
                             &mField));
        mirror();
    }
    // Occupancy.cpp:722-766 (Real = float unless realBytes = 8)
    void saveToFile(const std::string &fileName, int realBytes = 4) const {
        check(tsr_field_save(mTree->getScene().getContext().get(), mField, fileName.c_str(), realBytes));
    }
    // Occupancy.h:57-59,167-175 (Occupancy.cpp:129-136): the field follows added or erased
    // samples (indices into the tree's node-ordered samples); erase before Tree::eraseSamples
    void addSamples(const uint32 *chosenSamples, const uint32 count) {
        check(tsr_occupancy_add_samples(mTree->getScene().getContext().get(), mField, mTree->get(),
                                        mTree->getScene().get(), chosenSamples, count));
        mirror();
    }
    void eraseSamples(const uint32 *doomedSamples, const uint32 count) {
        check(tsr_occupancy_erase_samples(mTree->getScene().getContext().get(), mField, mTree->get(),
                                          mTree->getScene().get(), doomedSamples, count));
        mirror();
    }

private:
    void mirror() {
        tsr_context *ctx = mTree->getScene().getContext().get();
        const uint32 L = mTree->getLeaves().getCount(), N = mTree->getNodes().getCount();
        mKernelSums.resize((size_t)2 * L);
        mConeLengths.resize((size_t)2 * L);
        mPriors.resize(L);
        mNodeStates.resize(N);
        mOccupancy.resize(L);
        check(tsr_field_download(ctx, mField, mKernelSums.data(), mConeLengths.data(), mPriors.data(),
                                 mNodeStates.data(), mOccupancy.data(), &mConeCount));
    }

public:
    ~Occupancy() { tsr_field_destroy(mField); }
    Occupancy(const Occupancy &) = delete;
    Occupancy &operator=(const Occupancy &) = delete;

    // Occupancy.cpp:570-587 (FP64)
    double getOccupancy(uint32 leafIdx) const { return mOccupancy[leafIdx]; }
    // Occupancy.cpp:589-597
    double getEmptinessCPD(uint32 leafIdx) const { return mKernelSums[leafIdx] / mConeCount; }
    double getSamplenessCPD(uint32 leafIdx) const { return mKernelSums[mPriors.size() + leafIdx] / mConeCount; }
    // Occupancy.h:197-210
    double getPriorForEmptiness(uint32 leafIdx) const { return mPriors[leafIdx]; }
    double getPriorForSampleness(uint32 leafIdx) const { return 1.0 - mPriors[leafIdx]; }
    double getObservationPD(uint32 leafIdx) const {
        return getPriorForEmptiness(leafIdx) * getEmptinessCPD(leafIdx) +
               getPriorForSampleness(leafIdx) * getSamplenessCPD(leafIdx);
    }
    Real getBandwidthFactor() const { return mTree->getParams().bandwidth_factor; }
    Real getConfidenceThreshold() const { return mTree->getParams().confidence_threshold; }
    const uint32 *getNodeStates() const { return mNodeStates.data(); }
    float getKernelSum(DataType type, uint32 leafIdx) const { return mKernelSums[type * mPriors.size() + leafIdx]; }
    float getConeLength(DataType type, uint32 leafIdx) const {
        return mConeLengths[type * mPriors.size() + leafIdx];
    }
    double getConeCount() const { return mConeCount; }
    const Tree &getTree() const { return *mTree; }
    tsr_field *get() const { return mField; }

    // Occupancy.cpp:523-528
    const DualMarchingCells &extractCrust(uint32 minTriangleIsleSize = 1000) {
        mCrust.reset(new DualMarchingCells(*this, minTriangleIsleSize));
        return *mCrust;
    }
    const DualMarchingCells *getCrust() const { return mCrust.get(); }

private:
    const Tree *mTree;
    tsr_field *mField = nullptr;
    std::vector<float> mKernelSums, mConeLengths;
    std::vector<double> mPriors;
    std::vector<uint32> mNodeStates;
    std::vector<double> mOccupancy;
    double mConeCount = 0;
    std::unique_ptr<DualMarchingCells> mCrust;
};

inline DualMarchingCells::DualMarchingCells(const Occupancy &occupancy, uint32 minTriangleIsleSize)
    : mMinIsle(minTriangleIsleSize) {
    const Tree &tree = occupancy.getTree();
    tsr_context *ctx = tree.getScene().getContext().get();
    tsr_mesh *mesh = nullptr;
    check(tsr_dmc_extract(ctx, tree.get(), occupancy.get(), &mesh));
    std::unique_ptr<tsr_mesh, void (*)(tsr_mesh *)> owner(mesh, tsr_mesh_destroy);
    check(tsr_crust_postprocess(ctx, mesh, minTriangleIsleSize));
    tsr_mesh_info info;
    check(tsr_mesh_get_info(mesh, &info));
    mSurface.mPositions.resize((size_t)3 * info.vertex_count);
    mSurface.mScales.resize(info.vertex_count);
    mSurface.mEdgeKeys.resize((size_t)2 * info.vertex_count);
    mSurface.mIndices.resize((size_t)3 * info.triangle_count);
    check(tsr_mesh_download(ctx, mesh, mSurface.mPositions.data(), mSurface.mScales.data(), mSurface.mEdgeKeys.data(),
                            mSurface.mIndices.data()));
    mSurface.mNormals.resize((size_t)3 * info.vertex_count);
    check(tsr_mesh_download_normals(ctx, mesh, mSurface.mNormals.data()));
}

}  // namespace tsrgpu

#endif
