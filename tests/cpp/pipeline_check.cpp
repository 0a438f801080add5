// Scene::reconstruct (SR/Scene/Scene.cpp:169-197) through the C++ class mirror:
// Tree -> Occupancy(tree) -> extractCrust().getSurface(), on a seeded synthetic
// scene.  Checks the accessors against each other and prints one JSON line of
// counts and checksums that the Python binding of the same run must reproduce.
//   usage: pipeline_check <config> <resolution_scale> [seed] [files beginning: save + reload]
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "tsr_gpu.hpp"

using namespace tsrgpu;

static int failures = 0;
#define EXPECT(cond)                                                                  \
    do {                                                                              \
        if (!(cond)) {                                                                \
            std::fprintf(stderr, "pipeline_check: %s failed (line %d)\n", #cond, __LINE__); \
            ++failures;                                                               \
        }                                                                             \
    } while (0)

static uint64_t fnv(const void *p, size_t n, uint64_t h = 1469598103934665603ull) {
    const unsigned char *b = (const unsigned char *)p;
    for (size_t i = 0; i < n; ++i) h = (h ^ b[i]) * 1099511628211ull;
    return h;
}

int main(int argc, char **argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s <config> <resolution_scale> [seed]\n", argv[0]);
        return 2;
    }
    const int config = std::atoi(argv[1]);
    const double res = std::atof(argv[2]);
    const uint64_t seed = argc > 3 ? std::strtoull(argv[3], nullptr, 10) : 0;
    try {
        tsr_synthetic *syn = nullptr;
        check(tsr_synthetic_create(config, seed, res, &syn));
        tsr_samples_view sv;
        tsr_cameras_view cv;
        check(tsr_synthetic_views(syn, &sv, &cv));

        Context ctx(0);
        Scene scene(ctx, sv, cv);
        Tree tree(scene);
        const Nodes &nodes = tree.getNodes();
        const Leaves &leaves = tree.getLeaves();
        const DualCells &cells = tree.getDualCells();

        // Nodes / Leaves / Scope consistency: descending by getChildScope from the
        // root reaches every leaf with the scope Leaves::getScope reports
        uint32 visited = 0;
        std::vector<Scope> stack{tree.getRootScope()};
        while (!stack.empty()) {
            const Scope s = stack.back();
            stack.pop_back();
            const uint32 nd = s.getNodeIndex();
            if (nodes.isLeaf(nd)) {
                const Scope &ls = leaves.getScope(leaves.getLeafIndex(nd));
                EXPECT(ls.getNodeIndex() == nd);
                EXPECT(std::fabs(ls.getSize() - s.getSize()) <= 1e-6f * s.getSize());
                EXPECT(std::fabs(ls.getMinimumCoordinates().x - s.getMinimumCoordinates().x) <= 1e-5f * tree.getRootNodeSize());
                ++visited;
                continue;
            }
            for (uint32 c = 0; c < 8; ++c) stack.push_back(nodes.getChildScope(s, c));
        }
        EXPECT(visited == leaves.getCount());
        // Tree::getLeafIndex at every leaf centre finds that leaf
        for (uint32 l = 0; l < leaves.getCount(); ++l) EXPECT(tree.getLeafIndex(leaves.getScope(l).getCenterPosition()) == l);
        for (uint32 i = 0; i < cells.getIndexCount(); ++i) EXPECT(cells.getIndices()[i] < leaves.getCount());
        // sample depth = depth of the node holding the sample
        uint64_t depthSum = 0;
        for (uint32 i = 0; i < sv.count; ++i) depthSum += tree.getSampleDepth(i);

        Occupancy occupancy(&tree);
        uint32 positive = 0;
        double posSum = 0;
        for (uint32 l = 0; l < leaves.getCount(); ++l) {
            const double o = occupancy.getOccupancy(l);
            // Occupancy.cpp:570-587 restated on the accessors
            const double e = occupancy.getKernelSum(Occupancy::EMPTINESS, l), s = occupancy.getKernelSum(Occupancy::SAMPLENESS, l);
            double expect = -(double)1e-6f;
            if (!(e + s < occupancy.getConfidenceThreshold())) {
                const double pe = occupancy.getPriorForEmptiness(l) * occupancy.getEmptinessCPD(l);
                const double ps = occupancy.getPriorForSampleness(l) * occupancy.getSamplenessCPD(l);
                if (pe < ps) expect = ps - pe;
            }
            EXPECT(o == expect);
            if (o > 0) {
                ++positive;
                posSum += o;
            }
        }
        // all-leaf priors (Occupancy.cpp:530-568): every leaf whose occupancy reads a prior has one
        uint32 finitePriors = 0;
        for (uint32 l = 0; l < leaves.getCount(); ++l) {
            const bool finite = std::isfinite(occupancy.getPriorForEmptiness(l));
            finitePriors += finite;
            if (occupancy.getKernelSum(Occupancy::EMPTINESS, l) + occupancy.getKernelSum(Occupancy::SAMPLENESS, l) >=
                occupancy.getConfidenceThreshold())
                EXPECT(finite && std::isfinite(occupancy.getObservationPD(l)));
        }
        // Leaves::getNeighbors: symmetric lists
        uint64_t neighbourSum = 0;
        for (uint32 l = 0; l < leaves.getCount(); ++l) {
            uint32 n = 0;
            const uint32 *nb = leaves.getNeighbors(n, l);
            neighbourSum += n;
            for (uint32 j = 0; j < n; ++j) {
                uint32 m = 0;
                const uint32 *back = leaves.getNeighbors(m, nb[j]);
                bool found = false;
                for (uint32 k = 0; k < m; ++k) found |= back[k] == l;
                EXPECT(found);
            }
        }
        // Tree::saveToFiles -> samples file -> Tree(const Path&); Occupancy::saveToFile -> Occupancy(const Path&)
        if (argc > 4) {
            const std::string base = argv[4];
            tree.saveToFiles(base);
            occupancy.saveToFile(base + ".occupancy", 8);
            const std::string sp = base + ".samples";
            uint32 n = 0, mc = 0;
            check(tsr_samples_file_info(sp.c_str(), &n, &mc));
            std::vector<float> a[8];
            for (auto &v : a) v.resize(n);
            std::vector<uint32> cams((size_t)n * mc);
            check(tsr_samples_file_read(sp.c_str(), a[0].data(), a[1].data(), a[2].data(), a[3].data(), a[4].data(),
                                        a[5].data(), a[6].data(), a[7].data(), cams.data()));
            tsr_samples_view lv{n, mc, a[0].data(), a[1].data(), a[2].data(), a[3].data(), a[4].data(), a[5].data(),
                                a[6].data(), a[7].data(), cams.data()};
            Scene loadedScene(ctx, lv, cv);
            Tree loaded(loadedScene, base);
            EXPECT(loaded.getNodes().getCount() == nodes.getCount());
            EXPECT(loaded.getDualCells().getIndexCount() == cells.getIndexCount());
            for (uint32 i = 0; i < cells.getIndexCount(); ++i)
                EXPECT(loaded.getDualCells().getIndices()[i] == cells.getIndices()[i]);
            for (uint32 nd = 0; nd < nodes.getCount(); ++nd)
                EXPECT(loaded.getNodes().getChildBlock(nd) == nodes.getChildBlock(nd));
            Occupancy reloaded(&loaded, base + ".occupancy");
            for (uint32 l = 0; l < leaves.getCount(); ++l)
                EXPECT(reloaded.getOccupancy(l) == occupancy.getOccupancy(l));
        }
        const DualMarchingCells &crust = occupancy.extractCrust(1000);
        EXPECT(occupancy.getCrust() == &crust);
        const FlexibleMesh &mesh = crust.getSurface();
        for (uint32 i = 0; i < mesh.getIndexCount(); ++i) EXPECT(mesh.getIndices()[i] < mesh.getVertexCount());
        for (uint32 v = 0; v < mesh.getVertexCount(); ++v) {   // unit normals (or zero on collapsed triangles)
            const double *n = mesh.getNormals() + 3 * (size_t)v;
            const double l = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
            EXPECT(l == 0.0 || std::fabs(l - 1.0) < 1e-9);
        }

        std::printf("{\"samples\": %u, \"nodes\": %u, \"leaves\": %u, \"cells\": %u, \"max_depth\": %u, "
                    "\"sample_depth_sum\": %llu, \"positive_leaves\": %u, \"positive_sum\": %.17g, "
                    "\"cone_count\": %.17g, \"vertices\": %u, \"triangles\": %u, \"cells_hash\": %llu, "
                    "\"triangles_hash\": %llu, \"positions_hash\": %llu, \"finite_priors\": %u, "
                    "\"neighbour_entries\": %llu, \"failures\": %d}\n",
                    sv.count, nodes.getCount(), leaves.getCount(), cells.getCellCount(), tree.getMaxDepth(),
                    (unsigned long long)depthSum, positive, posSum, occupancy.getConeCount(), mesh.getVertexCount(),
                    mesh.getTriangleCount(),
                    (unsigned long long)fnv(cells.getIndices(), (size_t)cells.getIndexCount() * 4),
                    (unsigned long long)fnv(mesh.getIndices(), (size_t)mesh.getIndexCount() * 4),
                    (unsigned long long)fnv(mesh.getPositionsWS(), (size_t)mesh.getVertexCount() * 24), finitePriors,
                    (unsigned long long)neighbourSum, failures);
        tsr_synthetic_destroy(syn);
    } catch (const Exception &e) {
        std::fprintf(stderr, "pipeline_check: %s (status %d)\n", e.what(), (int)e.getStatus());
        return 1;
    }
    return failures ? 1 : 0;
}
