/*
 * tsr_gpu.h - C ABI of the MI355X (gfx950) occupancy-field + dual-marching-cells
 * library.  Plain C types only; every handle owns device memory on one HIP
 * stream.  The entry points replace the C++ class interfaces that TSR's
 * Scene::reconstruct calls on its hot path (paths relative to
 * SurfaceReconstruction/SurfaceReconstruction/):
 *
 *   tsr_tree_build          <- Tree::Tree()                 Scene/Tree/Tree.cpp:26-37
 *                              (Nodes ctor Nodes.cpp:186-194, Leaves ctor Leaves.cpp:22-28,
 *                               DualCells ctor DualCells.cpp:21-31, Samples::reorder Samples.cpp:825-835)
 *   tsr_occupancy_compute   <- Occupancy::Occupancy(const Tree*)  SurfaceExtraction/Occupancy.cpp:78-127
 *   tsr_dmc_extract         <- DualMarchingCells::extractMesh     SurfaceExtraction/DualMarchingCells.cpp:41-111
 *   tsr_*_download          <- the accessors of Nodes/Leaves/DualCells/Occupancy/FlexibleMesh
 *                              (Nodes.h:66-89, Leaves.h:36-67, DualCells.h:33-35, Occupancy.h:49-90)
 *
 * Errors: every call returns a tsr_status; tsr_last_error() gives the message
 * (thread-local).  The status codes map the reference's exceptions (see the
 * enum comments).  Calls on one context are serialised on its stream; handles
 * created on one context must not be used with another.
 */
#ifndef TSR_GPU_H
#define TSR_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum tsr_status {
    TSR_OK = 0,
    TSR_ERR_INVALID_ARGUMENT = 1,
    TSR_ERR_NO_SAMPLES = 2,            /* Scene.cpp:137-151: reconstruct() returns false          */
    TSR_ERR_SAMPLE_OUTSIDE_ROOT = 3,   /* Nodes.cpp:237-238 findContainer exception                */
    TSR_ERR_NODE_OVERFLOW = 4,         /* Nodes.cpp:399-407 more than 2^32-2 nodes                  */
    TSR_ERR_TREE_TOO_DEEP = 5,         /* octree deeper than 19 levels                              */
    TSR_ERR_LINK_OVERFLOW = 6,         /* Samples.h:395-402 more than 2^32-1 camera links          */
    TSR_ERR_NON_MANIFOLD_EDGE = 7,     /* FlexibleMesh.cpp:1227-1260 addEdgeConflict "BOOM"        */
    TSR_ERR_HIP = 8,                   /* HIP runtime failure                                      */
    TSR_ERR_OUT_OF_MEMORY = 9,
    TSR_ERR_ZERO_SIZE_NODE = 10,       /* Nodes.cpp:303-304 node of size 0                          */
    TSR_ERR_SAMPLE_NOT_INSERTED = 11,  /* Nodes.cpp:615-616                                          */
    TSR_ERR_INVALID_SCALE = 12,        /* Samples.cpp:280-288 non-positive sample scale             */
    TSR_ERR_FILE = 13                  /* FileCorruptionException: missing, short or invalid file   */
} tsr_status;

/* Parameters of the path, names as in App.cfg:63-73 (defaults = App.cfg values). */
typedef struct tsr_params {
    float bandwidth_factor;              /* Occupancy::bandwidthFactor = 2            */
    float confidence_threshold;          /* Occupancy::confidenceThreshold = 100      */
    float relative_radius_for_priors;    /* Occupancy::relativeRadiusForPriors = 25   */
    float relative_sample_cone_length;   /* Occupancy::relativeSampleConeLength = 1   */
    float max_relative_sampling_distance;/* Samples::maxRelativeSamplingDistance = 2  */
} tsr_params;

/* Occupancy accumulation modes (tsr_occupancy_compute flags). */
enum {
    TSR_ACCUM_KAHAN_SEGMENTED = 0,   /* FP32 compensated, deterministic segmented reduction (default) */
    TSR_ACCUM_FP32_ATOMIC = 1,       /* plain FP32 atomics (precision study only, nondeterministic)  */
    TSR_CULL_LEARNED = 0x100,        /* certified priors, first stage per leaf from the routing MLP */
    TSR_PRIORS_ALL_LEAVES = 0x200,   /* priors for every leaf; by default only for the leaves whose
                                        occupancy reads them (E + S >= confidence threshold, the
                                        others get NaN)                                             */
    TSR_PRIORS_CERTIFIED = 0x400,    /* certified priors, every leaf from the first stage (below)   */
    TSR_EMPTINESS_TRAVERSAL = 0x800  /* E pass as one cone traversal per link (the reference's loop
                                        order); by default the same contributions are gathered per
                                        sampleness-flagged leaf from a per-camera direction index  */
};
/* Certified priors (no reference counterpart).  The mesh reads of a leaf's occupancy only
 * its class: -EPSILON (pE >= pS), in (0, EPSILON], or above EPSILON (DualMarchingCell.cpp:
 * 51-71 places a crossing vertex at the corner whose |value| <= EPSILON is tested first).
 * Stage 0 bounds each queried leaf's prior by a sphere traversal that does not open nodes
 * 3 levels above the leaf's depth (their subtree sums bound what the exact traversal would
 * add), stage 1 with the cap 2 levels above; the occupancy formula on both ends of the
 * interval (the judge) keeps a leaf whose class the interval decides, with the interval
 * midpoint as its prior.  Stage 2 computes the exact prior (the plain evaluation, the
 * fallback) for the rest.  The DMC mesh equals that of the plain field; priors and positive
 * occupancies of certified leaves are within their certified intervals, not exact. */

typedef struct tsr_context tsr_context;
typedef struct tsr_scene tsr_scene;
typedef struct tsr_tree tsr_tree;
typedef struct tsr_field tsr_field;
typedef struct tsr_mesh tsr_mesh;
typedef struct tsr_synthetic tsr_synthetic;
typedef struct tsr_culler tsr_culler;

/* Float32 SoA view of the samples (Samples.h:300-312) in host memory. */
typedef struct tsr_samples_view {
    uint32_t count;
    uint32_t max_cameras_per_sample;   /* Samples::getMaxCamerasPerSample            */
    const float *px, *py, *pz;         /* positions                                   */
    const float *nx, *ny, *nz;         /* unit normals                                */
    const float *scale;                /* sample scale (> 0)                          */
    const float *confidence;
    const uint32_t *parent_cameras;    /* count * max_cameras_per_sample; >= camera count = invalid */
} tsr_samples_view;

/* Camera centres (Cameras::getPositionWS, Cameras.h:161-176). */
typedef struct tsr_cameras_view {
    uint32_t count;
    const float *x, *y, *z;
} tsr_cameras_view;

typedef struct tsr_tree_info {
    uint32_t node_count;     /* Nodes::getCount          */
    uint32_t leaf_count;     /* Leaves::getCount         */
    uint32_t cell_count;     /* DualCells::getCellCount  */
    uint32_t max_depth;
    float root_min[3];       /* Tree::getRootScope       */
    float root_size;         /* Tree::getRootNodeSize    */
} tsr_tree_info;

typedef struct tsr_mesh_info {
    uint32_t vertex_count;
    uint32_t triangle_count;
} tsr_mesh_info;

void tsr_params_default(tsr_params *out);
const char *tsr_last_error(void);

/* Context: one device, one HIP stream (default priority). */
tsr_status tsr_context_create(int device, tsr_context **out);
void tsr_context_destroy(tsr_context *ctx);
tsr_status tsr_context_synchronize(tsr_context *ctx);
/* Milliseconds of the last tree build / occupancy / DMC phases (HIP events):
 * [0] tree, [1] S pass, [2] E pass, [3] priors, [4] occupancy, [5] DMC.  n <= 6. */
tsr_status tsr_context_phase_times(tsr_context *ctx, float *ms, int n);
/* Accumulated live timing (HIP events on the context stream) of instrumented
 * kernels since the last reset, with their algorithmic bytes (each input or
 * output element counted once per launch) and work units (queries or links):
 * kernel 0 = priors sphere query, 1/2 = emptiness cone pass (one traversal per link,
 * staging its records) / re-traversal of the links with more records than staging
 * slots, 3/4 = the same for the sampleness cones, 5 = emptiness pass gathered per
 * sampleness-flagged leaf (the default E pass; units = leaves). */
tsr_status tsr_context_kernel_stats(tsr_context *ctx, int kernel, double *ms, uint64_t *launches, uint64_t *bytes,
                                    uint64_t *units);
tsr_status tsr_context_reset_kernel_stats(tsr_context *ctx);
/* Plain device copy kernel of `bytes` bytes (float4 loads and stores) on the
 * context stream: the reference point for scaling HBM traffic counters. */
tsr_status tsr_calibration_copy(tsr_context *ctx, uint64_t bytes);
/* Device memory held by the library in this process (all contexts and handles,
 * cached blocks included) and its high-water mark; reset_peak != 0 restarts the
 * mark at the current value after reading it.  No reference counterpart: it
 * backs the footprint figures of large scenes. */
tsr_status tsr_device_memory(uint64_t *current_bytes, uint64_t *peak_bytes, int reset_peak);

/* Scene = device copy of samples + cameras (the Scene singleton's data). */
tsr_status tsr_scene_upload(tsr_context *ctx, const tsr_samples_view *samples, const tsr_cameras_view *cameras,
                            tsr_scene **out);
void tsr_scene_destroy(tsr_scene *scene);
/* Download the scene's samples (input order). Arrays may be NULL. */
tsr_status tsr_scene_download(tsr_context *ctx, const tsr_scene *scene, float *px, float *py, float *pz, float *nx,
                              float *ny, float *nz, float *scale, float *confidence, uint32_t *parent_cameras);

/* Tree::Tree(): octree nodes, leaves and dual cells built on the GPU.  The tree
 * keeps its own copy of the samples permuted into node order (Nodes.cpp:594-636,
 * Samples::reorder); the scene's arrays stay in input order, so building again
 * from the same scene gives the same tree. */
tsr_status tsr_tree_build(tsr_context *ctx, const tsr_scene *scene, const tsr_params *params, tsr_tree **out);
tsr_status tsr_tree_get_info(const tsr_tree *tree, tsr_tree_info *info);
/* Statistics of the sampling refinement (Nodes.cpp:433-517; no reference counterpart):
 * [0] sampling units (2 per sample), [1] units that refined, [2] nodes they created,
 * [3] visited-leaf lookups answered on the device (leaves outside the transferred part of
 * the visit index; TSR_HOT_RADIUS, in max sampling distances, sizes that part). */
tsr_status tsr_tree_build_stats(const tsr_tree *tree, uint64_t *stats4);
/* Host copies of the tables (NULL pointers are skipped):
 *   children/parents/sample_counts/sample_starts: node_count (Nodes.h:143-147)
 *   node_to_leaf: node_count (Leaves.h:90); leaf_nodes: leaf_count
 *   leaf_scopes: 4 floats per leaf = min x,y,z, size (Scope, Scope.h:36-38)
 *   cells: 8 * cell_count leaf indices (DualCells.h:67)
 *   sample_target: new index of each original sample (Nodes.cpp:625-626), one per sample the
 *   tree was built from (0xFFFFFFFF for samples erased since)
 *   node_depth: node_count */
tsr_status tsr_tree_download(tsr_context *ctx, const tsr_tree *tree, uint32_t *children, uint32_t *parents,
                             uint32_t *sample_counts, uint32_t *sample_starts, uint32_t *node_to_leaf,
                             uint32_t *leaf_nodes, float *leaf_scopes, uint32_t *cells, uint32_t *sample_target,
                             uint32_t *node_depth);
/* The tree's samples in node order (the reference's reordered Samples). */
tsr_status tsr_tree_samples_download(tsr_context *ctx, const tsr_tree *tree, float *px, float *py, float *pz,
                                     float *nx, float *ny, float *nz, float *scale, float *confidence,
                                     uint32_t *parent_cameras);
void tsr_tree_destroy(tsr_tree *tree);

/* Leaves::createNeighborsEdges / getNeighbors (Leaves.cpp:82-180, Leaves.h:36-67): the
 * face-neighbour graph of the leaves as CSR -- offsets: leaf_count + 1, neighbors: total;
 * leaf l's neighbours are neighbors[offsets[l] .. offsets[l + 1]) in the reference's order.
 * Built on the GPU on first request (or read from a leaves file) and kept with the tree. */
tsr_status tsr_tree_neighbor_count(tsr_context *ctx, tsr_tree *tree, uint64_t *total);
tsr_status tsr_tree_neighbors_download(tsr_context *ctx, tsr_tree *tree, uint32_t *offsets, uint32_t *neighbors);

/* On-disk formats (Tree::saveToFiles / Tree(const Path&), Nodes.cpp:770-807,
 * Leaves.cpp:230-292, DualCells.cpp:359-385; Occupancy.cpp:663-766; Samples.cpp:693-765):
 * the reference's payloads in its order behind a 16-byte header ("TSRF", kind, version 0,
 * bytes per Real).  NULL paths are skipped.  Loading fails with TSR_ERR_FILE on missing,
 * short, inconsistent or foreign files.
 * tsr_tree_load: the scene must hold the tree's samples in node order (as samples files
 * store them, tsr_tree_samples_save); nodes must be in depth-first child-block order.
 * tsr_field_save: real_bytes 4 (the reference's Real = float) or 8 (lossless for this
 * library's FP64 priors). */
tsr_status tsr_tree_save(tsr_context *ctx, tsr_tree *tree, const char *nodes_path, const char *leaves_path,
                         const char *cells_path);
/* Cameras::saveToFile / loadFromFile (Cameras.cpp:116-177; header kind 6): the camera count
 * and one CameraData record (CameraData.h:21-36) per camera, as the struct lays it out:
 * 13 Reals -- orientation quaternion (BaseProject's Quaternion member order is not in the
 * reference; stored as given), position, principal point, 2 distortion coefficients, focal
 * length, image aspect ratio -- then the view ID (padded to 8 bytes when Real is double).
 * Host only: data13 holds 13 doubles per camera. */
tsr_status tsr_cameras_file_write(const char *path, uint32_t count, const double *data13, const uint32_t *view_ids,
                                  int real_bytes);
tsr_status tsr_cameras_file_info(const char *path, uint32_t *count);
tsr_status tsr_cameras_file_read(const char *path, double *data13, uint32_t *view_ids);
tsr_status tsr_tree_load(tsr_context *ctx, const tsr_scene *scene, const tsr_params *params, const char *nodes_path,
                         const char *leaves_path, const char *cells_path, tsr_tree **out);
tsr_status tsr_tree_samples_save(tsr_context *ctx, const tsr_tree *tree, const char *path);
tsr_status tsr_samples_file_info(const char *path, uint32_t *count, uint32_t *max_cameras_per_sample);
/* host arrays of count (parent_cameras: count * max_cameras_per_sample); NULL = skip */
tsr_status tsr_samples_file_read(const char *path, float *px, float *py, float *pz, float *nx, float *ny, float *nz,
                                 float *scale, float *confidence, uint32_t *parent_cameras);

/* Occupancy::Occupancy(tree): S pass, flag propagation, E pass, priors and the
 * FP64 occupancy ratio.  flags: TSR_ACCUM_* | optional TSR_CULL_LEARNED (then
 * `culler` must be non-NULL). */
tsr_status tsr_occupancy_compute(tsr_context *ctx, const tsr_tree *tree, const tsr_scene *scene,
                                 const tsr_params *params, int flags, const tsr_culler *culler, tsr_field **out);
/* kernel_sums / cone_lengths: 2 * leaf_count floats, [0..L) = EMPTINESS, [L..2L) = SAMPLENESS
 * (Occupancy.h:137-138); priors: leaf_count doubles (mPriorsForEmptiness, kept in FP64 because
 * the occupancy ratio uses 1 - prior, which float rounding of a prior near 1 would spoil);
 * node_states: node_count;
 * occupancy: leaf_count doubles (getOccupancy, Occupancy.cpp:570-587); cone_count: M. */
tsr_status tsr_field_download(tsr_context *ctx, const tsr_field *field, float *kernel_sums, float *cone_lengths,
                              double *priors, uint32_t *node_states, double *occupancy, double *cone_count);
/* counts of (link, leaf) kernel contributions of the last passes: [0] = E, [1] = S */
tsr_status tsr_field_hit_counts(const tsr_field *field, uint64_t *hits2);
/* Occupancy::addSamples / eraseSamples (Occupancy.h:57-59,167-175; Occupancy.cpp:129-136,
 * 138-248): the field follows a change of samples.  The S pass, the S flags up the tree and
 * the gated E pass run over the camera links of the chosen samples only, and their per-leaf
 * kernel sums, cone lengths and cone count M are added to (erase: subtracted from) the
 * field's; node flags are only ever set, as in the reference.  Priors (in the field's priors
 * mode) and the occupancy ratio are then recomputed.  samples: indices into the tree's
 * samples (node order, as Samples after Nodes::reorderSamples).  Scene::eraseSamples
 * (Scene.cpp:270-296) erases from the field first, then from the tree. */
tsr_status tsr_occupancy_add_samples(tsr_context *ctx, tsr_field *field, const tsr_tree *tree, const tsr_scene *scene,
                                     const uint32_t *samples, uint32_t count);
tsr_status tsr_occupancy_erase_samples(tsr_context *ctx, tsr_field *field, const tsr_tree *tree,
                                       const tsr_scene *scene, const uint32_t *samples, uint32_t count);
/* Tree::eraseSamples (Tree.cpp:50-53, Nodes.cpp:644-671) with Samples::compact: drops the
 * tree's samples whose inliers[i] == 0 (node order, one byte per sample) and moves every
 * node's sample count and start accordingly; the input-to-node sample map maps erased samples
 * to 0xFFFFFFFF.  The scene (input-order samples, cameras) is not changed. */
tsr_status tsr_tree_erase_samples(tsr_context *ctx, tsr_tree *tree, const uint8_t *inliers);
/* Priors stages (6 counts): queried leaves; leaves entering stage 0, 1, 2 (2 = exact);
 * leaves the routing MLP sent straight to stage 1, to stage 2.  A plain field has every
 * query at stage 2. */
tsr_status tsr_field_prior_stages(const tsr_field *field, uint32_t *stages6);
/* Per leaf (leaf_count bytes): the stage whose evaluation decided the leaf (0, 1: certified
 * by the capped bound 3 or 2 levels up; 2: exact prior), 255 if its prior was not queried.
 * On a TSR_PRIORS_CERTIFIED field these are the routing MLP's training targets. */
tsr_status tsr_field_prior_levels(tsr_context *ctx, const tsr_field *field, uint8_t *levels);
void tsr_field_destroy(tsr_field *field);
/* Occupancy::saveToFile / Occupancy(const Path&) for a tree (see the on-disk formats above);
 * loading recomputes the occupancy ratio from the loaded sums and priors. */
tsr_status tsr_field_save(tsr_context *ctx, const tsr_field *field, const char *path, int real_bytes);
tsr_status tsr_field_load(tsr_context *ctx, const tsr_tree *tree, const tsr_params *params, const char *path,
                          tsr_field **out);

/* DualMarchingCells::extractMesh on the GPU (raw crust mesh, before isle
 * removal / hole filling).  Fails with TSR_ERR_NON_MANIFOLD_EDGE like
 * FlexibleMesh::addEdgeConflict. */
tsr_status tsr_dmc_extract(tsr_context *ctx, const tsr_tree *tree, const tsr_field *field, tsr_mesh **out);
/* DualMarchingCells over an explicit per-leaf occupancy (leaf_count doubles in host
 * memory) instead of a computed field: the crust of any scalar field on the tree's dual
 * cells (DualMarchingCells.cpp:41-111 with Occupancy::getOccupancy replaced by the given
 * values).  Fails with TSR_ERR_NON_MANIFOLD_EDGE like tsr_dmc_extract. */
tsr_status tsr_dmc_extract_values(tsr_context *ctx, const tsr_tree *tree, const double *occupancy, tsr_mesh **out);
/* tsr_dmc_extract into caller-provided DEVICE buffers, allocating no device memory.  No
 * counterpart in the reference: it is for callers that cannot allocate (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 *   positions: 3 * max_vertices doubles, scales: max_vertices doubles, edge_keys:
 *   2 * max_vertices, triangles: 3 * max_triangles (same layouts as tsr_mesh_download).
 *   vertex_count / triangle_count (host) receive the mesh size, also when the call fails
 *   because a capacity is too small (TSR_ERR_INVALID_ARGUMENT; the buffers are then
 *   partly written).
 * The temporaries live in one fixed device scratch array of the library of 256 MiB per
 * GPU (held until the process ends), enough for meshes of about two million triangles; a
 * mesh that needs more fails with TSR_ERR_INVALID_ARGUMENT.  Every call on a GPU uses that
 * GPU's array, so calls that overlap on one GPU (several contexts or threads) race. */
tsr_status tsr_dmc_extract_noalloc(tsr_context *ctx, const tsr_tree *tree, const tsr_field *field,
                                   uint32_t max_vertices, uint32_t max_triangles, double *positions, double *scales,
                                   uint32_t *edge_keys, uint32_t *triangles, uint32_t *vertex_count,
                                   uint32_t *triangle_count);
/* The rest of DualMarchingCells' constructor (DualMarchingCells.cpp:26-39) on a mesh in
 * place: isle removal (FlexibleMesh::deleteIsolatedGeometry, FlexibleMesh.cpp:1411-1426,
 * IslesEraser.cpp: vertex components below min_triangle_isle_size vertices go), hole filling
 * (FlexibleMesh::fillHoles, FlexibleMesh.cpp:1580-1932; new vertices have scale 0 and edge
 * keys 0xffffffff) and angle-weighted vertex normals (Mesh.cpp:145-202).  An empty crust
 * after isle removal is left without holes filled or normals, as in the reference. */
tsr_status tsr_crust_postprocess(tsr_context *ctx, tsr_mesh *mesh, uint32_t min_triangle_isle_size);
/* normals: 3 doubles per vertex (zeros before tsr_crust_postprocess) */
tsr_status tsr_mesh_download_normals(tsr_context *ctx, const tsr_mesh *mesh, double *normals);
tsr_status tsr_mesh_get_info(const tsr_mesh *mesh, tsr_mesh_info *info);
/* positions: 3 doubles per vertex; scales: 1 double per vertex; edge_keys: 2
 * leaf indices per vertex (EdgeVertexIndex); triangles: 3 vertex ids each. */
tsr_status tsr_mesh_download(tsr_context *ctx, const tsr_mesh *mesh, double *positions, double *scales,
                             uint32_t *edge_keys, uint32_t *triangles);
void tsr_mesh_destroy(tsr_mesh *mesh);

/* Learned routing of the certified priors (no reference counterpart): an MLP 16-64-64-64-1
 * (ReLU, sigmoid output; BF16 MFMA with FP32 accumulation) predicts from per-leaf features
 * the stage of the certified cascade that decides a queried leaf (trained on stage / 2).
 * With TSR_CULL_LEARNED each queried leaf starts at stage floor(2 p + 1 - threshold)
 * (clamped to 0..2) instead of 0, skipping bound evaluations that would not decide it.
 * The network never changes a result: every leaf is still decided by the judge or the
 * exact fallback.  Weights are a flat FP32 array W1[64x16], b1[64], W2[64x64], b2[64],
 * W3[64x64], b3[64], W4[1x64], b4[1] (W row-major [out][in]). */
#define TSR_MLP_PARAM_COUNT (64 * 16 + 64 + 64 * 64 + 64 + 64 * 64 + 64 + 64 + 1)
#define TSR_LEAF_FEATURES 16
tsr_status tsr_culler_create(tsr_context *ctx, const float *weights, float threshold, tsr_culler **out);
tsr_status tsr_culler_set_weights(tsr_context *ctx, tsr_culler *culler, const float *weights, float threshold);
void tsr_culler_destroy(tsr_culler *culler);
/* TSR_LEAF_FEATURES floats per leaf (row-major, leaf_count rows) from a tree and its field
 * (kernel sums, cone lengths and the subtree sums the priors traverse). */
tsr_status tsr_leaf_features(tsr_context *ctx, const tsr_tree *tree, const tsr_field *field, float *features_host);
/* MLP forward on host features (n rows): sigmoid outputs. */
tsr_status tsr_mlp_forward(tsr_context *ctx, const float *weights, const float *features, uint32_t n, float *probs);
/* Loss (weighted BCE, summed) and its gradient w.r.t. the weights for one batch,
 * computed by the MFMA forward/backward kernels: the training building block. */
tsr_status tsr_mlp_gradient(tsr_context *ctx, const float *weights, const float *features, const float *labels,
                            const float *row_weights, uint32_t n, float *grad, double *loss);
/* Adam training on the device: epochs over n rows (features, labels, per-row
 * weights in host memory), mini-batches of `batch` rows, shuffled with `seed`.
 * weights is read as the initial state and overwritten with the result;
 * loss_per_epoch receives the mean weighted loss of every epoch. */
tsr_status tsr_mlp_train(tsr_context *ctx, float *weights, const float *features, const float *labels,
                         const float *row_weights, uint32_t n, uint32_t epochs, uint32_t batch, float learning_rate,
                         uint64_t seed, float *loss_per_epoch);

/* Sample preparation from depth maps (DepthImage.cpp:191-413 triangulate / triangulateBlock /
 * createFlexibleMesh, Mesh.cpp:63-124 vertex scales, Scene.cpp:484-573 loadDepthMeshes /
 * loadCameraIndices, Samples.cpp:224-273 addSamplesViaMeshes): every 2x2 pixel block with 3 or 4
 * depths > 0 gets one triangle or the two along the diagonal of smaller depth difference, minus
 * triangles with an edge across a depth discontinuity (depth difference above 5 pixel footprints
 * of the closer corner, sqrt(2) times that on the diagonal); vertex ids follow the pixels' first
 * use in block order.  Vertices get world positions (camera centre + unit ray * depth along the
 * ray), angle-weighted normals, the mean distance to their mesh neighbours as scale, and parent
 * cameras from a views image (views_channels per pixel) or, without one, camera_index alone.
 * Each vertex is a sample of confidence 1 (Samples.cpp:258-269).  The pixel ray is
 * ((x + 0.5 - ppx W) / fx, (y + 0.5 - ppy H) / fy, 1) in view space (x right, y down, z forward),
 * fx = focal_length * max(W, H), fy = fx * aspect_ratio, rotated to world space by the transpose
 * of `rotation` (world-to-view, row-major): BaseProject's PinholeCamera, whose mapping this
 * replaces, is not part of the reference. */
typedef struct tsr_depth_mesh tsr_depth_mesh;
typedef struct tsr_depth_view {
    uint32_t width, height;
    const float *depths;          /* width * height, row-major; depth along the pixel ray, <= 0: none */
    float position[3];            /* camera centre (world) */
    float rotation[9];            /* world-to-view rotation, row-major */
    float focal_length;           /* normalised by max(width, height) */
    float principal_point[2];     /* normalised; (0.5, 0.5) = image centre */
    float aspect_ratio;
    uint32_t camera_index;        /* this view's camera: the parent link without a views image */
    uint32_t views_channels;      /* parent cameras per pixel in `views`, 0 = no views image */
    const uint32_t *views;        /* width * height * views_channels camera indices (0xFFFFFFFF: none) or NULL */
} tsr_depth_view;
tsr_status tsr_depth_mesh_create(tsr_context *ctx, const tsr_depth_view *view, tsr_depth_mesh **out);
tsr_status tsr_depth_mesh_info(const tsr_depth_mesh *mesh, uint32_t *vertex_count, uint32_t *triangle_count,
                               uint32_t *cameras_per_sample);
/* positions / normals: 3 doubles per vertex; scales: 1; triangles: 3 vertex ids each;
 * parent_cameras: cameras_per_sample per vertex; pixel_to_vertex: width * height (0xFFFFFFFF:
 * pixel not triangulated).  NULL pointers are skipped. */
tsr_status tsr_depth_mesh_download(tsr_context *ctx, const tsr_depth_mesh *mesh, double *positions, double *normals,
                                   double *scales, uint32_t *triangles, uint32_t *parent_cameras,
                                   uint32_t *pixel_to_vertex);
void tsr_depth_mesh_destroy(tsr_depth_mesh *mesh);

/* Seeded synthetic scenes (analytic stand-in for SyntheticScene.cpp): config 0..4,
 * resolution_scale multiplies the image size. Views stay valid until destroy. */
tsr_status tsr_synthetic_create(int config, uint64_t seed, double resolution_scale, tsr_synthetic **out);
tsr_status tsr_synthetic_views(const tsr_synthetic *scene, tsr_samples_view *samples, tsr_cameras_view *cameras);
void tsr_synthetic_destroy(tsr_synthetic *scene);

#ifdef __cplusplus
}
#endif

#endif /* TSR_GPU_H */
