// C ABI entry points (include/tsr_gpu.h): argument checks, handle lifetime,
// exception -> status translation, host <-> device copies.
#include <atomic>
#include <cstring>
#include <string>

#include "handles.h"

using namespace tsr;

namespace {

thread_local std::string g_lastError;

tsr_status fail(tsr_status s, const std::string &m) {
    g_lastError = m;
    return s;
}

template <typename F>
tsr_status guarded(F f) {
    try {
        f();
        g_lastError.clear();
        return TSR_OK;
    } catch (const tsr_error &e) {
        return fail(e.status, e.what());
    } catch (const std::bad_alloc &) {
        return fail(TSR_ERR_OUT_OF_MEMORY, "host allocation failed");
    } catch (const std::exception &e) {
        return fail(TSR_ERR_HIP, e.what());
    }
}

template <typename T>
void upload(DBuf<T> &dst, const T *src, size_t n, hipStream_t st) {
    dst.alloc(n);
    if (n) HIPCHECK(hipMemcpyAsync(dst.p, src, n * sizeof(T), hipMemcpyHostToDevice, st));
}

template <typename T>
void download(T *dst, const DBuf<T> &src, size_t n, hipStream_t st) {
    if (dst && n) HIPCHECK(hipMemcpyAsync(dst, src.p, n * sizeof(T), hipMemcpyDeviceToHost, st));
}

void activate(const tsr_context *ctx) { HIPCHECK(hipSetDevice(ctx->device)); }

}  // namespace

void tsrSetLastError(const char *msg) { g_lastError = msg ? msg : ""; }

extern "C" {

void tsr_params_default(tsr_params *out) {
    if (!out) return;
    out->bandwidth_factor = 2.0f;
    out->confidence_threshold = 100.0f;
    out->relative_radius_for_priors = 25.0f;
    out->relative_sample_cone_length = 1.0f;
    out->max_relative_sampling_distance = 2.0f;
}

const char *tsr_last_error(void) { return g_lastError.c_str(); }

tsr_status tsr_context_create(int device, tsr_context **out) {
    if (!out) return fail(TSR_ERR_INVALID_ARGUMENT, "null output pointer");
    *out = nullptr;
    return guarded([&] {
        int count = 0;
        HIPCHECK(hipGetDeviceCount(&count));
        if (device < 0 || device >= count) throw tsr_error(TSR_ERR_INVALID_ARGUMENT, "no such HIP device");
        HIPCHECK(hipSetDevice(device));
        tsr_context *c = new tsr_context();
        c->device = device;
        c->pool = tsr::makePool(device);
        c->pool2 = tsr::makePool(device);
        (void)tsr::noAllocScratch(device);   // this GPU's copy of the no-allocation scratch array
        HIPCHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        HIPCHECK(hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking));
        for (hipEvent_t &e : c->ev) HIPCHECK(hipEventCreate(&e));
        for (hipEvent_t &e : c->kev) HIPCHECK(hipEventCreate(&e));
        *out = c;
    });
}

void tsr_context_destroy(tsr_context *ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    ctx->scratch.buf.release();
    ctx->cone.release();
    if (ctx->pool) ctx->pool->trim();   // blocks still held by live handles are freed by their owners
    ctx->pool.reset();
    (void)hipStreamSynchronize(ctx->stream2);
    if (ctx->pool2) ctx->pool2->trim();
    ctx->pool2.reset();
    for (void *p : ctx->pinned)
        if (p) (void)hipHostFree(p);
    for (hipEvent_t &e : ctx->ev) (void)hipEventDestroy(e);
    for (hipEvent_t &e : ctx->kev) (void)hipEventDestroy(e);
    (void)hipStreamSynchronize(ctx->stream2);
    (void)hipStreamDestroy(ctx->stream);
    (void)hipStreamDestroy(ctx->stream2);
    delete ctx;
}

tsr_status tsr_context_synchronize(tsr_context *ctx) {
    if (!ctx) return fail(TSR_ERR_INVALID_ARGUMENT, "null context");
    return guarded([&] { HIPCHECK(hipStreamSynchronize(ctx->stream)); });
}

tsr_status tsr_context_phase_times(tsr_context *ctx, float *ms, int n) {
    if (!ctx || !ms || n < 0 || n > 6) return fail(TSR_ERR_INVALID_ARGUMENT, "bad arguments");
    return guarded([&] {
        HIPCHECK(hipStreamSynchronize(ctx->stream));
        for (int p = 0; p < n; ++p) {
            float t = 0.f;
            if (hipEventElapsedTime(&t, ctx->ev[2 * p], ctx->ev[2 * p + 1]) != hipSuccess) t = 0.f;
            ms[p] = t;
        }
        (void)hipGetLastError();
    });
}

tsr_status tsr_context_kernel_stats(tsr_context *ctx, int kernel, double *ms, uint64_t *launches, uint64_t *bytes,
                                    uint64_t *units) {
    if (!ctx || kernel < 0 || kernel >= 8) return fail(TSR_ERR_INVALID_ARGUMENT, "bad arguments");
    const tsr_kernel_stat &k = ctx->kstat[kernel];
    if (ms) *ms = k.ms;
    if (launches) *launches = k.launches;
    if (bytes) *bytes = k.bytes;
    if (units) *units = k.units;
    return TSR_OK;
}

tsr_status tsr_context_reset_kernel_stats(tsr_context *ctx) {
    if (!ctx) return fail(TSR_ERR_INVALID_ARGUMENT, "null context");
    for (tsr_kernel_stat &k : ctx->kstat) k = tsr_kernel_stat();
    return TSR_OK;
}

tsr_status tsr_calibration_copy(tsr_context *ctx, uint64_t bytes) {
    if (!ctx || bytes == 0) return fail(TSR_ERR_INVALID_ARGUMENT, "bad arguments");
    return guarded([&] {
        activate(ctx);
        calibrationCopy(ctx, bytes);
    });
}

tsr_status tsr_device_memory(uint64_t *current_bytes, uint64_t *peak_bytes, int reset_peak) {
    MemCounter &m = memCounter();
    if (current_bytes) *current_bytes = m.current.load();
    if (peak_bytes) *peak_bytes = m.peak.load();
    if (reset_peak) m.peak.store(m.current.load());
    return TSR_OK;
}

tsr_status tsr_scene_upload(tsr_context *ctx, const tsr_samples_view *s, const tsr_cameras_view *c,
                            tsr_scene **out) {
    if (!ctx || !s || !c || !out) return fail(TSR_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    if (s->count == 0) return fail(TSR_ERR_NO_SAMPLES, "no samples");
    if (c->count == 0) return fail(TSR_ERR_NO_SAMPLES, "no cameras");
    if (s->max_cameras_per_sample == 0) return fail(TSR_ERR_INVALID_ARGUMENT, "max_cameras_per_sample must be >= 1");
    // Samples.h:395-402: camera-to-sample links are indexed by uint32 (checked before any array is read)
    if ((uint64_t)s->count * s->max_cameras_per_sample >= 0xffffffffull)
        return fail(TSR_ERR_LINK_OVERFLOW, "Number of traced rays exceeds maximum count. Supported maximum is 2^32-1.");
    if (!s->px || !s->py || !s->pz || !s->nx || !s->ny || !s->nz || !s->scale || !s->confidence || !s->parent_cameras ||
        !c->x || !c->y || !c->z)
        return fail(TSR_ERR_INVALID_ARGUMENT, "null array in view");
    return guarded([&] {
        activate(ctx);
        tsr_scene *sc = new tsr_scene();
        try {
            static std::atomic<unsigned long long> sceneIds{0};
            sc->id = ++sceneIds;
            hipStream_t st = ctx->stream;
            const u32 n = s->count;
            sc->ctx = ctx;
            sc->device = ctx->device;
            sc->n = n;
            sc->maxCams = s->max_cameras_per_sample;
            sc->camCount = c->count;
            upload(sc->px, s->px, n, st); upload(sc->py, s->py, n, st); upload(sc->pz, s->pz, n, st);
            upload(sc->nx, s->nx, n, st); upload(sc->ny, s->ny, n, st); upload(sc->nz, s->nz, n, st);
            upload(sc->scale, s->scale, n, st); upload(sc->conf, s->confidence, n, st);
            upload(sc->cams, s->parent_cameras, (size_t)n * sc->maxCams, st);
            upload(sc->cx, c->x, c->count, st); upload(sc->cy, c->y, c->count, st); upload(sc->cz, c->z, c->count, st);
            sc->hpx.assign(s->px, s->px + n); sc->hpy.assign(s->py, s->py + n); sc->hpz.assign(s->pz, s->pz + n);
            sc->hnx.assign(s->nx, s->nx + n); sc->hny.assign(s->ny, s->ny + n); sc->hnz.assign(s->nz, s->nz + n);
            sc->hscale.assign(s->scale, s->scale + n);
            HIPCHECK(hipStreamSynchronize(st));
        } catch (...) {
            delete sc;
            throw;
        }
        *out = sc;
    });
}

void tsr_scene_destroy(tsr_scene *scene) {
    if (!scene) return;
    (void)hipSetDevice(scene->device);
    (void)hipDeviceSynchronize();
    delete scene;
}

tsr_status tsr_scene_download(tsr_context *ctx, const tsr_scene *sc, float *px, float *py, float *pz, float *nx,
                              float *ny, float *nz, float *scale, float *confidence, uint32_t *parent_cameras) {
    if (!ctx || !sc) return fail(TSR_ERR_INVALID_ARGUMENT, "null argument");
    return guarded([&] {
        activate(ctx);
        hipStream_t st = ctx->stream;
        const u32 n = sc->n;
        download(px, sc->px, n, st); download(py, sc->py, n, st); download(pz, sc->pz, n, st);
        download(nx, sc->nx, n, st); download(ny, sc->ny, n, st); download(nz, sc->nz, n, st);
        download(scale, sc->scale, n, st); download(confidence, sc->conf, n, st);
        download(parent_cameras, sc->cams, (size_t)n * sc->maxCams, st);
        HIPCHECK(hipStreamSynchronize(st));
    });
}

tsr_status tsr_tree_build(tsr_context *ctx, const tsr_scene *scene, const tsr_params *params, tsr_tree **out) {
    if (!ctx || !scene || !out) return fail(TSR_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    tsr_params prm;
    tsr_params_default(&prm);
    if (params) prm = *params;
    if (!(prm.max_relative_sampling_distance > 0.0f)) return fail(TSR_ERR_INVALID_ARGUMENT, "max_relative_sampling_distance must be > 0");
    return guarded([&] {
        activate(ctx);
        tsr::PoolScope scope(ctx->pool);
        tsr_tree *t = new tsr_tree();
        t->ctx = ctx;
        t->device = ctx->device;
        try {
            recordPhase(ctx, 0, true);
            buildTree(ctx, scene, prm, t);
            recordPhase(ctx, 0, false);
            HIPCHECK(hipStreamSynchronize(ctx->stream));
        } catch (...) {
            delete t;
            throw;
        }
        *out = t;
    });
}

tsr_status tsr_tree_get_info(const tsr_tree *t, tsr_tree_info *info) {
    if (!t || !info) return fail(TSR_ERR_INVALID_ARGUMENT, "null argument");
    info->node_count = t->nodeCount;
    info->leaf_count = t->leafCount;
    info->cell_count = t->cellCount;
    info->max_depth = t->maxDepth;
    for (int a = 0; a < 3; ++a) info->root_min[a] = t->frame.rootMinF[a];
    info->root_size = t->frame.rootSizeF;
    return TSR_OK;
}

tsr_status tsr_tree_build_stats(const tsr_tree *t, uint64_t *stats4) {
    if (!t || !stats4) return fail(TSR_ERR_INVALID_ARGUMENT, "null argument");
    stats4[0] = t->samplingUnits;
    stats4[1] = t->samplingEvents;
    stats4[2] = t->samplingCreated;
    stats4[3] = t->samplingMisses;
    return TSR_OK;
}

tsr_status tsr_tree_download(tsr_context *ctx, const tsr_tree *t, uint32_t *children, uint32_t *parents,
                             uint32_t *sample_counts, uint32_t *sample_starts, uint32_t *node_to_leaf,
                             uint32_t *leaf_nodes, float *leaf_scopes, uint32_t *cells, uint32_t *sample_target,
                             uint32_t *node_depth) {
    if (!ctx || !t) return fail(TSR_ERR_INVALID_ARGUMENT, "null argument");
    return guarded([&] {
        activate(ctx);
        hipStream_t st = ctx->stream;
        const u32 N = t->nodeCount, L = t->leafCount;
        if (children) {   // Nodes::mChildren: first child index or INVALID for leaves
            std::vector<u32> c(N);
            HIPCHECK(hipMemcpyAsync(c.data(), t->child.p, (size_t)N * 4, hipMemcpyDeviceToHost, st));
            HIPCHECK(hipStreamSynchronize(st));
            for (u32 i = 0; i < N; ++i) children[i] = (c[i] & LEAF_BIT) ? INVALID : c[i];
        }
        download(parents, t->parents, N, st);
        download(sample_counts, t->sampleCounts, N, st);
        download(sample_starts, t->sampleStarts, N, st);
        download(node_to_leaf, t->nodeToLeaf, N, st);
        download(leaf_nodes, t->leafNode, L, st);
        download(cells, t->cells, 8 * (size_t)t->cellCount, st);
        download(sample_target, t->sampleTarget, t->sampleTarget.n, st);
        if (node_depth) {
            std::vector<uint8_t> d(N);
            HIPCHECK(hipMemcpyAsync(d.data(), t->nodeDepth.p, N, hipMemcpyDeviceToHost, st));
            HIPCHECK(hipStreamSynchronize(st));
            for (u32 i = 0; i < N; ++i) node_depth[i] = d[i];
        }
        if (leaf_scopes) {   // Scope: min corner + size, rounded to float
            std::vector<u64> loc(L);
            HIPCHECK(hipMemcpyAsync(loc.data(), t->leafLoc.p, (size_t)L * 8, hipMemcpyDeviceToHost, st));
            HIPCHECK(hipStreamSynchronize(st));
            for (u32 l = 0; l < L; ++l) {
                u32 k[3];
                locCoords(loc[l], k[0], k[1], k[2]);
                const u32 d = locDepth(loc[l]);
                for (int a = 0; a < 3; ++a) leaf_scopes[4 * (size_t)l + a] = (float)t->frame.B(a, k[a], d);
                leaf_scopes[4 * (size_t)l + 3] = (float)t->frame.h(d);
            }
        }
        HIPCHECK(hipStreamSynchronize(st));
    });
}

tsr_status tsr_tree_samples_download(tsr_context *ctx, const tsr_tree *t, float *px, float *py, float *pz, float *nx,
                                     float *ny, float *nz, float *scale, float *confidence, uint32_t *parent_cameras) {
    if (!ctx || !t) return fail(TSR_ERR_INVALID_ARGUMENT, "null argument");
    return guarded([&] {
        activate(ctx);
        hipStream_t st = ctx->stream;
        const u32 n = t->sampleCount;
        download(px, t->px, n, st); download(py, t->py, n, st); download(pz, t->pz, n, st);
        download(nx, t->nx, n, st); download(ny, t->ny, n, st); download(nz, t->nz, n, st);
        download(scale, t->scale, n, st); download(confidence, t->conf, n, st);
        download(parent_cameras, t->cams, (size_t)n * t->maxCams, st);
        HIPCHECK(hipStreamSynchronize(st));
    });
}

void tsr_tree_destroy(tsr_tree *t) {
    if (!t) return;
    (void)hipSetDevice(t->device);
    (void)hipDeviceSynchronize();
    delete t;
}

tsr_status tsr_tree_neighbor_count(tsr_context *ctx, tsr_tree *tree, uint64_t *total) {
    if (!ctx || !tree || !total) return fail(TSR_ERR_INVALID_ARGUMENT, "null argument");
    return guarded([&] {
        activate(ctx);
        tsr::PoolScope scope(ctx->pool);
        leafNeighbours(ctx, tree);
        *total = tree->nbrCount;
    });
}

tsr_status tsr_tree_neighbors_download(tsr_context *ctx, tsr_tree *tree, uint32_t *offsets, uint32_t *neighbors) {
    if (!ctx || !tree) return fail(TSR_ERR_INVALID_ARGUMENT, "null argument");
    return guarded([&] {
        activate(ctx);
        tsr::PoolScope scope(ctx->pool);
        leafNeighbours(ctx, tree);
        download(offsets, tree->nbrOffsets, (size_t)tree->leafCount + 1, ctx->stream);
        download(neighbors, tree->nbrs, tree->nbrCount, ctx->stream);
        HIPCHECK(hipStreamSynchronize(ctx->stream));
    });
}

tsr_status tsr_tree_save(tsr_context *ctx, tsr_tree *tree, const char *nodes_path, const char *leaves_path,
                         const char *cells_path) {
    if (!ctx || !tree) return fail(TSR_ERR_INVALID_ARGUMENT, "null argument");
    return guarded([&] {
        activate(ctx);
        tsr::PoolScope scope(ctx->pool);
        saveTree(ctx, tree, nodes_path, leaves_path, cells_path);
    });
}

tsr_status tsr_tree_load(tsr_context *ctx, const tsr_scene *scene, const tsr_params *params, const char *nodes_path,
                         const char *leaves_path, const char *cells_path, tsr_tree **out) {
    if (!ctx || !scene || !nodes_path || !out) return fail(TSR_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    tsr_params prm;
    tsr_params_default(&prm);
    if (params) prm = *params;
    return guarded([&] {
        activate(ctx);
        tsr::PoolScope scope(ctx->pool);
        tsr_tree *t = new tsr_tree();
        t->ctx = ctx;
        t->device = ctx->device;
        try {
            loadTree(ctx, scene, prm, nodes_path, leaves_path, cells_path, t);
        } catch (...) {
            delete t;
            throw;
        }
        *out = t;
    });
}

tsr_status tsr_tree_samples_save(tsr_context *ctx, const tsr_tree *tree, const char *path) {
    if (!ctx || !tree || !path) return fail(TSR_ERR_INVALID_ARGUMENT, "null argument");
    return guarded([&] {
        activate(ctx);
        saveTreeSamples(ctx, tree, path);
    });
}

tsr_status tsr_samples_file_info(const char *path, uint32_t *count, uint32_t *max_cameras_per_sample) {
    if (!path || !count || !max_cameras_per_sample) return fail(TSR_ERR_INVALID_ARGUMENT, "null argument");
    return guarded([&] { samplesFileInfo(path, *count, *max_cameras_per_sample); });
}

tsr_status tsr_samples_file_read(const char *path, float *px, float *py, float *pz, float *nx, float *ny, float *nz,
                                 float *scale, float *confidence, uint32_t *parent_cameras) {
    if (!path) return fail(TSR_ERR_INVALID_ARGUMENT, "null argument");
    return guarded([&] { samplesFileRead(path, px, py, pz, nx, ny, nz, scale, confidence, parent_cameras); });
}

tsr_status tsr_cameras_file_write(const char *path, uint32_t count, const double *data13, const uint32_t *view_ids,
                                  int real_bytes) {
    if (!path || (count && (!data13 || !view_ids))) return fail(TSR_ERR_INVALID_ARGUMENT, "null argument");
    if (real_bytes != 4 && real_bytes != 8) return fail(TSR_ERR_INVALID_ARGUMENT, "real_bytes must be 4 or 8");
    return guarded([&] { camerasFileWrite(path, count, data13, view_ids, real_bytes); });
}

tsr_status tsr_cameras_file_info(const char *path, uint32_t *count) {
    if (!path || !count) return fail(TSR_ERR_INVALID_ARGUMENT, "null argument");
    return guarded([&] { camerasFileInfo(path, *count); });
}

tsr_status tsr_cameras_file_read(const char *path, double *data13, uint32_t *view_ids) {
    if (!path) return fail(TSR_ERR_INVALID_ARGUMENT, "null argument");
    return guarded([&] { camerasFileRead(path, data13, view_ids); });
}

tsr_status tsr_field_save(tsr_context *ctx, const tsr_field *field, const char *path, int real_bytes) {
    if (!ctx || !field || !path) return fail(TSR_ERR_INVALID_ARGUMENT, "null argument");
    if (real_bytes != 4 && real_bytes != 8) return fail(TSR_ERR_INVALID_ARGUMENT, "real_bytes must be 4 or 8");
    return guarded([&] {
        activate(ctx);
        saveField(ctx, field, path, real_bytes);
    });
}

tsr_status tsr_field_load(tsr_context *ctx, const tsr_tree *tree, const tsr_params *params, const char *path,
                          tsr_field **out) {
    if (!ctx || !tree || !path || !out) return fail(TSR_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    tsr_params prm;
    tsr_params_default(&prm);
    if (params) prm = *params;
    return guarded([&] {
        activate(ctx);
        tsr::PoolScope scope(ctx->pool);
        tsr_field *f = new tsr_field();
        f->ctx = ctx;
        f->device = ctx->device;
        try {
            loadField(ctx, tree, prm, path, f);
        } catch (...) {
            delete f;
            throw;
        }
        *out = f;
    });
}

tsr_status tsr_occupancy_compute(tsr_context *ctx, const tsr_tree *tree, const tsr_scene *scene,
                                 const tsr_params *params, int flags, const tsr_culler *culler, tsr_field **out) {
    if (!ctx || !tree || !scene || !out) return fail(TSR_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    if (tree->maxCams != scene->maxCams) return fail(TSR_ERR_INVALID_ARGUMENT, "scene and tree camera links differ");
    const int mode = flags & 0xff;
    if (flags & ~(0xff | TSR_CULL_LEARNED | TSR_PRIORS_ALL_LEAVES | TSR_PRIORS_CERTIFIED | TSR_EMPTINESS_TRAVERSAL))
        return fail(TSR_ERR_INVALID_ARGUMENT, "unknown flags");
    if (mode != TSR_ACCUM_KAHAN_SEGMENTED && mode != TSR_ACCUM_FP32_ATOMIC)
        return fail(TSR_ERR_INVALID_ARGUMENT, "unknown accumulation mode");
    if ((flags & TSR_CULL_LEARNED) && !culler) return fail(TSR_ERR_INVALID_ARGUMENT, "learned routing needs a culler");
    tsr_params prm;
    tsr_params_default(&prm);
    if (params) prm = *params;
    return guarded([&] {
        activate(ctx);
        tsr::PoolScope scope(ctx->pool);
        tsr_field *f = new tsr_field();
        f->ctx = ctx;
        f->device = ctx->device;
        try {
            computeOccupancy(ctx, tree, scene, prm, flags, (flags & TSR_CULL_LEARNED) ? culler : nullptr, f);
        } catch (...) {
            delete f;
            throw;
        }
        *out = f;
    });
}

tsr_status tsr_field_download(tsr_context *ctx, const tsr_field *f, float *kernel_sums, float *cone_lengths,
                              double *priors, uint32_t *node_states, double *occupancy, double *cone_count) {
    if (!ctx || !f) return fail(TSR_ERR_INVALID_ARGUMENT, "null argument");
    return guarded([&] {
        activate(ctx);
        hipStream_t st = ctx->stream;
        download(kernel_sums, f->ksum, 2 * (size_t)f->leafCount, st);
        download(cone_lengths, f->clen, 2 * (size_t)f->leafCount, st);
        download(priors, f->prior, f->leafCount, st);
        download(node_states, f->states, f->nodeCount, st);
        download(occupancy, f->occ, f->leafCount, st);
        if (cone_count) *cone_count = f->coneCount;
        HIPCHECK(hipStreamSynchronize(st));
    });
}

tsr_status tsr_field_hit_counts(const tsr_field *f, uint64_t *hits2) {
    if (!f || !hits2) return fail(TSR_ERR_INVALID_ARGUMENT, "null argument");
    hits2[0] = f->hits[0];
    hits2[1] = f->hits[1];
    return TSR_OK;
}

namespace {
tsr_status samplesChanged(tsr_context *ctx, tsr_field *f, const tsr_tree *tree, const tsr_scene *scene,
                          const uint32_t *samples, uint32_t count, bool removal) {
    if (!ctx || !f || !tree || !scene || (count && !samples)) return fail(TSR_ERR_INVALID_ARGUMENT, "null argument");
    if (f->leafCount != tree->leafCount || f->nodeCount != tree->nodeCount)
        return fail(TSR_ERR_INVALID_ARGUMENT, "field does not belong to the tree");
    if (tree->maxCams != scene->maxCams) return fail(TSR_ERR_INVALID_ARGUMENT, "scene and tree camera links differ");
    if (!f->states.p || !f->clen.p) return fail(TSR_ERR_INVALID_ARGUMENT, "field has no passes to update");
    return guarded([&] {
        activate(ctx);
        tsr::PoolScope scope(ctx->pool);
        updateOccupancy(ctx, tree, scene, samples, count, removal, f);
    });
}
}  // namespace

tsr_status tsr_occupancy_add_samples(tsr_context *ctx, tsr_field *field, const tsr_tree *tree, const tsr_scene *scene,
                                     const uint32_t *samples, uint32_t count) {
    return samplesChanged(ctx, field, tree, scene, samples, count, false);
}

tsr_status tsr_occupancy_erase_samples(tsr_context *ctx, tsr_field *field, const tsr_tree *tree,
                                       const tsr_scene *scene, const uint32_t *samples, uint32_t count) {
    return samplesChanged(ctx, field, tree, scene, samples, count, true);
}

tsr_status tsr_tree_erase_samples(tsr_context *ctx, tsr_tree *tree, const uint8_t *inliers) {
    if (!ctx || !tree || (tree->sampleCount && !inliers)) return fail(TSR_ERR_INVALID_ARGUMENT, "null argument");
    return guarded([&] {
        activate(ctx);
        tsr::PoolScope scope(ctx->pool);
        eraseTreeSamples(ctx, tree, inliers);
    });
}

tsr_status tsr_field_prior_stages(const tsr_field *f, uint32_t *stages6) {
    if (!f || !stages6) return fail(TSR_ERR_INVALID_ARGUMENT, "null argument");
    for (int i = 0; i < 6; ++i) stages6[i] = f->priorStages[i];
    return TSR_OK;
}

tsr_status tsr_field_prior_levels(tsr_context *ctx, const tsr_field *f, uint8_t *levels) {
    if (!ctx || !f || !levels) return fail(TSR_ERR_INVALID_ARGUMENT, "null argument");
    return guarded([&] {
        HIPCHECK(hipSetDevice(ctx->device));
        const u32 L = f->leafCount;
        if (f->levels.p) {
            HIPCHECK(hipMemcpyAsync(levels, f->levels.p, L, hipMemcpyDeviceToHost, ctx->stream));
            HIPCHECK(hipStreamSynchronize(ctx->stream));
            return;
        }
        // plain field: every queried prior is exact, the others are NaN
        std::vector<double> p(L);
        if (L) HIPCHECK(hipMemcpyAsync(p.data(), f->prior.p, (size_t)L * 8, hipMemcpyDeviceToHost, ctx->stream));
        HIPCHECK(hipStreamSynchronize(ctx->stream));
        for (u32 l = 0; l < L; ++l) levels[l] = p[l] == p[l] ? 2 : 255;
    });
}

void tsr_field_destroy(tsr_field *f) {
    if (!f) return;
    (void)hipSetDevice(f->device);
    (void)hipDeviceSynchronize();
    delete f;
}

tsr_status tsr_dmc_extract(tsr_context *ctx, const tsr_tree *tree, const tsr_field *field, tsr_mesh **out) {
    if (!ctx || !tree || !field || !out) return fail(TSR_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    if (field->leafCount != tree->leafCount) return fail(TSR_ERR_INVALID_ARGUMENT, "field does not belong to the tree");
    return guarded([&] {
        activate(ctx);
        tsr::PoolScope scope(ctx->pool);
        tsr_mesh *m = new tsr_mesh();
        m->ctx = ctx;
        m->device = ctx->device;
        try {
            recordPhase(ctx, 5, true);
            extractDMC(ctx, tree, field->occ.p, m);
            recordPhase(ctx, 5, false);
            HIPCHECK(hipStreamSynchronize(ctx->stream));
        } catch (...) {
            delete m;
            throw;
        }
        *out = m;
    });
}

tsr_status tsr_dmc_extract_values(tsr_context *ctx, const tsr_tree *tree, const double *occupancy, tsr_mesh **out) {
    if (!ctx || !tree || !occupancy || !out) return fail(TSR_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    return guarded([&] {
        activate(ctx);
        tsr::PoolScope scope(ctx->pool);
        DBuf<double> occ;
        upload(occ, occupancy, tree->leafCount, ctx->stream);
        tsr_mesh *m = new tsr_mesh();
        m->ctx = ctx;
        m->device = ctx->device;
        try {
            extractDMC(ctx, tree, occ.p, m);
            HIPCHECK(hipStreamSynchronize(ctx->stream));
        } catch (...) {
            delete m;
            throw;
        }
        *out = m;
    });
}

tsr_status tsr_dmc_extract_noalloc(tsr_context *ctx, const tsr_tree *tree, const tsr_field *field,
                                   uint32_t max_vertices, uint32_t max_triangles, double *positions, double *scales,
                                   uint32_t *edge_keys, uint32_t *triangles, uint32_t *vertex_count,
                                   uint32_t *triangle_count) {
    if (!ctx || !tree || !field || !vertex_count || !triangle_count)
        return fail(TSR_ERR_INVALID_ARGUMENT, "null argument");
    *vertex_count = 0;
    *triangle_count = 0;
    if ((max_vertices && (!positions || !scales || !edge_keys)) || (max_triangles && !triangles))
        return fail(TSR_ERR_INVALID_ARGUMENT, "null output buffer with a non-zero capacity");
    if (field->leafCount != tree->leafCount) return fail(TSR_ERR_INVALID_ARGUMENT, "field does not belong to the tree");
    return guarded([&] {
        activate(ctx);
        extractDMCNoAlloc(ctx, tree, field->occ.p, max_vertices, max_triangles, positions, scales, edge_keys,
                          triangles, vertex_count, triangle_count);
    });
}

tsr_status tsr_crust_postprocess(tsr_context *ctx, tsr_mesh *mesh, uint32_t min_triangle_isle_size) {
    if (!ctx || !mesh) return fail(TSR_ERR_INVALID_ARGUMENT, "null argument");
    return guarded([&] {
        activate(ctx);
        tsr::PoolScope scope(ctx->pool);
        postprocessCrust(ctx, mesh, min_triangle_isle_size, true, true);
    });
}

tsr_status tsr_mesh_download_normals(tsr_context *ctx, const tsr_mesh *m, double *normals) {
    if (!ctx || !m || !normals) return fail(TSR_ERR_INVALID_ARGUMENT, "null argument");
    return guarded([&] {
        activate(ctx);
        const size_t n = 3 * (size_t)m->vertexCount;
        if (m->normals.p && m->normals.n >= n) download(normals, m->normals, n, ctx->stream);
        else std::memset(normals, 0, n * sizeof(double));
        HIPCHECK(hipStreamSynchronize(ctx->stream));
    });
}

tsr_status tsr_mesh_get_info(const tsr_mesh *m, tsr_mesh_info *info) {
    if (!m || !info) return fail(TSR_ERR_INVALID_ARGUMENT, "null argument");
    info->vertex_count = m->vertexCount;
    info->triangle_count = m->triangleCount;
    return TSR_OK;
}

tsr_status tsr_mesh_download(tsr_context *ctx, const tsr_mesh *m, double *positions, double *scales,
                             uint32_t *edge_keys, uint32_t *triangles) {
    if (!ctx || !m) return fail(TSR_ERR_INVALID_ARGUMENT, "null argument");
    return guarded([&] {
        activate(ctx);
        hipStream_t st = ctx->stream;
        download(positions, m->pos, 3 * (size_t)m->vertexCount, st);
        download(scales, m->scale, m->vertexCount, st);
        download(edge_keys, m->keys, 2 * (size_t)m->vertexCount, st);
        download(triangles, m->tris, 3 * (size_t)m->triangleCount, st);
        HIPCHECK(hipStreamSynchronize(st));
    });
}

void tsr_mesh_destroy(tsr_mesh *m) {
    if (!m) return;
    (void)hipSetDevice(m->device);
    (void)hipDeviceSynchronize();
    delete m;
}

tsr_status tsr_synthetic_create(int config, uint64_t seed, double resolution_scale, tsr_synthetic **out) {
    if (!out) return fail(TSR_ERR_INVALID_ARGUMENT, "null output pointer");
    *out = nullptr;
    if (!(resolution_scale > 0.0)) return fail(TSR_ERR_INVALID_ARGUMENT, "resolution_scale must be > 0");
    return guarded([&] { *out = tsr_synth_build(config, seed, resolution_scale); });
}

tsr_status tsr_synthetic_views(const tsr_synthetic *s, tsr_samples_view *samples, tsr_cameras_view *cameras) {
    if (!s) return fail(TSR_ERR_INVALID_ARGUMENT, "null scene");
    if (samples) {
        samples->count = (uint32_t)s->px.size();
        samples->max_cameras_per_sample = 1;
        samples->px = s->px.data(); samples->py = s->py.data(); samples->pz = s->pz.data();
        samples->nx = s->nx.data(); samples->ny = s->ny.data(); samples->nz = s->nz.data();
        samples->scale = s->scale.data();
        samples->confidence = s->conf.data();
        samples->parent_cameras = s->cams.data();
    }
    if (cameras) {
        cameras->count = (uint32_t)s->cx.size();
        cameras->x = s->cx.data(); cameras->y = s->cy.data(); cameras->z = s->cz.data();
    }
    return TSR_OK;
}

void tsr_synthetic_destroy(tsr_synthetic *s) { delete s; }

}  // extern "C"
