// Sample preparation from depth maps (DepthImage.cpp:191-413, Mesh.cpp:63-124,
// Scene.cpp:484-573, Samples.cpp:224-273): a depth map is triangulated block by block,
// its vertices get world positions, angle-weighted normals and the mean distance to their
// neighbours as scale, and parent camera links from a views image or, without one, the
// depth map's own camera.  The vertices become samples of confidence 1.
//
// The pixel -> view ray mapping of BaseProject's PinholeCamera is not part of the reference;
// here the ray through the centre of pixel (x, y) is ((x + 0.5 - cx) / fx, (y + 0.5 - cy) / fy, 1)
// in view space (x right, y down, z forward), rotated to world space by the transpose of the
// world-to-view rotation, with fx = f max(W, H), fy = fx * aspect, (cx, cy) = principal point
// times (W, H).  Decisions and geometry are evaluated in FP64 (the reference's Real is float).
#include <algorithm>
#include <cmath>

#include "handles.h"

using namespace tsr;

struct tsr_depth_mesh {
    tsr_context *ctx = nullptr;
    int device = 0;
    u32 width = 0, height = 0, vertexCount = 0, triangleCount = 0, camerasPerSample = 1;
    DBuf<double> pos, normals, scales;
    DBuf<u32> tris, cams, pixelToVertex;
};

namespace tsr {

namespace {

constexpr u32 BLOCK = 256;
constexpr double DEPTH_DIFFERENCE_FACTOR = 5.0;   // DepthImage.cpp:25
constexpr uint8_t NO_TRIANGLE = 0xff;

struct DepthCam {
    double pos[3], rot[9];   // centre; world-to-view rotation, row-major
    double fx, fy, cx, cy;   // pixels
};

// DepthImage.cpp:227-234: the 4 triangles of a 2x2 block (corner k at (x + k % 2, y + k / 2))
__constant__ uint8_t BLOCK_TRIANGLES[4][3] = {{0, 2, 1}, {0, 3, 1}, {0, 2, 3}, {1, 2, 3}};

__device__ void viewRay(const DepthCam &c, u32 x, u32 y, double v[3]) {
    v[0] = ((double)x + 0.5 - c.cx) / c.fx;
    v[1] = ((double)y + 0.5 - c.cy) / c.fy;
    v[2] = 1.0;
}

// DepthImage.cpp:377-392 isDepthDiscontinuity
__device__ bool discontinuity(const double fp[4], const double d[4], u32 v0, u32 v1) {
    const u32 closer = d[v0] < d[v1] ? v0 : v1, further = d[v0] < d[v1] ? v1 : v0;
    double dd = DEPTH_DIFFERENCE_FACTOR;
    if (v0 + v1 == 3) dd *= 1.4142135623730951;   // diagonal edge
    return d[further] - d[closer] > fp[closer] * dd;
}

// DepthImage.cpp:225-351 triangulateBlock, the triangle choice: up to two block triangles
// per block (NO_TRIANGLE where none), tri[2 b], tri[2 b + 1]; count[b] valid ones
__global__ void k_depthBlocks(const float *__restrict__ depth, u32 W, u32 H, DepthCam c, uint8_t *__restrict__ tri,
                              u32 *__restrict__ count) {
    const u32 b = blockIdx.x * BLOCK + threadIdx.x;
    const u32 bw = W - 1;
    if (b >= bw * (H - 1)) return;
    const u32 x = b % bw, y = b / bw, p = y * W + x;
    const double d[4] = {depth[p], depth[p + 1], depth[p + W], depth[p + W + 1]};
    u32 mask = 0, avail = 0;
    for (u32 k = 0; k < 4; ++k)
        if (d[k] > 0.0) {
            mask |= 1u << k;
            ++avail;
        }
    uint8_t t[2] = {NO_TRIANGLE, NO_TRIANGLE};
    if (avail >= 3) {
        switch (mask) {
            case 7: t[0] = 0; break;
            case 11: t[0] = 1; break;
            case 13: t[0] = 2; break;
            case 14: t[0] = 3; break;
            case 15:
                if (fabs(d[0] - d[3]) < fabs(d[1] - d[2])) {
                    t[0] = 1;
                    t[1] = 2;
                } else {
                    t[0] = 0;
                    t[1] = 3;
                }
                break;
            default: break;
        }
    }
    // pixel footprints at the corners' depths (DepthImage.cpp:297-309): the ray length at
    // unit z scales depth along the ray to z, one pixel covers z / fx
    double fp[4] = {0.0, 0.0, 0.0, 0.0};
    for (u32 k = 0; k < 4; ++k) {
        if (!(d[k] > 0.0)) continue;
        double v[3];
        viewRay(c, x + (k & 1u), y + (k >> 1), v);
        fp[k] = d[k] / sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) / c.fx;
    }
    u32 n = 0;
    for (u32 s = 0; s < 2; ++s) {
        if (t[s] == NO_TRIANGLE) continue;
        const uint8_t *tr = BLOCK_TRIANGLES[t[s]];
        for (u32 e = 0; e < 3; ++e)
            if (discontinuity(fp, d, tr[e], tr[(e + 1) % 3])) {
                t[s] = NO_TRIANGLE;
                break;
            }
        n += t[s] != NO_TRIANGLE;
    }
    tri[2 * b] = t[0];
    tri[2 * b + 1] = t[1];
    count[b] = n;
}

// vertex ids in first-use order (DepthImage.cpp:336-349): a pixel's key is the position of its
// first use, block-major, then triangle slot, then corner
__global__ void k_firstUse(const uint8_t *__restrict__ tri, u32 W, u32 H, u32 *__restrict__ firstKey) {
    const u32 b = blockIdx.x * BLOCK + threadIdx.x;
    const u32 bw = W - 1;
    if (b >= bw * (H - 1)) return;
    const u32 x = b % bw, y = b / bw, p = y * W + x;
    for (u32 s = 0; s < 2; ++s) {
        const uint8_t t = tri[2 * b + s];
        if (t == NO_TRIANGLE) continue;
        for (u32 k = 0; k < 3; ++k) {
            const u32 corner = BLOCK_TRIANGLES[t][k];
            atomicMin(&firstKey[p + (corner >> 1) * W + (corner & 1u)], 6u * b + 3u * s + k);
        }
    }
}

__global__ void k_usedPixels(const u32 *__restrict__ firstKey, u32 n, u32 *__restrict__ flag,
                             u32 *__restrict__ pixel) {
    const u32 i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    flag[i] = firstKey[i] != INVALID ? 1u : 0u;
    pixel[i] = i;
}

__global__ void k_scatterIds(const u32 *__restrict__ pixelSorted, u32 V, u32 *__restrict__ pixelToVertex) {
    const u32 v = blockIdx.x * BLOCK + threadIdx.x;
    if (v < V) pixelToVertex[pixelSorted[v]] = v;
}

__global__ void k_emitDepthTriangles(const uint8_t *__restrict__ tri, const u32 *__restrict__ offsets, u32 W, u32 H,
                                     const u32 *__restrict__ pixelToVertex, u32 *__restrict__ tris) {
    const u32 b = blockIdx.x * BLOCK + threadIdx.x;
    const u32 bw = W - 1;
    if (b >= bw * (H - 1)) return;
    const u32 x = b % bw, y = b / bw, p = y * W + x;
    u32 o = offsets[b];
    for (u32 s = 0; s < 2; ++s) {
        const uint8_t t = tri[2 * b + s];
        if (t == NO_TRIANGLE) continue;
        for (u32 k = 0; k < 3; ++k) {
            const u32 corner = BLOCK_TRIANGLES[t][k];
            tris[3 * o + k] = pixelToVertex[p + (corner >> 1) * W + (corner & 1u)];
        }
        ++o;
    }
}

// DepthImage.cpp:415-438 setVertexPositions: centre + unit world ray * depth (along the ray)
__global__ void k_depthPositions(const u32 *__restrict__ pixelSorted, u32 V, const float *__restrict__ depth, u32 W,
                                 DepthCam c, double *__restrict__ pos) {
    const u32 v = blockIdx.x * BLOCK + threadIdx.x;
    if (v >= V) return;
    const u32 p = pixelSorted[v];
    double r[3];
    viewRay(c, p % W, p / W, r);
    double w[3];
    for (int a = 0; a < 3; ++a) w[a] = c.rot[a] * r[0] + c.rot[3 + a] * r[1] + c.rot[6 + a] * r[2];
    const double len = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
    for (int a = 0; a < 3; ++a) pos[3 * (size_t)v + a] = c.pos[a] + w[a] / len * (double)depth[p];
}

// directed neighbour pairs (a << 32 | b) of every triangle corner (FlexibleMesh::findVertexNeighbors)
__global__ void k_neighbourPairs(const u32 *__restrict__ tris, u32 T, u64 *__restrict__ pairs) {
    const u32 t = blockIdx.x * BLOCK + threadIdx.x;
    if (t >= T) return;
    const u32 v[3] = {tris[3 * t], tris[3 * t + 1], tris[3 * t + 2]};
    u32 o = 6 * t;
    for (u32 i = 0; i < 3; ++i)
        for (u32 j = 0; j < 3; ++j)
            if (i != j) pairs[o++] = ((u64)v[i] << 32) | v[j];
}

__global__ void k_pairRuns(const u64 *__restrict__ pairs, u32 n, u32 *__restrict__ start, u32 *__restrict__ end) {
    const u32 i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    const u32 a = (u32)(pairs[i] >> 32);
    if (i == 0 || (u32)(pairs[i - 1] >> 32) != a) start[a] = i;
    if (i + 1 == n || (u32)(pairs[i + 1] >> 32) != a) end[a] = i + 1;
}

// Mesh.cpp:63-124 computeVertexScales / computeAverageDistance: the mean distance to the
// distinct neighbours (a vertex without any: -REAL_MAX)
__global__ void k_vertexScales(const u64 *__restrict__ pairs, const u32 *__restrict__ start,
                               const u32 *__restrict__ end, u32 V, const double *__restrict__ pos,
                               double *__restrict__ scale) {
    const u32 v = blockIdx.x * BLOCK + threadIdx.x;
    if (v >= V) return;
    double sum = 0.0;
    u32 cnt = 0;
    for (u32 j = start[v]; j < end[v]; ++j) {
        if (j > start[v] && pairs[j] == pairs[j - 1]) continue;
        const u32 b = (u32)pairs[j];
        if (b == v) continue;
        double d2 = 0.0;
        for (int a = 0; a < 3; ++a) {
            const double d = pos[3 * (size_t)b + a] - pos[3 * (size_t)v + a];
            d2 += d * d;
        }
        sum += sqrt(d2);
        ++cnt;
    }
    scale[v] = cnt ? sum / cnt : -3.4028234663852886e38;
}

// Scene.cpp:526-573 loadCameraIndices: the views image's parent cameras of each vertex's pixel,
// or the depth map's own camera
__global__ void k_vertexCameras(const u32 *__restrict__ pixelSorted, u32 V, const u32 *__restrict__ views,
                                u32 channels, u32 cameraIndex, u32 *__restrict__ cams) {
    const u32 v = blockIdx.x * BLOCK + threadIdx.x;
    if (v >= V) return;
    if (!views) {
        cams[v] = cameraIndex;
        return;
    }
    const u32 p = pixelSorted[v];
    for (u32 c = 0; c < channels; ++c) cams[(size_t)v * channels + c] = views[(size_t)p * channels + c];
}

__global__ void k_fillU(u32 *p, size_t n, u32 v) {
    const size_t i = (size_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) p[i] = v;
}

}  // namespace

// DepthImage::triangulate + createFlexibleMesh and Scene::loadCameraIndices for one view
void depthMesh(tsr_context *ctx, const tsr_depth_view &view, tsr_depth_mesh *m) {
    hipStream_t st = ctx->stream;
    Scratch &sc = ctx->scratch;
    const u32 W = view.width, H = view.height;
    const u64 pixels64 = (u64)W * H;
    if (W < 2 || H < 2) throw tsr_error(TSR_ERR_INVALID_ARGUMENT, "a depth map needs at least 2 x 2 pixels");
    if (pixels64 * 6 >= (u64)INVALID) throw tsr_error(TSR_ERR_INVALID_ARGUMENT, "depth map too large");
    if (!(view.focal_length > 0.f) || !(view.aspect_ratio > 0.f))
        throw tsr_error(TSR_ERR_INVALID_ARGUMENT, "focal length and aspect ratio must be positive");
    const u32 P = (u32)pixels64, B = (W - 1) * (H - 1);
    DepthCam c;
    for (int a = 0; a < 3; ++a) c.pos[a] = view.position[a];
    for (int a = 0; a < 9; ++a) c.rot[a] = view.rotation[a];
    c.fx = (double)view.focal_length * (double)std::max(W, H);
    c.fy = c.fx * (double)view.aspect_ratio;
    c.cx = (double)view.principal_point[0] * W;
    c.cy = (double)view.principal_point[1] * H;
    m->width = W;
    m->height = H;
    DBuf<float> depth(P);
    HIPCHECK(hipMemcpyAsync(depth.p, view.depths, (size_t)P * 4, hipMemcpyHostToDevice, st));
    DBuf<uint8_t> tri(2 * (size_t)B);
    DBuf<u32> count(B), offsets(B), firstKey(P), flag(P), pixel(P), keysSel(P), pixSel(P), counter(1);
    k_depthBlocks<<<gridFor(B, BLOCK), BLOCK, 0, st>>>(depth.p, W, H, c, tri.p, count.p);
    exclusiveScan(sc, count.p, offsets.p, B, st);
    const u32 T = readScalar(offsets.p + B - 1, st) + readScalar(count.p + B - 1, st);
    k_fillU<<<gridFor(P, BLOCK), BLOCK, 0, st>>>(firstKey.p, P, INVALID);
    k_firstUse<<<gridFor(B, BLOCK), BLOCK, 0, st>>>(tri.p, W, H, firstKey.p);
    // used pixels ordered by their first use = vertex ids
    k_usedPixels<<<gridFor(P, BLOCK), BLOCK, 0, st>>>(firstKey.p, P, flag.p, pixel.p);
    selectFlagged(sc, firstKey.p, flag.p, keysSel.p, counter.p, P, st);
    selectFlagged(sc, pixel.p, flag.p, pixSel.p, counter.p, P, st);
    const u32 V = readScalar(counter.p, st);
    DBuf<u32> keysSorted(std::max(V, 1u)), pixSorted(std::max(V, 1u));
    if (V) sortPairs(sc, keysSel.p, keysSorted.p, pixSel.p, pixSorted.p, V, 0, 32, st);
    m->pixelToVertex.alloc(P);
    k_fillU<<<gridFor(P, BLOCK), BLOCK, 0, st>>>(m->pixelToVertex.p, P, INVALID);
    if (V) k_scatterIds<<<gridFor(V, BLOCK), BLOCK, 0, st>>>(pixSorted.p, V, m->pixelToVertex.p);
    m->tris.alloc(3 * (size_t)std::max(T, 1u));
    if (T) k_emitDepthTriangles<<<gridFor(B, BLOCK), BLOCK, 0, st>>>(tri.p, offsets.p, W, H, m->pixelToVertex.p,
                                                                     m->tris.p);
    m->pos.alloc(3 * (size_t)std::max(V, 1u));
    if (V) k_depthPositions<<<gridFor(V, BLOCK), BLOCK, 0, st>>>(pixSorted.p, V, depth.p, W, c, m->pos.p);
    m->vertexCount = V;
    m->triangleCount = T;
    // normals: angle-weighted over the depth mesh (Mesh::computeNormalsWeightedByAngles)
    if (V && T) {
        tsr_mesh tmpMesh;
        tmpMesh.vertexCount = V;
        tmpMesh.triangleCount = T;
        tmpMesh.pos = std::move(m->pos);
        tmpMesh.tris = std::move(m->tris);
        postprocessCrust(ctx, &tmpMesh, 0, false, true);
        m->pos = std::move(tmpMesh.pos);
        m->tris = std::move(tmpMesh.tris);
        m->normals = std::move(tmpMesh.normals);
    } else {
        m->normals.alloc(3 * (size_t)std::max(V, 1u));
        HIPCHECK(hipMemsetAsync(m->normals.p, 0, m->normals.bytes(), st));
    }
    // scales: mean distance to the distinct neighbours over the triangles
    m->scales.alloc(std::max(V, 1u));
    if (V) {
        const u32 np = 6 * T;
        DBuf<u32> start(V), end(V);
        HIPCHECK(hipMemsetAsync(start.p, 0, (size_t)V * 4, st));
        HIPCHECK(hipMemsetAsync(end.p, 0, (size_t)V * 4, st));
        if (np) {
            DBuf<u64> pairs(np), pairsS(np);
            k_neighbourPairs<<<gridFor(T, BLOCK), BLOCK, 0, st>>>(m->tris.p, T, pairs.p);
            sortKeys(sc, pairs.p, pairsS.p, np, 0, 64, st);
            k_pairRuns<<<gridFor(np, BLOCK), BLOCK, 0, st>>>(pairsS.p, np, start.p, end.p);
            k_vertexScales<<<gridFor(V, BLOCK), BLOCK, 0, st>>>(pairsS.p, start.p, end.p, V, m->pos.p, m->scales.p);
            HIPCHECK(hipStreamSynchronize(st));
        } else {
            k_vertexScales<<<gridFor(V, BLOCK), BLOCK, 0, st>>>(nullptr, start.p, end.p, V, m->pos.p, m->scales.p);
        }
    }
    // parent cameras
    const bool views = view.views && view.views_channels;
    m->camerasPerSample = views ? view.views_channels : 1;
    m->cams.alloc((size_t)std::max(V, 1u) * m->camerasPerSample);
    if (V) {
        DBuf<u32> vw;
        if (views) {
            vw.alloc((size_t)P * view.views_channels);
            HIPCHECK(hipMemcpyAsync(vw.p, view.views, vw.bytes(), hipMemcpyHostToDevice, st));
        }
        k_vertexCameras<<<gridFor(V, BLOCK), BLOCK, 0, st>>>(pixSorted.p, V, views ? vw.p : nullptr,
                                                             m->camerasPerSample, view.camera_index, m->cams.p);
    }
    HIPCHECK(hipStreamSynchronize(st));
}

}  // namespace tsr

namespace {
tsr_status depthFail(tsr_status s, const char *m) {
    tsrSetLastError(m);
    return s;
}
template <typename F>
tsr_status depthGuard(F f) {
    try {
        f();
        tsrSetLastError("");
        return TSR_OK;
    } catch (const tsr_error &e) {
        tsrSetLastError(e.what());
        return e.status;
    } catch (const std::exception &e) {
        tsrSetLastError(e.what());
        return TSR_ERR_HIP;
    }
}
}  // namespace

extern "C" {

tsr_status tsr_depth_mesh_create(tsr_context *ctx, const tsr_depth_view *view, tsr_depth_mesh **out) {
    if (!ctx || !view || !out || !view->depths) return depthFail(TSR_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    return depthGuard([&] {
        HIPCHECK(hipSetDevice(ctx->device));
        tsr::PoolScope scope(ctx->pool);
        tsr_depth_mesh *m = new tsr_depth_mesh();
        m->ctx = ctx;
        m->device = ctx->device;
        try {
            depthMesh(ctx, *view, m);
        } catch (...) {
            delete m;
            throw;
        }
        *out = m;
    });
}

tsr_status tsr_depth_mesh_info(const tsr_depth_mesh *m, uint32_t *vertex_count, uint32_t *triangle_count,
                               uint32_t *cameras_per_sample) {
    if (!m) return depthFail(TSR_ERR_INVALID_ARGUMENT, "null argument");
    if (vertex_count) *vertex_count = m->vertexCount;
    if (triangle_count) *triangle_count = m->triangleCount;
    if (cameras_per_sample) *cameras_per_sample = m->camerasPerSample;
    return TSR_OK;
}

tsr_status tsr_depth_mesh_download(tsr_context *ctx, const tsr_depth_mesh *m, double *positions, double *normals,
                                   double *scales, uint32_t *triangles, uint32_t *parent_cameras,
                                   uint32_t *pixel_to_vertex) {
    if (!ctx || !m) return depthFail(TSR_ERR_INVALID_ARGUMENT, "null argument");
    return depthGuard([&] {
        HIPCHECK(hipSetDevice(ctx->device));
        hipStream_t st = ctx->stream;
        const size_t V = m->vertexCount, T = m->triangleCount;
        auto get = [&](void *dst, const void *src, size_t bytes) {
            if (dst && bytes) HIPCHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, st));
        };
        get(positions, m->pos.p, 3 * V * 8);
        get(normals, m->normals.p, 3 * V * 8);
        get(scales, m->scales.p, V * 8);
        get(triangles, m->tris.p, 3 * T * 4);
        get(parent_cameras, m->cams.p, V * m->camerasPerSample * 4);
        get(pixel_to_vertex, m->pixelToVertex.p, (size_t)m->width * m->height * 4);
        HIPCHECK(hipStreamSynchronize(st));
    });
}

void tsr_depth_mesh_destroy(tsr_depth_mesh *m) {
    if (!m) return;
    (void)hipSetDevice(m->device);
    (void)hipDeviceSynchronize();
    delete m;
}

}  // extern "C"
