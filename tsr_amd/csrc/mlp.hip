// Learned node culling (no counterpart in the reference): per-node features, a
// 16-64-64-64-1 MLP with hand-written BF16 MFMA forward/backward kernels (FP32
// accumulation), Adam training on the device, training labels from a plain
// field, and the culling mask used by the emptiness pass.
//
// MFMA: __builtin_amdgcn_mfma_f32_16x16x16bf16_1k, one wavefront per 16-row tile.
// Lane l holds A[l % 16][4 (l / 16) + i], B[4 (l / 16) + i][l % 16] and
// D[4 (l / 16) + i][l % 16], i = 0..3.  Activations are re-laid out between
// layers through LDS.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <random>
#include <vector>

#include "handles.h"

using namespace tsr;

struct tsr_culler {
    tsr_context *ctx = nullptr;
    int device = 0;
    DBuf<float> weights;   // TSR_MLP_PARAM_COUNT FP32 master weights
    float threshold = 0.5f;
};

namespace tsr {

namespace {

constexpr u32 BLOCK = 256;
constexpr int NIN = 16, NH = 64;
constexpr int OFF_W1 = 0, OFF_B1 = OFF_W1 + NH * NIN, OFF_W2 = OFF_B1 + NH, OFF_B2 = OFF_W2 + NH * NH,
              OFF_W3 = OFF_B2 + NH, OFF_B3 = OFF_W3 + NH * NH, OFF_W4 = OFF_B3 + NH, OFF_B4 = OFF_W4 + NH,
              PARAMS = OFF_B4 + 1;
static_assert(PARAMS == TSR_MLP_PARAM_COUNT, "parameter layout");

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// float -> bf16 (round to nearest even), as raw bits
__device__ __forceinline__ short toBf16(float f) {
    const u32 u = __float_as_uint(f);
    return (short)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float fromBf16(short b) { return __uint_as_float(((u32)(unsigned short)b) << 16); }

constexpr int WAVES = 2;   // wavefronts per block of the MLP kernels

struct MlpShared {
    short w1[NH * NIN], w2[NH * NH], w3[NH * NH];   // bf16 weights, [out][in]
    float b1[NH], b2[NH], b3[NH], w4[NH], b4;
    struct Wave {
        short x[16 * NIN];      // input tile (bf16)
        short h1[16 * NH], h2[16 * NH], h3[16 * NH];   // activations (bf16)
        float z1[16 * NH], z2[16 * NH], z3[16 * NH];   // pre-activations
        float dz[16 * NH], dzb[16 * NH];   // gradient buffers (ping-pong)
        float p[16], dz4[16];
    } wave[WAVES];
};

// C = A (16 x K) * W^T for one 16-column output tile nt: A rows in LDS (bf16,
// row stride lda), W [out][in] bf16 in LDS (row stride K)
__device__ __forceinline__ f32x4 tileAWt(const short *A, int lda, const short *W, int K, int nt, int lane) {
    f32x4 c = {0.f, 0.f, 0.f, 0.f};
    const int r = lane & 15, q = lane >> 4;
    for (int kc = 0; kc < K; kc += 16) {
        bf16x4 a, b;
        for (int i = 0; i < 4; ++i) {
            a[i] = A[r * lda + kc + 4 * q + i];
            b[i] = W[(16 * nt + r) * K + kc + 4 * q + i];
        }
        c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
    }
    return c;
}

// one hidden layer: z = A W^T + b, h = relu(z), both written row-major [16][64]
__device__ void denseRelu(const short *A, int lda, int K, const short *W, const float *b, float *z, short *h,
                          int lane) {
    const int col = lane & 15, q = lane >> 4;
    for (int nt = 0; nt < NH / 16; ++nt) {
        const f32x4 c = tileAWt(A, lda, W, K, nt, lane);
        for (int i = 0; i < 4; ++i) {
            const int row = 4 * q + i, n = 16 * nt + col;
            const float v = c[i] + b[n];
            z[row * NH + n] = v;
            h[row * NH + n] = toBf16(fmaxf(v, 0.f));
        }
    }
}

__device__ void loadWeights(MlpShared &S, const float *__restrict__ w) {
    for (int i = threadIdx.x; i < NH * NIN; i += blockDim.x) S.w1[i] = toBf16(w[OFF_W1 + i]);
    for (int i = threadIdx.x; i < NH * NH; i += blockDim.x) {
        S.w2[i] = toBf16(w[OFF_W2 + i]);
        S.w3[i] = toBf16(w[OFF_W3 + i]);
    }
    for (int i = threadIdx.x; i < NH; i += blockDim.x) {
        S.b1[i] = w[OFF_B1 + i];
        S.b2[i] = w[OFF_B2 + i];
        S.b3[i] = w[OFF_B3 + i];
        S.w4[i] = fromBf16(toBf16(w[OFF_W4 + i]));
    }
    if (threadIdx.x == 0) S.b4 = w[OFF_B4];
    __syncthreads();
}

// forward of one 16-row tile starting at row0 (rows >= n are zero); writes W.p
__device__ void forwardTile(MlpShared &S, MlpShared::Wave &W, const float *__restrict__ X, u32 row0, u32 n,
                            int lane) {
    for (int i = lane; i < 16 * NIN; i += 64) {
        const u32 row = row0 + i / NIN;
        W.x[i] = toBf16(row < n ? X[(size_t)row * NIN + i % NIN] : 0.f);
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");   // wave-local LDS exchange
    denseRelu(W.x, NIN, NIN, S.w1, S.b1, W.z1, W.h1, lane);
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    denseRelu(W.h1, NH, NH, S.w2, S.b2, W.z2, W.h2, lane);
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    denseRelu(W.h2, NH, NH, S.w3, S.b3, W.z3, W.h3, lane);
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    // output layer (64 -> 1): 4 lanes per row, 16 products each, fixed-order combine
    const int row = lane >> 2, part = lane & 3;
    float acc = 0.f;
    for (int k = 16 * part; k < 16 * part + 16; ++k) acc += fromBf16(W.h3[row * NH + k]) * S.w4[k];
    acc += __shfl_xor(acc, 1);
    acc += __shfl_xor(acc, 2);
    if (part == 0) W.p[row] = 1.f / (1.f + expf(-(acc + S.b4)));
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
}

__global__ __launch_bounds__(64 * WAVES) void k_mlpForward(const float *__restrict__ X, u32 n,
                                                           const float *__restrict__ w, float *__restrict__ prob) {
    extern __shared__ char smem[];
    MlpShared &S = *reinterpret_cast<MlpShared *>(smem);
    loadWeights(S, w);
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    MlpShared::Wave &W = S.wave[wave];
    const u32 tiles = (n + 15) / 16;
    for (u32 t = blockIdx.x * WAVES + wave; t < tiles; t += gridDim.x * WAVES) {
        forwardTile(S, W, X, 16 * t, n, lane);
        if (lane < 16 && 16 * t + lane < n) prob[16 * t + lane] = W.p[lane];
    }
}

// dh = dz (16 x 64, fp32 in LDS) * W (W [out][in] bf16): the input gradient of a
// 64 -> 64 layer, masked by relu'(z) of the layer below; result to dzOut (fp32)
__device__ void backDense(const float *dz, const short *Wt, const float *zBelow, float *dzOut, int lane) {
    const int r = lane & 15, q = lane >> 4;
    f32x4 acc[4];
    for (int nt = 0; nt < 4; ++nt) acc[nt] = (f32x4){0.f, 0.f, 0.f, 0.f};
    for (int kc = 0; kc < NH; kc += 16) {
        bf16x4 a;
        for (int i = 0; i < 4; ++i) a[i] = toBf16(dz[r * NH + kc + 4 * q + i]);
        for (int nt = 0; nt < 4; ++nt) {
            bf16x4 b;
            for (int i = 0; i < 4; ++i) b[i] = Wt[(kc + 4 * q + i) * NH + 16 * nt + r];   // W[out=k][in=n]
            acc[nt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[nt], 0, 0, 0);
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    for (int nt = 0; nt < 4; ++nt)
        for (int i = 0; i < 4; ++i) {
            const int row = 4 * q + i, col = 16 * nt + r;
            dzOut[row * NH + col] = zBelow[row * NH + col] > 0.f ? acc[nt][i] : 0.f;
        }
}

// acc[mt][nt] += dz^T (64 x 16 rows) * H (16 rows x K): weight gradient tiles
template <int NT>
__device__ void accumWeightGrad(f32x4 (&acc)[4][NT], const float *dz, const short *H, int ldh, int lane) {
    const int r = lane & 15, q = lane >> 4;
    for (int mt = 0; mt < 4; ++mt) {
        bf16x4 a;
        for (int i = 0; i < 4; ++i) a[i] = toBf16(dz[(4 * q + i) * NH + 16 * mt + r]);   // A[m=out][k=row]
        for (int nt = 0; nt < NT; ++nt) {
            bf16x4 b;
            for (int i = 0; i < 4; ++i) b[i] = H[(4 * q + i) * ldh + 16 * nt + r];            // B[k=row][n=in]
            acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[mt][nt], 0, 0, 0);
        }
    }
}

template <int NT>
__device__ void flushWeightGrad(const f32x4 (&acc)[4][NT], float *g, int ld, int lane) {
    const int r = lane & 15, q = lane >> 4;
    for (int mt = 0; mt < 4; ++mt)
        for (int nt = 0; nt < NT; ++nt)
            for (int i = 0; i < 4; ++i) atomicAdd(&g[(16 * mt + 4 * q + i) * ld + 16 * nt + r], acc[mt][nt][i]);
}

// weighted BCE loss and gradient, summed over all rows into grad / loss
__global__ __launch_bounds__(64 * WAVES) void k_mlpGradient(const float *__restrict__ X, const float *__restrict__ y,
                                                            const float *__restrict__ rw, u32 n,
                                                            const float *__restrict__ w, float *__restrict__ grad,
                                                            double *__restrict__ loss) {
    extern __shared__ char smem[];
    MlpShared &S = *reinterpret_cast<MlpShared *>(smem);
    loadWeights(S, w);
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    MlpShared::Wave &W = S.wave[wave];
    f32x4 g1[4][1], g2[4][4], g3[4][4];
    for (int a = 0; a < 4; ++a) {
        g1[a][0] = (f32x4){0.f, 0.f, 0.f, 0.f};
        for (int b = 0; b < 4; ++b) g2[a][b] = g3[a][b] = (f32x4){0.f, 0.f, 0.f, 0.f};
    }
    float gb1 = 0.f, gb2 = 0.f, gb3 = 0.f, gw4 = 0.f, gb4 = 0.f;   // lane = hidden unit
    double lossAcc = 0.0;
    const u32 tiles = (n + 15) / 16;
    for (u32 t = blockIdx.x * WAVES + wave; t < tiles; t += gridDim.x * WAVES) {
        forwardTile(S, W, X, 16 * t, n, lane);
        if (lane < 16) {
            const u32 row = 16 * t + lane;
            float d = 0.f;
            if (row < n) {
                const float p = W.p[lane], yy = y[row], ww = rw[row];
                d = ww * (p - yy);   // d loss / d z4 for sigmoid + BCE
                const float pc = fminf(fmaxf(p, 1e-7f), 1.f - 1e-7f);
                lossAcc += (double)(-ww * (yy * logf(pc) + (1.f - yy) * logf(1.f - pc)));
            }
            W.dz4[lane] = d;
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        // output layer: dW4[k] = sum_r dz4[r] h3[r][k]; dz3 = dz4 w4 relu'(z3)
        {
            const int k = lane;
            float s = 0.f;
            for (int r = 0; r < 16; ++r) s += W.dz4[r] * fromBf16(W.h3[r * NH + k]);
            gw4 += s;
            if (lane == 0)
                for (int r = 0; r < 16; ++r) gb4 += W.dz4[r];
            for (int r = 0; r < 16; ++r) W.dz[r * NH + k] = W.z3[r * NH + k] > 0.f ? W.dz4[r] * S.w4[k] : 0.f;
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        {   // layer 3: dW3 += dz3^T h2, db3, dz2 = (dz3 W3) relu'(z2)
            float s = 0.f;
            for (int r = 0; r < 16; ++r) s += W.dz[r * NH + lane];
            gb3 += s;
            accumWeightGrad<4>(g3, W.dz, W.h2, NH, lane);
            backDense(W.dz, S.w3, W.z2, W.dzb, lane);
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        {   // layer 2
            float s = 0.f;
            for (int r = 0; r < 16; ++r) s += W.dzb[r * NH + lane];
            gb2 += s;
            accumWeightGrad<4>(g2, W.dzb, W.h1, NH, lane);
            backDense(W.dzb, S.w2, W.z1, W.dz, lane);
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        {   // layer 1 (input gradient not needed)
            float s = 0.f;
            for (int r = 0; r < 16; ++r) s += W.dz[r * NH + lane];
            gb1 += s;
            accumWeightGrad<1>(g1, W.dz, W.x, NIN, lane);
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    }
    flushWeightGrad<1>(g1, grad + OFF_W1, NIN, lane);
    flushWeightGrad<4>(g2, grad + OFF_W2, NH, lane);
    flushWeightGrad<4>(g3, grad + OFF_W3, NH, lane);
    atomicAdd(&grad[OFF_B1 + lane], gb1);
    atomicAdd(&grad[OFF_B2 + lane], gb2);
    atomicAdd(&grad[OFF_B3 + lane], gb3);
    atomicAdd(&grad[OFF_W4 + lane], gw4);
    if (lane == 0) atomicAdd(&grad[OFF_B4], gb4);
    for (int off = 32; off > 0; off >>= 1) lossAcc += __shfl_xor(lossAcc, off);
    if (lane == 0) atomicAdd(loss, lossAcc);
}

// Adam (Kingma & Ba, 2015) on the FP32 master weights; grad holds the batch sum
__global__ void k_adam(float *__restrict__ w, float *__restrict__ m, float *__restrict__ v,
                       const float *__restrict__ grad, float scale, float lr, float b1t, float b2t) {
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= PARAMS) return;
    const float g = grad[i] * scale;
    const float mi = 0.9f * m[i] + 0.1f * g, vi = 0.999f * v[i] + 0.001f * g * g;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / (1.f - b1t)) / (sqrtf(vi / (1.f - b2t)) + 1e-8f);
}

size_t mlpShared() { return sizeof(MlpShared); }

void mlpForward(tsr_context *ctx, const float *dW, const float *dX, u32 n, float *dP) {
    if (!n) return;
    const u32 tiles = (n + 15) / 16;
    const u32 grid = std::min<u32>((tiles + WAVES - 1) / WAVES, 4096);
    k_mlpForward<<<grid, 64 * WAVES, mlpShared(), ctx->stream>>>(dX, n, dW, dP);
}

void mlpGradient(tsr_context *ctx, const float *dW, const float *dX, const float *dY, const float *dRw, u32 n,
                 float *dGrad, double *dLoss) {
    HIPCHECK(hipMemsetAsync(dGrad, 0, PARAMS * 4, ctx->stream));
    HIPCHECK(hipMemsetAsync(dLoss, 0, 8, ctx->stream));
    if (!n) return;
    const u32 tiles = (n + 15) / 16;
    const u32 grid = std::min<u32>((tiles + WAVES - 1) / WAVES, 1024);
    k_mlpGradient<<<grid, 64 * WAVES, mlpShared(), ctx->stream>>>(dX, dY, dRw, n, dW, dGrad, dLoss);
}

// ---------------------------------------------------------------------------
// per-node features (TSR_NODE_FEATURES) from the tree, the sampleness pass
// and a 1/16-subsampled emptiness pass
// ---------------------------------------------------------------------------
constexpr int AGG = 14;   // cnt, s, s2, nx, ny, nz, px, py, pz, Shits, Ehits, leaves, Sleaves, Sksum

__global__ void k_ownAgg(const u32 *__restrict__ counts, const u32 *__restrict__ starts, u32 nodeCount,
                         const float *__restrict__ px, const float *__restrict__ py, const float *__restrict__ pz,
                         const float *__restrict__ nx, const float *__restrict__ ny, const float *__restrict__ nz,
                         const float *__restrict__ scale, float *__restrict__ agg) {
    const u32 nd = blockIdx.x * BLOCK + threadIdx.x;
    if (nd >= nodeCount) return;
    float a[AGG];
    for (int i = 0; i < AGG; ++i) a[i] = 0.f;
    const u32 c = counts[nd], s0 = starts[nd];
    for (u32 j = s0; j < s0 + c; ++j) {
        a[0] += 1.f;
        a[1] += scale[j];
        a[2] += scale[j] * scale[j];
        a[3] += nx[j]; a[4] += ny[j]; a[5] += nz[j];
        a[6] += px[j]; a[7] += py[j]; a[8] += pz[j];
    }
    for (int i = 0; i < AGG; ++i) agg[(size_t)nd * AGG + i] = a[i];
}

__global__ void k_leafAgg(const u32 *__restrict__ leafNode, u32 leafCount, const u32 *__restrict__ hitsS,
                          const u32 *__restrict__ hitsE, const float *__restrict__ ksumS,
                          const u32 *__restrict__ states, float *__restrict__ agg) {
    const u32 l = blockIdx.x * BLOCK + threadIdx.x;
    if (l >= leafCount) return;
    const u32 nd = leafNode[l];
    float *a = agg + (size_t)nd * AGG;
    a[9] = (float)hitsS[l];
    a[10] = (float)hitsE[l];
    a[11] = 1.f;
    a[12] = (states[nd] & FLAG_S) ? 1.f : 0.f;
    a[13] = ksumS[l];
}

__global__ void k_aggLevel(const u32 *__restrict__ child, const uint8_t *__restrict__ depth, u32 nodeCount, u32 level,
                           float *__restrict__ agg) {
    const u32 nd = blockIdx.x * BLOCK + threadIdx.x;
    if (nd >= nodeCount || depth[nd] != level) return;
    const u32 c = child[nd];
    if (c & LEAF_BIT) return;
    float *a = agg + (size_t)nd * AGG;
    for (u32 s = 0; s < 8; ++s)
        for (int i = 0; i < AGG; ++i) a[i] += agg[(size_t)(c + s) * AGG + i];
}

__global__ void k_features(Frame f, const u64 *__restrict__ nodeLoc, const u32 *__restrict__ parents,
                           const u32 *__restrict__ states, const float *__restrict__ agg, u32 nodeCount,
                           float *__restrict__ feat) {
    const u32 nd = blockIdx.x * BLOCK + threadIdx.x;
    if (nd >= nodeCount) return;
    const float *a = agg + (size_t)nd * AGG;
    const u64 loc = nodeLoc[nd];
    const u32 d = locDepth(loc);
    u32 k[3];
    locCoords(loc, k[0], k[1], k[2]);
    const float h = (float)f.h(d);
    const float cnt = a[0], inv = cnt > 0.f ? 1.f / cnt : 0.f;
    const float ms = a[1] * inv / h, ms2 = a[2] * inv / (h * h);
    const float mn[3] = {a[3] * inv, a[4] * inv, a[5] * inv};
    const float nlen = sqrtf(mn[0] * mn[0] + mn[1] * mn[1] + mn[2] * mn[2]);
    float dist = 0.f;
    if (cnt > 0.f) {
        float d2 = 0.f;
        for (int q = 0; q < 3; ++q) {
            const float c = (float)f.centre(q, k[q], d), pm = a[6 + q] * inv;
            d2 += (pm - c) * (pm - c);
        }
        dist = sqrtf(d2) / h;
    }
    const u32 par = parents[nd];
    const float *pa = agg + (size_t)(par == INVALID ? nd : par) * AGG;
    const float ps = pa[9], pe = 16.f * pa[10];
    float *o = feat + (size_t)nd * TSR_NODE_FEATURES;
    o[0] = (float)d / (float)MAX_DEPTH;
    o[1] = log1pf(cnt);
    o[2] = ms;
    o[3] = fmaxf(ms2 - ms * ms, 0.f);
    o[4] = mn[0]; o[5] = mn[1]; o[6] = mn[2];
    o[7] = cnt > 0.f ? 1.f - nlen : 0.f;
    o[8] = log1pf(a[9]);
    o[9] = log1pf(16.f * a[10]);
    o[10] = (ps - pe) / (ps + pe + 1.f);
    o[11] = log1pf(a[13]) / 30.f;
    o[12] = a[11] > 0.f ? a[12] / a[11] : 0.f;
    o[13] = log2f(fmaxf(a[11], 1.f)) / 20.f;
    o[14] = dist;
    o[15] = (states[nd] & FLAG_S) ? 1.f : 0.f;
}

// training labels: leaves with positive occupancy or at a sign-changing dual cell
__global__ void k_labelCells(const u32 *__restrict__ cells, u32 cellCount, const double *__restrict__ occ,
                             uint8_t *__restrict__ leafLabel) {
    const u32 c = blockIdx.x * BLOCK + threadIdx.x;
    if (c >= cellCount) return;
    bool pos = false, neg = false;
    for (int i = 0; i < 8; ++i) {
        const double v = occ[cells[8 * (size_t)c + i]];
        pos |= v > 0.0;
        neg |= !(v > 0.0);
    }
    if (pos && neg)
        for (int i = 0; i < 8; ++i) leafLabel[cells[8 * (size_t)c + i]] = 1;
}

__global__ void k_labelNodesLeaves(const u32 *__restrict__ leafNode, u32 leafCount, const double *__restrict__ occ,
                                   const uint8_t *__restrict__ leafLabel, uint8_t *__restrict__ nodeLabel) {
    const u32 l = blockIdx.x * BLOCK + threadIdx.x;
    if (l >= leafCount) return;
    nodeLabel[leafNode[l]] = (leafLabel[l] || occ[l] > 0.0) ? 1 : 0;
}

__global__ void k_labelLevel(const u32 *__restrict__ child, const uint8_t *__restrict__ depth, u32 nodeCount,
                             u32 level, uint8_t *__restrict__ nodeLabel) {
    const u32 nd = blockIdx.x * BLOCK + threadIdx.x;
    if (nd >= nodeCount || depth[nd] != level) return;
    const u32 c = child[nd];
    if (c & LEAF_BIT) return;
    uint8_t v = 0;
    for (u32 s = 0; s < 8; ++s) v |= nodeLabel[c + s];
    nodeLabel[nd] = v;
}

__global__ void k_cullMask(const float *__restrict__ prob, u32 nodeCount, const uint8_t *__restrict__ depth,
                           u32 minDepth, float threshold, uint8_t *__restrict__ mask) {
    const u32 nd = blockIdx.x * BLOCK + threadIdx.x;
    if (nd >= nodeCount) return;
    mask[nd] = (depth[nd] >= minDepth && prob[nd] < threshold) ? 1 : 0;
}

}  // namespace

void nodeFeatures(tsr_context *ctx, const tsr_tree *tree, const tsr_scene *scene, const tsr_params &prm,
                  const u32 *states, const u32 *hitsS, const float *ksumS, DBuf<float> &feat) {
    hipStream_t st = ctx->stream;
    const u32 N = tree->nodeCount, L = tree->leafCount;
    DBuf<u32> hitsE(L);
    emptinessSubsample(ctx, tree, scene, prm, states, 16, hitsE.p);
    DBuf<float> agg((size_t)N * AGG);
    k_ownAgg<<<gridFor(N, BLOCK), BLOCK, 0, st>>>(tree->sampleCounts.p, tree->sampleStarts.p, N, tree->px.p,
                                                  tree->py.p, tree->pz.p, tree->nx.p, tree->ny.p, tree->nz.p,
                                                  tree->scale.p, agg.p);
    k_leafAgg<<<gridFor(L, BLOCK), BLOCK, 0, st>>>(tree->leafNode.p, L, hitsS, hitsE.p, ksumS, states, agg.p);
    for (int lvl = (int)tree->maxDepth - 1; lvl >= 0; --lvl)
        k_aggLevel<<<gridFor(N, BLOCK), BLOCK, 0, st>>>(tree->child.p, tree->nodeDepth.p, N, (u32)lvl, agg.p);
    feat.alloc((size_t)N * TSR_NODE_FEATURES);
    k_features<<<gridFor(N, BLOCK), BLOCK, 0, st>>>(tree->frame, tree->nodeLoc.p, tree->parents.p, states, agg.p, N,
                                                    feat.p);
}

void cullerMask(tsr_context *ctx, const tsr_culler *culler, const tsr_tree *tree, const tsr_scene *scene,
                const tsr_params &prm, const u32 *states, const u32 *hitsS, const float *ksumS,
                DBuf<uint8_t> &mask) {
    DBuf<float> feat;
    nodeFeatures(ctx, tree, scene, prm, states, hitsS, ksumS, feat);
    const u32 N = tree->nodeCount;
    DBuf<float> prob(N);
    mlpForward(ctx, culler->weights.p, feat.p, N, prob.p);
    mask.alloc(N);
    k_cullMask<<<gridFor(N, BLOCK), BLOCK, 0, ctx->stream>>>(prob.p, N, tree->nodeDepth.p, 2, culler->threshold,
                                                             mask.p);
}

}  // namespace tsr

namespace {
tsr_status mlpFail(tsr_status s, const char *m) {
    tsrSetLastError(m);
    return s;
}
template <typename F>
tsr_status mlpGuard(F f) {
    try {
        f();
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

tsr_status tsr_culler_create(tsr_context *ctx, const float *weights, float threshold, tsr_culler **out) {
    if (!ctx || !weights || !out) return mlpFail(TSR_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    return mlpGuard([&] {
        HIPCHECK(hipSetDevice(ctx->device));
        tsr_culler *c = new tsr_culler();
        c->ctx = ctx;
        c->device = ctx->device;
        c->threshold = threshold;
        c->weights.alloc(PARAMS);
        HIPCHECK(hipMemcpy(c->weights.p, weights, PARAMS * 4, hipMemcpyHostToDevice));
        *out = c;
    });
}

tsr_status tsr_culler_set_weights(tsr_context *ctx, tsr_culler *c, const float *weights, float threshold) {
    if (!ctx || !c || !weights) return mlpFail(TSR_ERR_INVALID_ARGUMENT, "null argument");
    return mlpGuard([&] {
        HIPCHECK(hipSetDevice(ctx->device));
        HIPCHECK(hipMemcpy(c->weights.p, weights, PARAMS * 4, hipMemcpyHostToDevice));
        c->threshold = threshold;
    });
}

void tsr_culler_destroy(tsr_culler *c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    (void)hipDeviceSynchronize();
    delete c;
}

tsr_status tsr_node_features(tsr_context *ctx, const tsr_tree *tree, const tsr_scene *scene, const tsr_field *field,
                             float *features_host) {
    if (!ctx || !tree || !scene || !field || !features_host) return mlpFail(TSR_ERR_INVALID_ARGUMENT, "null argument");
    if (field->leafCount != tree->leafCount || !field->hitsS.p)
        return mlpFail(TSR_ERR_INVALID_ARGUMENT, "field does not belong to the tree");
    return mlpGuard([&] {
        HIPCHECK(hipSetDevice(ctx->device));
        tsr_params prm = field->params;
        DBuf<float> feat;
        nodeFeatures(ctx, tree, scene, prm, field->states.p, field->hitsS.p, field->ksum.p + tree->leafCount, feat);
        HIPCHECK(hipMemcpyAsync(features_host, feat.p, feat.bytes(), hipMemcpyDeviceToHost, ctx->stream));
        HIPCHECK(hipStreamSynchronize(ctx->stream));
    });
}

tsr_status tsr_node_labels(tsr_context *ctx, const tsr_tree *tree, const tsr_field *field, uint8_t *labels_host) {
    if (!ctx || !tree || !field || !labels_host) return mlpFail(TSR_ERR_INVALID_ARGUMENT, "null argument");
    if (field->leafCount != tree->leafCount) return mlpFail(TSR_ERR_INVALID_ARGUMENT, "field does not belong to the tree");
    return mlpGuard([&] {
        HIPCHECK(hipSetDevice(ctx->device));
        hipStream_t st = ctx->stream;
        const u32 N = tree->nodeCount, L = tree->leafCount, C = tree->cellCount;
        DBuf<uint8_t> leafLabel(L), nodeLabel(N);
        HIPCHECK(hipMemsetAsync(leafLabel.p, 0, L, st));
        HIPCHECK(hipMemsetAsync(nodeLabel.p, 0, N, st));
        if (C) k_labelCells<<<gridFor(C, BLOCK), BLOCK, 0, st>>>(tree->cells.p, C, field->occ.p, leafLabel.p);
        k_labelNodesLeaves<<<gridFor(L, BLOCK), BLOCK, 0, st>>>(tree->leafNode.p, L, field->occ.p, leafLabel.p,
                                                                nodeLabel.p);
        for (int lvl = (int)tree->maxDepth - 1; lvl >= 0; --lvl)
            k_labelLevel<<<gridFor(N, BLOCK), BLOCK, 0, st>>>(tree->child.p, tree->nodeDepth.p, N, (u32)lvl,
                                                              nodeLabel.p);
        HIPCHECK(hipMemcpyAsync(labels_host, nodeLabel.p, N, hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
    });
}

tsr_status tsr_mlp_forward(tsr_context *ctx, const float *weights, const float *features, uint32_t n, float *probs) {
    if (!ctx || !weights || (n && (!features || !probs))) return mlpFail(TSR_ERR_INVALID_ARGUMENT, "null argument");
    return mlpGuard([&] {
        HIPCHECK(hipSetDevice(ctx->device));
        hipStream_t st = ctx->stream;
        DBuf<float> w(PARAMS), x((size_t)n * NIN + 1), p(n + 1);
        HIPCHECK(hipMemcpyAsync(w.p, weights, PARAMS * 4, hipMemcpyHostToDevice, st));
        if (n) HIPCHECK(hipMemcpyAsync(x.p, features, (size_t)n * NIN * 4, hipMemcpyHostToDevice, st));
        mlpForward(ctx, w.p, x.p, n, p.p);
        if (n) HIPCHECK(hipMemcpyAsync(probs, p.p, (size_t)n * 4, hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
    });
}

tsr_status tsr_mlp_gradient(tsr_context *ctx, const float *weights, const float *features, const float *labels,
                            const float *row_weights, uint32_t n, float *grad, double *loss) {
    if (!ctx || !weights || !grad || !loss || (n && (!features || !labels || !row_weights)))
        return mlpFail(TSR_ERR_INVALID_ARGUMENT, "null argument");
    return mlpGuard([&] {
        HIPCHECK(hipSetDevice(ctx->device));
        hipStream_t st = ctx->stream;
        DBuf<float> w(PARAMS), x((size_t)n * NIN + 1), y(n + 1), rw(n + 1), g(PARAMS);
        DBuf<double> l(1);
        HIPCHECK(hipMemcpyAsync(w.p, weights, PARAMS * 4, hipMemcpyHostToDevice, st));
        if (n) {
            HIPCHECK(hipMemcpyAsync(x.p, features, (size_t)n * NIN * 4, hipMemcpyHostToDevice, st));
            HIPCHECK(hipMemcpyAsync(y.p, labels, (size_t)n * 4, hipMemcpyHostToDevice, st));
            HIPCHECK(hipMemcpyAsync(rw.p, row_weights, (size_t)n * 4, hipMemcpyHostToDevice, st));
        }
        mlpGradient(ctx, w.p, x.p, y.p, rw.p, n, g.p, l.p);
        HIPCHECK(hipMemcpyAsync(grad, g.p, PARAMS * 4, hipMemcpyDeviceToHost, st));
        HIPCHECK(hipMemcpyAsync(loss, l.p, 8, hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
    });
}

tsr_status tsr_mlp_train(tsr_context *ctx, float *weights, const float *features, const float *labels,
                         const float *row_weights, uint32_t n, uint32_t epochs, uint32_t batch, float learning_rate,
                         uint64_t seed, float *loss_per_epoch) {
    if (!ctx || !weights || !features || !labels || !row_weights || n == 0 || batch == 0)
        return mlpFail(TSR_ERR_INVALID_ARGUMENT, "bad arguments");
    return mlpGuard([&] {
        HIPCHECK(hipSetDevice(ctx->device));
        hipStream_t st = ctx->stream;
        // shuffled copy of the data set, one permutation per epoch (host, seeded)
        std::vector<u32> perm(n);
        std::iota(perm.begin(), perm.end(), 0u);
        std::mt19937_64 rng(seed);
        std::vector<float> xs((size_t)n * NIN), ys(n), ws(n);
        DBuf<float> w(PARAMS), m(PARAMS), v(PARAMS), g(PARAMS), x((size_t)n * NIN), y(n), rw(n);
        DBuf<double> l(1);
        HIPCHECK(hipMemcpyAsync(w.p, weights, PARAMS * 4, hipMemcpyHostToDevice, st));
        HIPCHECK(hipMemsetAsync(m.p, 0, PARAMS * 4, st));
        HIPCHECK(hipMemsetAsync(v.p, 0, PARAMS * 4, st));
        int step = 0;
        for (u32 e = 0; e < epochs; ++e) {
            std::shuffle(perm.begin(), perm.end(), rng);
            for (u32 i = 0; i < n; ++i) {
                std::memcpy(&xs[(size_t)i * NIN], &features[(size_t)perm[i] * NIN], NIN * 4);
                ys[i] = labels[perm[i]];
                ws[i] = row_weights[perm[i]];
            }
            HIPCHECK(hipMemcpyAsync(x.p, xs.data(), xs.size() * 4, hipMemcpyHostToDevice, st));
            HIPCHECK(hipMemcpyAsync(y.p, ys.data(), (size_t)n * 4, hipMemcpyHostToDevice, st));
            HIPCHECK(hipMemcpyAsync(rw.p, ws.data(), (size_t)n * 4, hipMemcpyHostToDevice, st));
            double epochLoss = 0.0, epochWeight = 0.0;
            for (u32 b0 = 0; b0 < n; b0 += batch) {
                const u32 nb = std::min(batch, n - b0);
                mlpGradient(ctx, w.p, x.p + (size_t)b0 * NIN, y.p + b0, rw.p + b0, nb, g.p, l.p);
                double wsum = 0.0;
                for (u32 i = b0; i < b0 + nb; ++i) wsum += ws[i];
                ++step;
                const float b1t = powf(0.9f, (float)step), b2t = powf(0.999f, (float)step);
                k_adam<<<gridFor(PARAMS, BLOCK), BLOCK, 0, st>>>(w.p, m.p, v.p, g.p, (float)(1.0 / std::max(wsum, 1e-12)),
                                                                 learning_rate, b1t, b2t);
                epochLoss += readScalar(l.p, st);
                epochWeight += wsum;
            }
            if (loss_per_epoch) loss_per_epoch[e] = (float)(epochLoss / std::max(epochWeight, 1e-12));
        }
        HIPCHECK(hipMemcpyAsync(weights, w.p, PARAMS * 4, hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
    });
}

}  // extern "C"
