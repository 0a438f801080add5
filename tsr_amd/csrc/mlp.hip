// Learned routing of the certified priors (no counterpart in the reference): per-leaf
// features, a 16-64-64-64-1 MLP with hand-written BF16 MFMA forward/backward kernels
// (FP32 accumulation), Adam training on the device, and the first stage of the certified
// cascade per queried leaf from the network's output.
//
// MFMA (gfx950): __builtin_amdgcn_mfma_f32_16x16x32_bf16 for the 64-wide products, one
// wavefront per 16-row tile: lane l holds A[l % 16][8 (l / 16) + i] and B[8 (l / 16) + i][l % 16]
// (i = 0..7) and D[4 (l / 16) + i][l % 16] (i = 0..3).  Products over 16 (the input layer and
// the weight gradients, whose inner dimension is the tile's 16 rows) use
// __builtin_amdgcn_mfma_f32_16x16x16bf16_1k with 4 elements per lane.  Activations are re-laid
// out between layers through LDS.
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
typedef short s16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
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
    if (K % 32 == 0) {
        for (int kc = 0; kc < K; kc += 32) {
            s16x8 a, b;
            for (int i = 0; i < 8; ++i) {
                a[i] = A[r * lda + kc + 8 * q + i];
                b[i] = W[(16 * nt + r) * K + kc + 8 * q + i];
            }
            c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b),
                                                        c, 0, 0, 0);
        }
        return c;
    }
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
    for (int kc = 0; kc < NH; kc += 32) {
        s16x8 a;
        for (int i = 0; i < 8; ++i) a[i] = toBf16(dz[r * NH + kc + 8 * q + i]);
        for (int nt = 0; nt < 4; ++nt) {
            s16x8 b;
            for (int i = 0; i < 8; ++i) b[i] = Wt[(kc + 8 * q + i) * NH + 16 * nt + r];   // W[out=k][in=n]
            acc[nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a),
                                                              __builtin_bit_cast(bf16x8, b), acc[nt], 0, 0, 0);
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

// the wave's weight gradient tiles into its own slot of the partial-gradient array (plain
// stores; k_reduceGrad adds the slots in a fixed order)
template <int NT>
__device__ void flushWeightGrad(const f32x4 (&acc)[4][NT], float *g, int ld, int lane) {
    const int r = lane & 15, q = lane >> 4;
    for (int mt = 0; mt < 4; ++mt)
        for (int nt = 0; nt < NT; ++nt)
            for (int i = 0; i < 4; ++i) g[(16 * mt + 4 * q + i) * ld + 16 * nt + r] = acc[mt][nt][i];
}

// weighted BCE loss and gradient of the rows a wave handles, into the wave's slot of
// partial (TSR_MLP_PARAM_COUNT floats) and lossPart (one double)
__global__ __launch_bounds__(64 * WAVES) void k_mlpGradient(const float *__restrict__ X, const float *__restrict__ y,
                                                            const float *__restrict__ rw, u32 n,
                                                            const float *__restrict__ w, float *__restrict__ partial,
                                                            double *__restrict__ lossPart) {
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
    const u32 slot = blockIdx.x * WAVES + wave;
    float *grad = partial + (size_t)slot * PARAMS;
    flushWeightGrad<1>(g1, grad + OFF_W1, NIN, lane);
    flushWeightGrad<4>(g2, grad + OFF_W2, NH, lane);
    flushWeightGrad<4>(g3, grad + OFF_W3, NH, lane);
    grad[OFF_B1 + lane] = gb1;
    grad[OFF_B2 + lane] = gb2;
    grad[OFF_B3 + lane] = gb3;
    grad[OFF_W4 + lane] = gw4;
    if (lane == 0) grad[OFF_B4] = gb4;
    for (int off = 32; off > 0; off >>= 1) lossAcc += __shfl_xor(lossAcc, off);
    if (lane == 0) lossPart[slot] = lossAcc;
}

// grad[i] = sum of the slots' partial gradients in slot order; loss likewise (deterministic)
__global__ void k_reduceGrad(const float *__restrict__ partial, u32 slots, const double *__restrict__ lossPart,
                             float *__restrict__ grad, double *__restrict__ loss) {
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i < PARAMS) {
        float g = 0.f;
        for (u32 s = 0; s < slots; ++s) g += partial[(size_t)s * PARAMS + i];
        grad[i] = g;
    }
    if (i == 0) {
        double l = 0.0;
        for (u32 s = 0; s < slots; ++s) l += lossPart[s];
        *loss = l;
    }
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

// gradient workspace: one partial gradient and loss per wave of the launch
constexpr u32 GRAD_BLOCKS = 512;
struct GradWork {
    DBuf<float> partial;
    DBuf<double> lossPart;
    GradWork() : partial((size_t)GRAD_BLOCKS * WAVES * PARAMS), lossPart(GRAD_BLOCKS * WAVES) {}
};

void mlpGradient(tsr_context *ctx, GradWork &work, const float *dW, const float *dX, const float *dY,
                 const float *dRw, u32 n, float *dGrad, double *dLoss) {
    const u32 tiles = (n + 15) / 16;
    const u32 grid = std::max<u32>(1, std::min<u32>((tiles + WAVES - 1) / WAVES, GRAD_BLOCKS));
    k_mlpGradient<<<grid, 64 * WAVES, mlpShared(), ctx->stream>>>(dX, dY, dRw, n, dW, work.partial.p,
                                                                  work.lossPart.p);
    k_reduceGrad<<<gridFor(PARAMS, BLOCK), BLOCK, 0, ctx->stream>>>(work.partial.p, grid * WAVES, work.lossPart.p,
                                                                    dGrad, dLoss);
}

// ---------------------------------------------------------------------------
// per-leaf features of the certified-priors router (TSR_LEAF_FEATURES)
// ---------------------------------------------------------------------------
// 0 depth; 1, 2 log E and S kernel sums; 3 E / (E + S); 4, 5 log E and S cone lengths; for
// the ancestors 1..4 levels up: 6..9 their subtree prior E / (E + S) of cone lengths (what
// the sphere query sums, seen from coarser nodes) and 10..13 the occupancy margin
// (pS - pE) / (pS + pE) at that prior; 14 the signed log distance of the occupancy at the
// 3-levels-up prior from EPSILON; 15 log10 of (E + S) / M (the scale of the occupancy).
__global__ void k_leafFeatures(const u32 *__restrict__ list, u32 n, const u32 *__restrict__ leafNode,
                               const u32 *__restrict__ parents, const uint8_t *__restrict__ nodeDepth,
                               const float *__restrict__ ksum, const float *__restrict__ clen, u32 leafCount,
                               const double2 *__restrict__ sums, double M, float *__restrict__ X) {
    const u32 i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    const u32 l = list ? list[i] : i;
    const u32 nd = leafNode[l];
    const double e = ksum[l], s = ksum[leafCount + l];
    float *x = X + (size_t)i * NIN;
    x[0] = (float)nodeDepth[nd] / (float)MAX_DEPTH;
    x[1] = log1pf((float)e) / 16.f;
    x[2] = log1pf((float)s) / 16.f;
    x[3] = e + s > 0.0 ? (float)(e / (e + s)) : 0.5f;
    x[4] = log1pf(clen[l]) / 16.f;
    x[5] = log1pf(clen[leafCount + l]) / 16.f;
    u32 a = nd;
    double v3 = 0.0;
    for (int k = 1; k <= 4; ++k) {
        const u32 pa = parents[a];
        if (pa != INVALID) a = pa;
        const double2 sm = sums[a];
        const double p = sm.x + sm.y > 0.0 ? sm.x / (sm.x + sm.y) : 0.5;
        const double pE = p * (e / M), pS = (1.0 - p) * (s / M);
        x[5 + k] = (float)p;
        x[9 + k] = pS + pE > 0.0 ? (float)((pS - pE) / (pS + pE)) : 0.f;
        if (k == 3) v3 = pS - pE;
    }
    const double r = (v3 - EPS_D) / EPS_D;
    x[14] = (float)(copysign(log1p(fabs(r)), r) / 20.0);
    x[15] = (float)(log10(fmax((e + s) / M, 1e-30)) / 10.0);
}

// first stage from the network output p: floor(2 p + 1 - t) clamped to 0..2 (the training
// target is stage / 2; a larger t keeps more queries at the cheap first stage)
__global__ void k_route(const u32 *__restrict__ queries, u32 n, const float *__restrict__ prob, float t,
                        uint8_t *__restrict__ stage, u32 *__restrict__ routed) {
    const u32 i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    const float v = floorf(2.f * prob[i] + 1.f - t);
    const u32 s = v <= 0.f ? 0u : (v >= 2.f ? 2u : (u32)v);
    stage[queries[i]] = (uint8_t)s;
    if (s) atomicAdd(&routed[s - 1], 1u);
}

}  // namespace

void leafFeatures(tsr_context *ctx, const tsr_tree *tree, const tsr_field *field, const PriorRecords &rec,
                  const u32 *list, u32 n, float *X) {
    if (!n) return;
    const double M = field->coneCount > 0.0 ? field->coneCount : 1.0;
    k_leafFeatures<<<gridFor(n, BLOCK), BLOCK, 0, ctx->stream>>>(list, n, tree->leafNode.p, tree->parents.p,
                                                                 tree->nodeDepth.p, field->ksum.p, field->clen.p,
                                                                 tree->leafCount, rec.sums(), M, X);
}

void leafRoute(tsr_context *ctx, const tsr_culler *culler, const tsr_tree *tree, const tsr_field *field,
               const PriorRecords &rec, const u32 *queries, u32 nq, uint8_t *stage, u32 *routed) {
    if (!nq) return;
    DBuf<float> X((size_t)nq * NIN), prob(nq);
    leafFeatures(ctx, tree, field, rec, queries, nq, X.p);
    mlpForward(ctx, culler->weights.p, X.p, nq, prob.p);
    k_route<<<gridFor(nq, BLOCK), BLOCK, 0, ctx->stream>>>(queries, nq, prob.p, culler->threshold, stage, routed);
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

tsr_status tsr_leaf_features(tsr_context *ctx, const tsr_tree *tree, const tsr_field *field, float *features_host) {
    if (!ctx || !tree || !field || !features_host) return mlpFail(TSR_ERR_INVALID_ARGUMENT, "null argument");
    if (field->leafCount != tree->leafCount || field->nodeCount != tree->nodeCount)
        return mlpFail(TSR_ERR_INVALID_ARGUMENT, "field does not belong to the tree");
    return mlpGuard([&] {
        HIPCHECK(hipSetDevice(ctx->device));
        const u32 L = tree->leafCount;
        if (!L) return;
        PriorRecords rec;
        priorRecords(ctx, tree, field, rec);
        DBuf<float> X((size_t)L * NIN);
        leafFeatures(ctx, tree, field, rec, nullptr, L, X.p);
        HIPCHECK(hipMemcpyAsync(features_host, X.p, X.bytes(), hipMemcpyDeviceToHost, ctx->stream));
        HIPCHECK(hipStreamSynchronize(ctx->stream));
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
        GradWork work;
        HIPCHECK(hipMemcpyAsync(w.p, weights, PARAMS * 4, hipMemcpyHostToDevice, st));
        if (n) {
            HIPCHECK(hipMemcpyAsync(x.p, features, (size_t)n * NIN * 4, hipMemcpyHostToDevice, st));
            HIPCHECK(hipMemcpyAsync(y.p, labels, (size_t)n * 4, hipMemcpyHostToDevice, st));
            HIPCHECK(hipMemcpyAsync(rw.p, row_weights, (size_t)n * 4, hipMemcpyHostToDevice, st));
        }
        mlpGradient(ctx, work, w.p, x.p, y.p, rw.p, n, g.p, l.p);
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
        GradWork work;
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
                mlpGradient(ctx, work, w.p, x.p + (size_t)b0 * NIN, y.p + b0, rw.p + b0, nb, g.p, l.p);
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
