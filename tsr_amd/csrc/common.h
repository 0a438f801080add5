// Definitions shared by the host and device code of the library: constants,
// the FP64 geometry frame of the octree, location codes, and the BaseProject
// stand-ins (EPSILON, plane distance, sphere/box and segment/box tests) that
// the reference's tree and occupancy code call.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define TSR_HD __host__ __device__ __forceinline__

namespace tsr {

typedef uint32_t u32;
typedef uint64_t u64;

constexpr float EPS_F = 1e-6f;                       // Math::EPSILON (BaseProject; value fixed here)
constexpr double EPS_D = (double)EPS_F;
constexpr u32 INVALID = 0xffffffffu;                 // Nodes::INVALID_INDEX (Nodes.cpp:30)
constexpr u32 MAX_DEPTH = 19;                        // deepest node level
constexpr u32 MAX_DEPTH_DIFFERENCE = 3;              // Occupancy.cpp:72
constexpr u32 FLAG_S = 1u;                           // NODE_FLAG_SAMPLENESS (Occupancy.h:38)
constexpr u32 FLAG_E = 2u;                           // NODE_FLAG_EMPTINESS (Occupancy.h:39)
constexpr u32 LEAF_BIT = 0x80000000u;                // child field of a leaf: LEAF_BIT | leaf index
constexpr double PI_D = 3.14159265358979323846;

// ---------------------------------------------------------------------------
// Location codes: loc = (1 << 3d) | morton_d(kx, ky, kz); child slot c has
// bits x = 1, y = 2, z = 4 as in Nodes::getCoords (Nodes.cpp:46-58).
// ---------------------------------------------------------------------------
TSR_HD u64 spread3(u64 v) {   // 21 bits -> every third bit
    v &= 0x1fffffull;
    v = (v | (v << 32)) & 0x1f00000000ffffull;
    v = (v | (v << 16)) & 0x1f0000ff0000ffull;
    v = (v | (v << 8)) & 0x100f00f00f00f00full;
    v = (v | (v << 4)) & 0x10c30c30c30c30c3ull;
    v = (v | (v << 2)) & 0x1249249249249249ull;
    return v;
}
TSR_HD u32 compact3(u64 v) {
    v &= 0x1249249249249249ull;
    v = (v ^ (v >> 2)) & 0x10c30c30c30c30c3ull;
    v = (v ^ (v >> 4)) & 0x100f00f00f00f00full;
    v = (v ^ (v >> 8)) & 0x1f0000ff0000ffull;
    v = (v ^ (v >> 16)) & 0x1f00000000ffffull;
    v = (v ^ (v >> 32)) & 0x1fffffull;
    return (u32)v;
}
TSR_HD u64 makeLoc(u32 kx, u32 ky, u32 kz, u32 d) {
    return (1ull << (3 * d)) | spread3(kx) | (spread3(ky) << 1) | (spread3(kz) << 2);
}
TSR_HD u32 locDepth(u64 loc) {   // position of the leading marker bit / 3
    u32 b = 63u - (u32)__builtin_clzll(loc);
#ifdef __HIP_DEVICE_COMPILE__
    // Hide the bit position's small range from the optimiser: knowing b < 64 it divides
    // with 16-bit multiplies and byte selects, and on gfx950 that form gave sporadic wrong
    // results (whole wavefronts of wrong pre-order keys, a few per million, varying from
    // run to run); the 32-bit multiply-high form has given none.
    asm volatile("" : "+v"(b));
#endif
    return b / 3u;
}
TSR_HD void locCoords(u64 loc, u32 &kx, u32 &ky, u32 &kz) {
    const u32 d = locDepth(loc);
    const u64 m = loc ^ (1ull << (3 * d));
    kx = compact3(m);
    ky = compact3(m >> 1);
    kz = compact3(m >> 2);
}
// pre-order DFS key of a node: Morton code padded to MAX_DEPTH, depth in the low bits
TSR_HD u64 dfsKey(u64 loc) {
    const u32 d = locDepth(loc);
    const u64 m = loc ^ (1ull << (3 * d));
    return (m << (3 * (MAX_DEPTH - d) + 5)) | (u64)d;
}

// ---------------------------------------------------------------------------
// FP64 frame: boundary k at depth d is B(k,d) = rootMin + k * (rootSize 2^-d),
// one rounding; centres are B(2k+1, d+1).  Identical on host and device.
// ---------------------------------------------------------------------------
struct Frame {
    double rootMin[3];
    double rootSize;
    float rootMinF[3];
    float rootSizeF;
    TSR_HD double h(u32 d) const { return ldexp(rootSize, -(int)d); }
    // k < 2^21 (grid coordinates at depth <= MAX_DEPTH + 1): a 32-bit integer, one conversion
    TSR_HD double B(int a, u32 k, u32 d) const { return rootMin[a] + (double)k * h(d); }
    TSR_HD double centre(int a, u32 k, u32 d) const { return B(a, 2 * k + 1, d + 1); }
};

// Samples::getDistanceToPlane (Samples.cpp:451-455): dot(x - p, n)
TSR_HD double planeDistance(double x, double y, double z, const double p[3], const double n[3]) {
    return (x - p[0]) * n[0] + (y - p[1]) * n[1] + (z - p[2]) * n[2];
}

// Per-unit data of the sampling refinement (Nodes.cpp:271-337): sample position,
// normal and max sampling distance, all promoted from float32.
struct SamplingUnit {
    double p[3], n[3], msd;
};

// Nodes.cpp:471-488 fittingSamplingNodeCenter for node (k, d)
TSR_HD bool fittingCentre(const Frame &f, u32 kx, u32 ky, u32 kz, u32 d, bool positive, const SamplingUnit &u) {
    const double cx = f.B(0, 2ull * kx + 1, d + 1), cy = f.B(1, 2ull * ky + 1, d + 1), cz = f.B(2, 2ull * kz + 1, d + 1);
    const double dx = cx - u.p[0], dy = cy - u.p[1], dz = cz - u.p[2];
    const double distSq = dx * dx + dy * dy + dz * dz;
    if (distSq > u.msd * u.msd) return false;
    const bool onPositive = planeDistance(cx, cy, cz, u.p, u.n) > 0.0;
    return onPositive == positive;
}

// Slab clip of segment st->en against a box, parameters in [0,1]
// (CollisionDetection::intersectAABBWithLineSegment stand-in, Nodes.cpp:319).
TSR_HD bool segmentBox(double &tIn, double &tOut, const double st[3], const double en[3], const double lo[3],
                       const double hi[3]) {
    tIn = 0.0;
    tOut = 1.0;
    for (int a = 0; a < 3; ++a) {
        const double dir = en[a] - st[a];
        if (dir == 0.0) {
            if (st[a] < lo[a] || st[a] > hi[a]) return false;
            continue;
        }
        double t1 = (lo[a] - st[a]) / dir, t2 = (hi[a] - st[a]) / dir;
        if (t1 > t2) { const double tmp = t1; t1 = t2; t2 = tmp; }
        if (t1 > tIn) tIn = t1;
        if (t2 < tOut) tOut = t2;
    }
    return tIn <= tOut;
}

// Tree::getSampleDepth restated exactly (Tree.cpp:86-91 / Nodes.cpp:250):
// smallest d with rootSize 2^-d <= max sampling distance; MAX_DEPTH+1 = too deep.
TSR_HD u32 containerDepth(const Frame &f, double msd) {
    for (u32 d = 0; d <= MAX_DEPTH; ++d)
        if (f.h(d) <= msd) return d;
    return MAX_DEPTH + 1;
}

// Integer cell of coordinate x at depth d, consistent with the FP64 boundaries:
// the largest k in [0, 2^d) with B(k, d) <= x.
TSR_HD u32 cellCoord(const Frame &f, int a, double x, u32 d) {
    const u64 lim = 1ull << d;
    double q = floor((x - f.rootMin[a]) / f.h(d));
    if (!(q >= 0.0)) q = 0.0;
    if (q > (double)(lim - 1)) q = (double)(lim - 1);
    u64 k = (u64)q;
    while (k > 0 && f.B(a, k, d) > x) --k;
    while (k + 1 < lim && f.B(a, k + 1, d) <= x) ++k;
    return (u32)k;
}

// ---------------------------------------------------------------------------
// Cone (ObliqueCircularCone stand-in) and the cone kernels, Occupancy.cpp:279-439.
//   apex a, end centre e, outside normal n of the end disc, end radius R, cut length L:
//   tComplete(x) = dot(x - a, n) / dot(e - a, n); disc at t centred a + (e - a) t, radius R t;
//   cut cone t in [tStart, 1], tStart = max(0, 1 - L/|e - a|), getLength() = min(L, |e - a|),
//   tCut = (tComplete - tStart) / (1 - tStart); degenerate iff dot(e - a, n) <= 0 or |e-a| = 0 or R <= 0.
// ---------------------------------------------------------------------------
struct ConeD {
    double a[3], v[3], n[3];
    double R, invDenom, tStart, invCut, length;
    bool degenerate;
};

TSR_HD void coneInit(ConeD &c, const double apex[3], const double endNormal[3], const double end[3], double R,
                     double L) {
    for (int i = 0; i < 3; ++i) {
        c.a[i] = apex[i];
        c.n[i] = endNormal[i];
        c.v[i] = end[i] - apex[i];
    }
    const double denom = c.v[0] * c.n[0] + c.v[1] * c.n[1] + c.v[2] * c.n[2];
    const double A = sqrt(c.v[0] * c.v[0] + c.v[1] * c.v[1] + c.v[2] * c.v[2]);
    c.R = R;
    c.degenerate = !(denom > 0.0) || !(A > 0.0) || !(R > 0.0);
    c.invDenom = 0; c.tStart = 0; c.invCut = 1; c.length = 0;
    if (c.degenerate) return;
    c.invDenom = 1.0 / denom;
    if (L >= A) {
        c.tStart = 0.0;
        c.length = A;
        c.invCut = 1.0;
    } else {
        c.tStart = 1.0 - L / A;
        c.length = L;
        c.invCut = 1.0 / (1.0 - c.tStart);
    }
}

// Occupancy.cpp:279-346 addKernel: scaled kernel at x, or 0 when rejected.  The kernel is
// the product of the ray kernel (Occupancy.cpp:429-439 computeRayKernel, a negative
// parabola 3/(2Z) (1 - z^2/Z^2) on [0, Z)) and the disc kernel (Occupancy.cpp:392-400
// computeCircularDiscKernelFromSquared, 2/(pi R^2) (1 - r^2/R^2) inside R), formed with
// the two reciprocals 1/Z and 1/R^2 instead of four quotients (a few ulp of FP64 apart).
TSR_HD double coneKernel(const ConeD &c, const double x[3], bool sampleness, double confidence) {
    const double d0 = x[0] - c.a[0], d1 = x[1] - c.a[1], d2 = x[2] - c.a[2];
    const double tc = (d0 * c.n[0] + d1 * c.n[1] + d2 * c.n[2]) * c.invDenom;
    const double tcut = (tc - c.tStart) * c.invCut;
    if (tcut < 0.0 || tcut >= 1.0) return 0.0;
    const double q0 = x[0] - (c.a[0] + c.v[0] * tc), q1 = x[1] - (c.a[1] + c.v[1] * tc),
                 q2 = x[2] - (c.a[2] + c.v[2] * tc);
    const double r2 = q0 * q0 + q1 * q1 + q2 * q2;
    const double pcdR = c.R * tc;
    const double R2 = pcdR * pcdR;
    if (r2 >= R2) return 0.0;
    double Z, z, norm;
    if (!sampleness) {
        Z = c.length;
        z = tcut * Z;
        norm = 1.0;
    } else {
        const double tc2 = fabs(tcut - 0.5) * 2.0;
        Z = 0.5 * c.length;
        z = tc2 * Z;
        norm = 0.5;
    }
    if (z < 0.0 || z >= Z) return 0.0;
    const double invZ = 1.0 / Z, invR2 = 1.0 / R2;
    const double zq = z * invZ;
    const double ray = 1.5 * invZ * (1.0 - zq * zq);
    const double disc = (2.0 / PI_D) * invR2 * (1.0 - r2 * invR2);
    const double k = ray * norm * disc;
    if (EPS_D > k) return 0.0;
    const double sk = confidence * k;
    if (EPS_D > sk) return 0.0;
    return sk;
}

// Conservative cone/box culling in FP32 (stand-in for
// ObliqueCircularCone::haveSeparatingAxis, ConeNodesChecker.cpp:65): the cut cone
// lies inside the union of balls B(a + v t, R t), t in [tStart, 1]; a box is
// culled only if its bounding ball misses that union by a safety margin.
// Besides, a point with a non-zero kernel has tComplete in [tStart, 1), so dot(x - a, n)
// lies in the slab [tStart D, D] with D = dot(e - a, n) > 0: a cube whose projection on n
// (centre +- half (|n_x| + |n_y| + |n_z|)) misses the slab holds no such point.  For a
// sampleness cone the slab is a few sample scales thick around the sample's tangent plane.
struct ConeCull {
    float a[3], vhat[3];
    float A, R, tStart, slope;   // slope = R / sqrt(A^2 - R^2), < 0 when R >= A
    float n[3], slabLo, slabHi;  // end normal and the slab of dot(x - a, n)
};

TSR_HD void cullInit(ConeCull &k, const ConeD &c) {
    const double A = sqrt(c.v[0] * c.v[0] + c.v[1] * c.v[1] + c.v[2] * c.v[2]);
    for (int i = 0; i < 3; ++i) {
        k.a[i] = (float)c.a[i];
        k.vhat[i] = (float)(c.v[i] / A);
        k.n[i] = (float)c.n[i];
    }
    k.A = (float)A;
    k.R = (float)c.R;
    k.tStart = (float)c.tStart;
    k.slope = c.R < A ? (float)(c.R / sqrt(A * A - c.R * c.R)) : -1.0f;
    const double D = c.v[0] * c.n[0] + c.v[1] * c.n[1] + c.v[2] * c.n[2];
    k.slabLo = (float)(c.tStart * D);
    k.slabHi = (float)D;
}

// true if the cube [lo, lo + size]^3 cannot meet the cone.  The square roots and the
// quotient use the hardware approximations (v_sqrt_f32, v_rcp_f32, about 1 ulp): like the
// FP32 rounding of the inputs, their error is orders of magnitude below the absolute
// margin (1e-5 of the root side), so the test stays conservative.
__device__ __forceinline__ bool cullSeparated(const ConeCull &k, float lx, float ly, float lz, float size,
                                              float absMargin) {
    const float half = 0.5f * size;
    const float b0 = lx + half - k.a[0], b1 = ly + half - k.a[1], b2 = lz + half - k.a[2];
    // the slab along the end normal (margin: absolute, plus the FP32 rounding of the
    // projections, relative 1e-6)
    const float cn = b0 * k.n[0] + b1 * k.n[1] + b2 * k.n[2];
    const float ext = half * (fabsf(k.n[0]) + fabsf(k.n[1]) + fabsf(k.n[2]));
    const float sm = absMargin + 1e-6f * (fabsf(cn) + ext + fabsf(k.slabHi));
    if (cn + ext < k.slabLo - sm || cn - ext > k.slabHi + sm) return true;
    const float u = b0 * k.vhat[0] + b1 * k.vhat[1] + b2 * k.vhat[2];
    const float p0 = b0 - u * k.vhat[0], p1 = b1 - u * k.vhat[1], p2 = b2 - u * k.vhat[2];
    const float w2 = p0 * p0 + p1 * p1 + p2 * p2;
    const float w = __builtin_amdgcn_sqrtf(w2);
    float ts = k.slope < 0.0f ? 1.0f : (u + k.slope * w) * __builtin_amdgcn_rcpf(k.A);
    ts = fminf(1.0f, fmaxf(k.tStart, ts));
    const float du = u - ts * k.A;
    const float f = __builtin_amdgcn_sqrtf(du * du + w2) - k.R * ts;
    const float rho = half * 1.7320508f * 1.01f + absMargin;
    return f > rho;
}

}  // namespace tsr
