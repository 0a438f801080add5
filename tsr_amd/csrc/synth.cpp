// Seeded synthetic scene generator for the occupancy + DMC path.
//
// Replaces the reference's Embree-rendered SyntheticScene
// (SR/Scene/SyntheticScene.cpp:45-85,173-214,305-336) with analytic ray casting
// of spheres and oriented boxes.  Per pixel of each pinhole camera (up = +y,
// looking at the scene centre; horizontal field of view 60 degrees for the
// sphere scene, fitted to the scene's bounding sphere for the others):
//   depth d of the first analytic hit, noisy depth d' = d (1 + sigma g),
//   position = camera + dir d', normal = ground-truth normal at the noiseless hit,
//   scale = pixel footprint d' 2 tan(fov/2) / W, confidence = 1,
//   one camera link to the rendering camera (Scene.cpp:564-572 fallback).
// Outliers replace the position by a uniform point in the scene box and the
// normal by a random unit vector (scale copied from the pixel).
// Random numbers: Philox4x32-10 (Salmon et al., SC'11), counter = (pixel,
// camera, stream, 0), key = seed; Gaussians by Box-Muller.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "tsr_capi_internal.h"

namespace {

typedef uint32_t u32;
typedef uint64_t u64;

struct Philox {
    static void round(u32 ctr[4], const u32 key[2]) {
        const u64 p0 = (u64)0xD2511F53u * ctr[0], p1 = (u64)0xCD9E8D57u * ctr[2];
        const u32 hi0 = (u32)(p0 >> 32), lo0 = (u32)p0, hi1 = (u32)(p1 >> 32), lo1 = (u32)p1;
        const u32 n0 = hi1 ^ ctr[1] ^ key[0], n1 = lo1, n2 = hi0 ^ ctr[3] ^ key[1], n3 = lo0;
        ctr[0] = n0; ctr[1] = n1; ctr[2] = n2; ctr[3] = n3;
    }
    static void gen(u32 out[4], u32 c0, u32 c1, u32 c2, u32 c3, u64 seed) {
        u32 ctr[4] = {c0, c1, c2, c3};
        u32 key[2] = {(u32)seed, (u32)(seed >> 32)};
        for (int r = 0; r < 10; ++r) {
            round(ctr, key);
            key[0] += 0x9E3779B9u;
            key[1] += 0xBB67AE85u;
        }
        std::memcpy(out, ctr, sizeof(ctr));
    }
    static double uniform(u32 v) { return ((double)v + 0.5) * (1.0 / 4294967296.0); }
};

struct Vec { double x, y, z; };
inline Vec operator+(Vec a, Vec b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline Vec operator-(Vec a, Vec b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline Vec operator*(Vec a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline double dot(Vec a, Vec b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline Vec cross(Vec a, Vec b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline Vec normalize(Vec a) { return a * (1.0 / std::sqrt(dot(a, a))); }

struct Prim {
    int type;      // 0 sphere, 1 oriented box
    Vec c;         // centre
    double r;      // sphere radius
    Vec ax[3];     // box axes (orthonormal)
    double he[3];  // box half extents
};

bool hitSphere(const Prim &p, Vec o, Vec d, double &t, Vec &n) {
    const Vec oc = o - p.c;
    const double b = dot(oc, d), c = dot(oc, oc) - p.r * p.r;
    const double disc = b * b - c;
    if (disc < 0) return false;
    const double s = std::sqrt(disc);
    double t0 = -b - s;
    if (t0 <= 1e-9) t0 = -b + s;
    if (t0 <= 1e-9) return false;
    t = t0;
    n = normalize((o + d * t) - p.c);
    if (dot(n, d) > 0) n = n * -1.0;
    return true;
}

bool hitBox(const Prim &p, Vec o, Vec d, double &t, Vec &n) {
    const Vec oc = o - p.c;
    double tin = -1e300, tout = 1e300;
    int inAxis = -1;
    double inSign = 0;
    for (int a = 0; a < 3; ++a) {
        const double oa = dot(oc, p.ax[a]), da = dot(d, p.ax[a]);
        if (std::fabs(da) < 1e-300) {
            if (oa < -p.he[a] || oa > p.he[a]) return false;
            continue;
        }
        double t1 = (-p.he[a] - oa) / da, t2 = (p.he[a] - oa) / da;
        double s1 = -1.0;
        if (t1 > t2) { std::swap(t1, t2); s1 = 1.0; }
        if (t1 > tin) { tin = t1; inAxis = a; inSign = s1; }
        if (t2 < tout) tout = t2;
    }
    if (tin > tout || tout <= 1e-9 || tin <= 1e-9 || inAxis < 0) return false;
    t = tin;
    n = p.ax[inAxis] * inSign;
    return true;
}

struct Camera { Vec pos, fwd, right, up; double tanHalf; };

Camera lookAt(Vec pos, Vec target) {
    Camera c;
    c.pos = pos;
    c.fwd = normalize(target - pos);
    Vec up = {0, 1, 0};
    if (std::fabs(dot(up, c.fwd)) > 0.99) up = {0, 0, 1};
    c.right = normalize(cross(c.fwd, up));
    c.up = cross(c.right, c.fwd);
    c.tanHalf = std::tan(M_PI / 6.0);   // 60 degree horizontal field of view
    return c;
}

Vec randomUnit(u32 a, u32 b) {
    const double z = 2.0 * Philox::uniform(a) - 1.0;
    const double phi = 2.0 * M_PI * Philox::uniform(b);
    const double s = std::sqrt(std::max(0.0, 1.0 - z * z));
    return {s * std::cos(phi), s * std::sin(phi), z};
}

void rotationFrom(Vec ax[3], u32 r0, u32 r1, u32 r2) {
    // random rotation: random unit axis u, random in-plane angle
    const Vec u = randomUnit(r0, r1);
    Vec tmp = std::fabs(u.x) < 0.9 ? Vec{1, 0, 0} : Vec{0, 1, 0};
    Vec v = normalize(cross(u, tmp));
    Vec w = cross(u, v);
    const double ang = 2.0 * M_PI * Philox::uniform(r2);
    const double c = std::cos(ang), s = std::sin(ang);
    ax[0] = u;
    ax[1] = v * c + w * s;
    ax[2] = w * c - v * s;
}

struct SceneSpec {
    std::vector<Prim> prims;
    std::vector<Camera> cams;
    int W = 0, H = 0;
    double sigma = 0, outlierRate = 0;
    Vec boxMin{0, 0, 0}, boxMax{0, 0, 0};
};

void primBox(const Prim &p, Vec &mn, Vec &mx) {
    if (p.type == 0) {
        mn = p.c - Vec{p.r, p.r, p.r};
        mx = p.c + Vec{p.r, p.r, p.r};
        return;
    }
    Vec e{0, 0, 0};
    for (int a = 0; a < 3; ++a) {
        e.x += std::fabs(p.ax[a].x) * p.he[a];
        e.y += std::fabs(p.ax[a].y) * p.he[a];
        e.z += std::fabs(p.ax[a].z) * p.he[a];
    }
    mn = p.c - e;
    mx = p.c + e;
}

Prim makeBox(Vec c, const Vec ax[3], double h0, double h1, double h2) {
    Prim p{};
    p.type = 1; p.c = c;
    for (int a = 0; a < 3; ++a) p.ax[a] = ax[a];
    p.he[0] = h0; p.he[1] = h1; p.he[2] = h2;
    return p;
}

Prim makeSphere(Vec c, double r) {
    Prim p{};
    p.type = 0; p.c = c; p.r = r;
    return p;
}

// Cameras at random directions and distances U[dmin,dmax] around `centre`;
// the horizontal field of view is fitted so that the sphere of radius
// `sceneRadius` spans the image width (multi-scale sampling of one scene).
void primBox(const Prim &p, Vec &lo, Vec &hi);

// Cameras look at `centre` from distances U[dmin, dmax]; the field of view fits a
// sphere of `sceneRadius` (90 degrees for a camera inside 1.1 x that radius).  With
// clearance > 0 a camera position closer than `clearance` to any primitive's box
// is drawn again (next counter), so that no depth map starts right at a surface.
void randomCameras(SceneSpec &sc, int count, double dmin, double dmax, Vec centre, double sceneRadius, u64 seed,
                   double clearance = 0.0) {
    for (int k = 0; k < count; ++k) {
        for (u32 attempt = 0;; ++attempt) {
            u32 r[4];
            Philox::gen(r, (u32)k, 0xCA3u, 7u + 1000u * attempt, 0u, seed);
            const Vec dir = randomUnit(r[0], r[1]);
            const double dist = dmin + (dmax - dmin) * Philox::uniform(r[2]);
            const Vec pos = centre + dir * dist;
            bool clear = true;
            for (size_t q = 0; clear && clearance > 0.0 && q < sc.prims.size(); ++q) {
                Vec a, b;
                primBox(sc.prims[q], a, b);
                const double dx = std::max({a.x - pos.x, 0.0, pos.x - b.x}), dy = std::max({a.y - pos.y, 0.0, pos.y - b.y}),
                             dz = std::max({a.z - pos.z, 0.0, pos.z - b.z});
                clear = dx * dx + dy * dy + dz * dz >= clearance * clearance;
            }
            if (!clear && attempt < 1000) continue;
            Camera c = lookAt(pos, centre);
            c.tanHalf = dist > 1.1 * sceneRadius
                            ? sceneRadius / std::sqrt(dist * dist - sceneRadius * sceneRadius)
                            : 1.0;
            sc.cams.push_back(c);
            break;
        }
    }
}

void randomPrimitives(SceneSpec &sc, int count, double lo, double hi, double sizeScale, u64 seed) {
    for (int k = 0; k < count; ++k) {
        u32 r[4], q[4];
        Philox::gen(r, (u32)k, 0x9817u, 11u, 0u, seed);
        Philox::gen(q, (u32)k, 0x9817u, 12u, 0u, seed);
        const Vec c{lo + (hi - lo) * Philox::uniform(r[0]), lo + (hi - lo) * Philox::uniform(r[1]),
                    lo + (hi - lo) * Philox::uniform(r[2])};
        const int type = (int)(r[3] % 3u);
        Vec ax[3];
        rotationFrom(ax, q[0], q[1], q[2]);
        if (type == 0) {
            sc.prims.push_back(makeSphere(c, sizeScale * (0.1 + 0.4 * Philox::uniform(q[3]))));
        } else if (type == 1) {
            u32 s[4];
            Philox::gen(s, (u32)k, 0x9817u, 13u, 0u, seed);
            sc.prims.push_back(makeBox(c, ax, sizeScale * (0.1 + 0.3 * Philox::uniform(s[0])),
                                       sizeScale * (0.1 + 0.3 * Philox::uniform(s[1])),
                                       sizeScale * (0.1 + 0.3 * Philox::uniform(s[2]))));
        } else {
            u32 s[4];
            Philox::gen(s, (u32)k, 0x9817u, 14u, 0u, seed);
            const double thick = 0.005 + 0.015 * Philox::uniform(s[2]);
            sc.prims.push_back(makeBox(c, ax, sizeScale * (0.15 + 0.35 * Philox::uniform(s[0])),
                                       sizeScale * (0.15 + 0.35 * Philox::uniform(s[1])), 0.5 * thick));
        }
    }
}

SceneSpec makeScene(int config, u64 seed, double resScale) {
    SceneSpec sc;
    auto res = [&](int w, int h) {
        sc.W = std::max(4, (int)std::lround(w * resScale));
        sc.H = std::max(3, (int)std::lround(h * resScale));
    };
    switch (config) {
    case 0: {  // unit sphere, 12 cameras on the y = 0 ring of radius 3
        sc.prims.push_back(makeSphere({0, 0, 0}, 1.0));
        for (int k = 0; k < 12; ++k) {
            const double a = k * M_PI / 6.0;
            sc.cams.push_back(lookAt({3.0 * std::cos(a), 0.0, 3.0 * std::sin(a)}, {0, 0, 0}));
        }
        res(128, 96);
        sc.sigma = 0.005;
        break;
    }
    case 1:
    case 2: {  // thin wedge (10 degree dihedral, sheets 1x1x0.01) plus a sphere r = 0.4
        for (int s = 0; s < 2; ++s) {
            const double ang = (s == 0 ? -5.0 : 5.0) * M_PI / 180.0;
            Vec ax[3] = {{std::cos(ang), std::sin(ang), 0}, {-std::sin(ang), std::cos(ang), 0}, {0, 0, 1}};
            sc.prims.push_back(makeBox(ax[0] * 0.5, ax, 0.5, 0.005, 0.5));
        }
        sc.prims.push_back(makeSphere({-0.55, 0.0, 0.0}, 0.4));
        randomCameras(sc, 64, 2.0, 8.0, {0, 0, 0}, 0.8, seed);
        res(640, 480);
        sc.sigma = 0.003;
        sc.outlierRate = 0.01;
        break;
    }
    case 3: {  // random union of 3-8 primitives, 32 cameras
        u32 r[4];
        Philox::gen(r, 0, 0x3333u, 1u, 0u, seed);
        randomPrimitives(sc, 3 + (int)(r[0] % 6u), -0.8, 0.8, 1.0, seed);
        randomCameras(sc, 32, 2.0, 8.0, {0, 0, 0}, 1.2, seed);
        res(320, 240);
        sc.sigma = 0.003;
        sc.outlierRate = 0.01;
        break;
    }
    case 4: {  // 30 random primitives in [-4,4]^3, 120 cameras at 3..20
        randomPrimitives(sc, 30, -3.0, 3.0, 2.0, seed);
        randomCameras(sc, 120, 3.0, 20.0, {0, 0, 0}, 4.5, seed, 0.5);
        res(1024, 768);
        sc.sigma = 0.003;
        sc.outlierRate = 0.02;
        break;
    }
    default:
        throw tsr_error(TSR_ERR_INVALID_ARGUMENT, "unknown synthetic scene config");
    }
    Vec mn{1e300, 1e300, 1e300}, mx{-1e300, -1e300, -1e300};
    for (const Prim &p : sc.prims) {
        Vec a, b;
        primBox(p, a, b);
        mn = {std::min(mn.x, a.x), std::min(mn.y, a.y), std::min(mn.z, a.z)};
        mx = {std::max(mx.x, b.x), std::max(mx.y, b.y), std::max(mx.z, b.z)};
    }
    sc.boxMin = mn;
    sc.boxMax = mx;
    return sc;
}

}  // namespace

tsr_synthetic *tsr_synth_build(int config, uint64_t seed, double resScale) {
    SceneSpec sc = makeScene(config, seed, resScale);
    tsr_synthetic *out = new tsr_synthetic();
    const int W = sc.W, H = sc.H, K = (int)sc.cams.size();
    const double aspect = (double)H / (double)W;
    for (int k = 0; k < K; ++k) {
        out->cx.push_back((float)sc.cams[k].pos.x);
        out->cy.push_back((float)sc.cams[k].pos.y);
        out->cz.push_back((float)sc.cams[k].pos.z);
    }
    // per camera, per pixel row: collect hits in order (camera, row, column)
    std::vector<std::vector<float>> rows((size_t)K * H);
    std::vector<std::vector<u32>> rowCam((size_t)K * H);
#pragma omp parallel for schedule(dynamic, 4)
    for (long long kr = 0; kr < (long long)K * H; ++kr) {
        const int k = (int)(kr / H), y = (int)(kr % H);
        const Camera &cam = sc.cams[k];
        const double tanHalf = cam.tanHalf;
        std::vector<float> &buf = rows[kr];
        for (int x = 0; x < W; ++x) {
            const double u = (2.0 * (x + 0.5) / W - 1.0) * tanHalf;
            const double v = (1.0 - 2.0 * (y + 0.5) / H) * tanHalf * aspect;
            const Vec dir = normalize(cam.fwd + cam.right * u + cam.up * v);
            double best = 1e300;
            Vec bn{0, 0, 0};
            for (const Prim &p : sc.prims) {
                double t;
                Vec n;
                const bool hit = p.type == 0 ? hitSphere(p, cam.pos, dir, t, n) : hitBox(p, cam.pos, dir, t, n);
                if (hit && t < best) { best = t; bn = n; }
            }
            if (best >= 1e299) continue;
            const u32 pix = (u32)(y * W + x);
            u32 r[4];
            Philox::gen(r, pix, (u32)k, 1u, 0u, seed);
            const double g = std::sqrt(-2.0 * std::log(Philox::uniform(r[0]))) * std::cos(2.0 * M_PI * Philox::uniform(r[1]));
            const double dn = best * (1.0 + sc.sigma * g);
            Vec pos = cam.pos + dir * dn;
            Vec nrm = bn;
            const double scale = dn * 2.0 * tanHalf / W;
            if (sc.outlierRate > 0 && Philox::uniform(r[2]) < sc.outlierRate) {
                u32 q[4];
                Philox::gen(q, pix, (u32)k, 2u, 0u, seed);
                pos = {sc.boxMin.x + (sc.boxMax.x - sc.boxMin.x) * Philox::uniform(q[0]),
                       sc.boxMin.y + (sc.boxMax.y - sc.boxMin.y) * Philox::uniform(q[1]),
                       sc.boxMin.z + (sc.boxMax.z - sc.boxMin.z) * Philox::uniform(q[2])};
                nrm = randomUnit(q[3], r[3]);
            }
            const float rec[7] = {(float)pos.x, (float)pos.y, (float)pos.z, (float)nrm.x, (float)nrm.y, (float)nrm.z,
                                  (float)scale};
            buf.insert(buf.end(), rec, rec + 7);
        }
    }
    size_t n = 0;
    for (auto &r : rows) n += r.size() / 7;
    out->px.resize(n); out->py.resize(n); out->pz.resize(n);
    out->nx.resize(n); out->ny.resize(n); out->nz.resize(n);
    out->scale.resize(n); out->conf.assign(n, 1.0f); out->cams.resize(n);
    size_t o = 0;
    for (size_t kr = 0; kr < rows.size(); ++kr) {
        const std::vector<float> &b = rows[kr];
        for (size_t j = 0; j < b.size(); j += 7, ++o) {
            out->px[o] = b[j]; out->py[o] = b[j + 1]; out->pz[o] = b[j + 2];
            out->nx[o] = b[j + 3]; out->ny[o] = b[j + 4]; out->nz[o] = b[j + 5];
            out->scale[o] = b[j + 6];
            out->cams[o] = (u32)(kr / H);
        }
    }
    out->width = W;
    out->height = H;
    return out;
}
