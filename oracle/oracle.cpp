// oracle.cpp — CPU oracle for the LuminaGI tracer path. TEST INFRASTRUCTURE ONLY.
//
// This file is the checker, never the product: only tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg may load liblgi_oracle.so. It restates, in plain C++,
// the reference algorithms (paths relative to the reference repository):
//   BuildBVH                       src/bvh.cpp:11-211 (stable tie order, see orc_bvh_build)
//   ray_aabb / ray_tri             src/pathtracer.cu:61-107
//   traverse_closest / _occluded   src/pathtracer.cu:109-191
//   xorshift32 / randf             src/pathtracer.cu:193-204
//   onb / sample_cos_hemi          src/pathtracer.cu:206-228
//   shade_direct                   src/pathtracer.cu:230-272
//   primary_ray / make_ray         src/pathtracer.cu:284-293, include/vecmath.h:25-44
//   trace_path                     src/pathtracer.cu:295-331
//   accumulate_kernel seeding      src/pathtracer.cu:345-364, chunking :699-704
//   gbuffer_kernel                 src/pathtracer.cu:366-411
//   tonemap_kernel                 src/pathtracer.cu:413-429
//   overwrite_camera               src/main.cpp:29-59
// Deviations chosen by design and shared with the GPU path:
//   * cosf/sinf are replaced by a Cody-Waite reduction + Cephes minimax polynomials
//     (orc_sincos) so CPU and GPU produce identical bits;
//   * the default RNG is PCG32 (O'Neill, pcg-random.org) with one stream per
//     (pixel, sample); the reference xorshift32 chunked seeding is kept as a policy;
//   * per-pixel sample sums are Kahan-compensated FP32 (the reference uses a plain
//     FP32 sum per 8-sample chunk); orc_render also returns an FP64 sum.
// Built with -ffp-contract=off so that every a*b+c rounds twice, as on the GPU build.
//
// Parity status of each function (checks live in tests/test_oracle_*.py):
//   orc_morton3       pinned: brute-force bit interleave over all 1024 inputs per axis
//   orc_bvh_build     pinned: golden node arrays from the compiled reference bvh.cpp,
//                     brute-force invariants, hand-worked tiny cases
//   orc_pcg32_*       pinned: PCG32 published known-answer vector (seed 42, stream 54)
//   orc_sincos        pinned: libm sin/cos within 2 ulp over a dense sweep
//   orc_render        pinned: brute-force closest hit (all triangles) on small scenes,
//                     closed forms (sun on an unoccluded plane, point-light falloff),
//                     statistical cos-hemisphere moments, FP64 sum agreement
//   orc_gbuffer/tonemap pinned: closed-form encodings
//   trace_path bounces pinned: closed interior of a sphere lit by a point light at its centre
//                     (every hit has the same direct term D, so the 2-bounce mean is
//                     D (1 + rho + rho^2) exactly when Russian roulette is unbiased)
//   traverse_occluded pinned: shadow footprint of an opaque plate over a ground plane
//
// Planted mistakes: building with -DORC_MUTANT=k (k > 0) plants one known mistake so that the
// tests can show the closed-form pins reject it (1: Russian roulette without the 1/q weight,
// 2: sun shadow test inverted, 3: bounce origin offset along -N, 4: G-buffer normal not turned
// towards the camera). The library the tests load is built with ORC_MUTANT = 0.
#ifndef ORC_MUTANT
#define ORC_MUTANT 0
#endif
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#if defined(__FP_FAST_FMA) || defined(__FAST_MATH__)
#error "oracle must be built without fast-math / FP contraction"
#endif

extern "C" {

struct OTri {          // reference Triangle layout (include/scene.h:39-45), 100 B
    float v[3][3];
    float n[3][3];
    float uv[3][2];
    int32_t material;
};
struct ONode {         // reference BvhNode layout (include/bvh.h:8-16), 40 B
    float bbMin[3];
    float bbMax[3];
    int32_t left, right, firstTri, triCount;
};
struct OPoint { float pos[3]; float color[3]; float intensity; float radius; };
struct OSun { float dir[3]; float color[3]; float intensity; };
struct OCam {          // lgi_camera layout
    float pos[3];
    float c2w[16];
    float fovYDeg, aspect, nearZ, farZ;
    int32_t width, height;
};

}  // extern "C"

static_assert(sizeof(OTri) == 100, "Triangle layout");
static_assert(sizeof(ONode) == 40, "BvhNode layout");

namespace {

// ---------------------------------------------------------------- vector math
// include/vecmath.h:9-44
struct V3 { float x, y, z; };
inline V3 v3(float x, float y, float z) { return {x, y, z}; }
inline V3 vfrom(const float p[3]) { return {p[0], p[1], p[2]}; }
inline V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 mul(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 had(V3 a, V3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline float len(V3 a) { return std::sqrt(dot(a, a)); }
inline V3 normalize(V3 a) {
    float l = len(a);
    return (l > 1e-20f) ? mul(a, 1.0f / l) : v3(0, 0, 1);
}
struct Ray { V3 o, d, inv; };
inline Ray make_ray(V3 o, V3 d) {
    Ray r;
    r.o = o;
    r.d = normalize(d);
    r.inv = {1.0f / r.d.x, 1.0f / r.d.y, 1.0f / r.d.z};
    return r;
}

// ---------------------------------------------------------------- sin/cos
// Cody-Waite reduction by pi/2 (three-part constant, exact first product for the small
// quadrant counts used here) followed by the Cephes single-precision minimax polynomials
// for sin and cos on [-pi/4, pi/4] (S. Moshier, Cephes Math Library, sinf.c).
// Every operation is a separately rounded IEEE op; the GPU path evaluates the same
// expression tree.
const float kPio2Hi = 1.5703125f;                 // 8 significant bits: k*hi exact for k < 2^16
const float kPio2Mid = 4.83751296997070312e-04f;  // next bits of pi/2
const float kPio2Lo = 7.54978995489188216e-08f;   // remaining bits of pi/2
const float kTwoOverPi = 0.636619772367581343f;

inline void own_sincos(float x, float* s, float* c) {
    float kf = std::nearbyint(x * kTwoOverPi);   // round-half-even, as rintf on the GPU
    int k = (int)kf;
    float r = ((x - kf * kPio2Hi) - kf * kPio2Mid) - kf * kPio2Lo;
    float z = r * r;
    float ps = ((-1.9515295891e-4f * z + 8.3321608736e-3f) * z - 1.6666654611e-1f) * z * r + r;
    float pc = ((2.443315711809948e-5f * z - 1.388731625493765e-3f) * z + 4.166664568298827e-2f) * z * z
               - 0.5f * z + 1.0f;
    switch (k & 3) {
        case 0: *s = ps; *c = pc; break;
        case 1: *s = pc; *c = -ps; break;
        case 2: *s = -ps; *c = -pc; break;
        default: *s = -pc; *c = ps; break;
    }
}

// ---------------------------------------------------------------- RNGs
// xorshift32 / randf: src/pathtracer.cu:193-204
inline uint32_t xorshift32(uint32_t& s) {
    s ^= s << 13;
    s ^= s >> 17;
    s ^= s << 5;
    return s;
}
// PCG32 (XSH-RR 64/32), M. O'Neill, "PCG: A Family of Simple Fast Space-Efficient
// Statistically Good Algorithms for Random Number Generation" (2014), pcg_basic.c.
const uint64_t kPcgMult = 6364136223846793005ULL;
struct Pcg { uint64_t state, inc; };
inline uint32_t pcg_next(Pcg& g) {
    uint64_t old = g.state;
    g.state = old * kPcgMult + g.inc;
    uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xs >> rot) | (xs << ((0u - rot) & 31u));
}
inline void pcg_seed(Pcg& g, uint64_t initstate, uint64_t initseq) {
    g.state = 0u;
    g.inc = (initseq << 1u) | 1u;
    pcg_next(g);
    g.state += initstate;
    pcg_next(g);
}
// SplitMix64 finaliser (Steele, Lea, Flood 2014) used to decorrelate stream seeds.
inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

// Per-path random source: either PCG32 seeded per (pixel, sample) or the reference's
// xorshift32 state that continues across the samples of one 8-sample chunk.
struct Rng {
    int pcg;
    Pcg g;
    uint32_t xs;
    inline float next() {
        uint32_t x = pcg ? pcg_next(g) : xorshift32(xs);
        return (float)(x >> 8) * (1.0f / 16777216.0f);
    }
};
inline void pcg_for_sample(Pcg& g, uint32_t seed, uint32_t pixel, uint32_t sample) {
    uint64_t stream = ((uint64_t)seed << 32) | pixel;
    pcg_seed(g, mix64(stream ^ ((uint64_t)sample << 40) ^ 0x9E3779B97F4A7C15ULL), stream);
}

// ---------------------------------------------------------------- scene view
struct Counters {
    uint64_t node_visits = 0, tri_tests = 0, closest_hits = 0, rays = 0, shadow_rays = 0;
    void add(const Counters& o) {
        node_visits += o.node_visits; tri_tests += o.tri_tests; closest_hits += o.closest_hits;
        rays += o.rays; shadow_rays += o.shadow_rays;
    }
};

// Diffuse textures: RGBA8 texels of all textures one after another, per material
// {texel offset, width, height, textured}, and the 8-bit sRGB -> linear table.
struct TexView {
    const uint32_t* texels = nullptr;
    const int32_t* desc = nullptr;
    const float* srgb = nullptr;
};

struct SceneView {
    const ONode* nodes;
    const OTri* tris;            // permuted (leaf) order
    const float* matAlbedo;      // 3 per material
    const OPoint* points;
    int numPoints;
    OSun sun;
    TexView tex;
};

struct CamView {
    V3 pos, fwd, left, up;
    float tanHalfFov, aspect;
    int width, height;
};

// src/pathtracer.cu:61-80
inline bool ray_aabb(const Ray& r, const float mn[3], const float mx[3], float tmax_in) {
    float t0 = (mn[0] - r.o.x) * r.inv.x;
    float t1 = (mx[0] - r.o.x) * r.inv.x;
    float lo = std::fmin(t0, t1), hi = std::fmax(t0, t1);
    t0 = (mn[1] - r.o.y) * r.inv.y;
    t1 = (mx[1] - r.o.y) * r.inv.y;
    lo = std::fmax(lo, std::fmin(t0, t1));
    hi = std::fmin(hi, std::fmax(t0, t1));
    t0 = (mn[2] - r.o.z) * r.inv.z;
    t1 = (mx[2] - r.o.z) * r.inv.z;
    lo = std::fmax(lo, std::fmin(t0, t1));
    hi = std::fmin(hi, std::fmax(t0, t1));
    if (hi < std::fmax(lo, 0.0f) || lo > tmax_in) return false;
    return true;
}

// src/pathtracer.cu:82-107 (Moller-Trumbore)
inline bool ray_tri(const Ray& r, const OTri& t, float tmax, float& outT, float& outU, float& outV) {
    const float kEps = 1e-6f;
    V3 v0 = vfrom(t.v[0]);
    V3 e1 = sub(vfrom(t.v[1]), v0);
    V3 e2 = sub(vfrom(t.v[2]), v0);
    V3 p = cross(r.d, e2);
    float det = dot(e1, p);
    if (det > -kEps && det < kEps) return false;
    float invDet = 1.0f / det;
    V3 s = sub(r.o, v0);
    float u = dot(s, p) * invDet;
    if (u < 0.0f || u > 1.0f) return false;
    V3 q = cross(s, e1);
    float v = dot(r.d, q) * invDet;
    if (v < 0.0f || u + v > 1.0f) return false;
    float tt = dot(e2, q) * invDet;
    if (tt <= kEps || tt >= tmax) return false;
    outT = tt; outU = u; outV = v;
    return true;
}

// src/pathtracer.cu:109-160. Stack of 64, push left then right (right popped first).
inline int traverse_closest(const SceneView& s, const Ray& r, float tmax, float& outT, V3& outN,
                            float& outTU, float& outTV, Counters& c) {
    int stack[64];
    int sp = 0;
    stack[sp++] = 0;
    float bestT = tmax;
    int bestTri = -1;
    float bestU = 0, bestV = 0;
    c.rays++;
    while (sp > 0) {
        int ni = stack[--sp];
        const ONode& n = s.nodes[ni];
        c.node_visits++;
        if (!ray_aabb(r, n.bbMin, n.bbMax, bestT)) continue;
        if (n.triCount > 0) {
            for (int i = 0; i < n.triCount; ++i) {
                int ti = n.firstTri + i;
                float tt, uu, vv;
                c.tri_tests++;
                if (ray_tri(r, s.tris[ti], bestT, tt, uu, vv)) {
                    bestT = tt; bestTri = ti; bestU = uu; bestV = vv;
                }
            }
        } else if (sp + 2 <= 64) {
            stack[sp++] = n.left;
            stack[sp++] = n.right;
        }
    }
    if (bestTri < 0) return -1;
    c.closest_hits++;
    const OTri& t = s.tris[bestTri];
    float w = 1.0f - bestU - bestV;
    outN = normalize(v3(w * t.n[0][0] + bestU * t.n[1][0] + bestV * t.n[2][0],
                        w * t.n[0][1] + bestU * t.n[1][1] + bestV * t.n[2][1],
                        w * t.n[0][2] + bestU * t.n[1][2] + bestV * t.n[2][2]));
    outT = bestT;
    outTU = w * t.uv[0][0] + bestU * t.uv[1][0] + bestV * t.uv[2][0];   // src/pathtracer.cu:156-157
    outTV = w * t.uv[0][1] + bestU * t.uv[1][1] + bestV * t.uv[2][1];
    return bestTri;
}

// src/pathtracer.cu:162-191
inline bool traverse_occluded(const SceneView& s, const Ray& r, float tmax, Counters& c) {
    int stack[64];
    int sp = 0;
    stack[sp++] = 0;
    c.shadow_rays++;
    while (sp > 0) {
        int ni = stack[--sp];
        const ONode& n = s.nodes[ni];
        c.node_visits++;
        if (!ray_aabb(r, n.bbMin, n.bbMax, tmax)) continue;
        if (n.triCount > 0) {
            for (int i = 0; i < n.triCount; ++i) {
                float tt, uu, vv;
                c.tri_tests++;
                if (ray_tri(r, s.tris[n.firstTri + i], tmax, tt, uu, vv)) return true;
            }
        } else if (sp + 2 <= 64) {
            stack[sp++] = n.left;
            stack[sp++] = n.right;
        }
    }
    return false;
}

// src/pathtracer.cu:206-214 (Frisvad)
inline void onb(V3 N, V3& T, V3& B) {
    if (N.z < -0.99999f) { T = v3(0, -1, 0); B = v3(-1, 0, 0); return; }
    float a = 1.0f / (1.0f + N.z);
    float b = -N.x * N.y * a;
    T = v3(1.0f - N.x * N.x * a, b, -N.x);
    B = v3(b, 1.0f - N.y * N.y * a, -N.y);
}

// src/pathtracer.cu:216-228 (own sin/cos, see above)
inline V3 sample_cos_hemi(V3 N, Rng& rng) {
    float u1 = rng.next();
    float u2 = rng.next();
    float r = std::sqrt(u1);
    float phi = 2.0f * 3.14159265358979f * u2;
    float sn, cs;
    own_sincos(phi, &sn, &cs);
    float x = r * cs;
    float y = r * sn;
    float z = std::sqrt(std::fmax(0.0f, 1.0f - u1));
    V3 T, B;
    onb(N, T, B);
    return normalize(add(add(mul(T, x), mul(B, y)), mul(N, z)));
}

// src/pathtracer.cu:230-272
inline V3 shade_direct(const SceneView& s, V3 P, V3 N, V3 albedo, Counters& c) {
    const float kShadowBias = 1e-3f;
    V3 color = v3(0, 0, 0);
    {
        V3 toLight = normalize(mul(vfrom(s.sun.dir), -1.0f));
        float ndotl = std::fmax(0.0f, dot(N, toLight));
        if (ndotl > 0.0f) {
            Ray shadow = make_ray(add(P, mul(N, kShadowBias)), toLight);
            if (traverse_occluded(s, shadow, 1e30f, c) == (ORC_MUTANT == 2)) {
                float I = s.sun.intensity * ndotl;
                color = add(color, had(albedo, mul(vfrom(s.sun.color), I)));
            }
        }
    }
    for (int i = 0; i < s.numPoints; ++i) {
        const OPoint& L = s.points[i];
        V3 toLightVec = sub(vfrom(L.pos), P);
        float dist = len(toLightVec);
        V3 toLight = mul(toLightVec, 1.0f / dist);
        float ndotl = std::fmax(0.0f, dot(N, toLight));
        if (ndotl <= 0.0f || dist >= L.radius * 3.0f) continue;
        Ray shadow = make_ray(add(P, mul(N, kShadowBias)), toLight);
        if (traverse_occluded(s, shadow, dist - kShadowBias, c)) continue;
        float r = L.radius;
        float atten = std::fmax(0.0f, 1.0f - dist / r);
        atten = atten * atten;
        float I = L.intensity * ndotl * atten;
        color = add(color, had(albedo, mul(vfrom(L.color), I)));
    }
    return color;
}

// src/pathtracer.cu:284-293
inline Ray primary_ray(const CamView& cam, int px, int py, float jx, float jy) {
    float ndcx = (2.0f * (px + jx) / cam.width) - 1.0f;
    float ndcy = 1.0f - (2.0f * (py + jy) / cam.height);
    float rx = ndcx * cam.aspect * cam.tanHalfFov;
    float ry = ndcy * cam.tanHalfFov;
    V3 dirCam = v3(1.0f, -rx, ry);
    V3 dirWorld = add(add(mul(cam.fwd, dirCam.x), mul(cam.left, dirCam.y)), mul(cam.up, dirCam.z));
    return make_ray(cam.pos, dirWorld);
}

// sample_albedo, src/pathtracer.cu:274-282, with the texture object of :563-576: normalised
// coordinates (u, 1 - v) wrapped into [0, 1), texel centres at (i + 0.5) / size, RGBA8 texels
// converted from sRGB to linear before bilinear filtering, and filter weights in 1/256 steps
// (rounded to nearest), as a texture unit with 8-bit fractional weights computes them.
inline float tex_weight(float f) { return std::floor(f * 256.0f + 0.5f) * (1.0f / 256.0f); }

inline V3 sample_albedo(const SceneView& s, int mat, float tu, float tv) {
    if (!s.tex.desc || !s.tex.desc[4 * mat + 3]) return vfrom(&s.matAlbedo[3 * mat]);
    const uint32_t* tex = s.tex.texels + s.tex.desc[4 * mat];
    const int W = s.tex.desc[4 * mat + 1], H = s.tex.desc[4 * mat + 2];
    float u = tu - std::floor(tu);
    float v = (1.0f - tv) - std::floor(1.0f - tv);
    if (!(u >= 0.0f && u < 1.0f)) u = 0.0f;
    if (!(v >= 0.0f && v < 1.0f)) v = 0.0f;
    const float x = u * (float)W - 0.5f, y = v * (float)H - 0.5f;
    const float xf = std::floor(x), yf = std::floor(y);
    const float a = tex_weight(x - xf), b = tex_weight(y - yf);
    int x0 = (int)xf, y0 = (int)yf;
    if (x0 < 0) x0 += W;
    if (y0 < 0) y0 += H;
    x0 = std::min(std::max(x0, 0), W - 1);
    y0 = std::min(std::max(y0, 0), H - 1);
    const int x1 = x0 + 1 == W ? 0 : x0 + 1, y1 = y0 + 1 == H ? 0 : y0 + 1;
    auto texel = [&](int xx, int yy) {
        const uint32_t p = tex[(size_t)yy * W + xx];
        return v3(s.tex.srgb[p & 255u], s.tex.srgb[(p >> 8) & 255u], s.tex.srgb[(p >> 16) & 255u]);
    };
    const V3 top = add(mul(texel(x0, y0), 1.0f - a), mul(texel(x1, y0), a));
    const V3 bot = add(mul(texel(x0, y1), 1.0f - a), mul(texel(x1, y1), a));
    return add(mul(top, 1.0f - b), mul(bot, b));
}

// src/pathtracer.cu:295-331. no_rr reproduces the sorted path's semantics (:486-547).
inline V3 trace_path(const SceneView& s, Ray ray, int maxBounces, Rng& rng, bool no_rr, Counters& c) {
    V3 L = v3(0, 0, 0);
    V3 throughput = v3(1, 1, 1);
    for (int bounce = 0; bounce <= maxBounces; ++bounce) {
        float hitT;
        V3 N;
        float tu, tv;
        int tri = traverse_closest(s, ray, 1e30f, hitT, N, tu, tv, c);
        if (tri < 0) break;
        if (dot(N, ray.d) > 0.0f) N = mul(N, -1.0f);
        V3 P = add(ray.o, mul(ray.d, hitT));
        V3 albedo = sample_albedo(s, s.tris[tri].material, tu, tv);
        L = add(L, had(throughput, shade_direct(s, P, N, albedo, c)));
        if (bounce == maxBounces) break;
        if (bounce >= 1 && !no_rr) {
            float q = std::fmin(0.95f, std::fmax(throughput.x, std::fmax(throughput.y, throughput.z)));
            if (rng.next() > q) break;
            if (ORC_MUTANT != 1) throughput = mul(throughput, 1.0f / q);
        }
        V3 newDir = sample_cos_hemi(N, rng);
        ray = make_ray(add(P, mul(N, ORC_MUTANT == 3 ? -1e-3f : 1e-3f)), newDir);
        throughput = had(throughput, albedo);
    }
    return L;
}

CamView make_cam(const OCam& c, float tanHalfFov) {
    CamView v;
    v.pos = vfrom(c.pos);
    v.fwd = v3(c.c2w[0], c.c2w[1], c.c2w[2]);
    v.left = v3(c.c2w[4], c.c2w[5], c.c2w[6]);
    v.up = v3(c.c2w[8], c.c2w[9], c.c2w[10]);
    v.tanHalfFov = tanHalfFov;
    v.aspect = c.aspect;
    v.width = c.width;
    v.height = c.height;
    return v;
}

inline void kahan_add(float& sum, float& comp, float x) {
    float y = x - comp;
    float t = sum + y;
    comp = (t - sum) - y;
    sum = t;
}

// ---------------------------------------------------------------- BVH build (src/bvh.cpp)
const int kLeafMaxTris = 4;  // src/bvh.cpp:11

uint32_t spread3(uint32_t v) {  // src/bvh.cpp:14-21
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}
uint32_t quant10(float f) {  // src/bvh.cpp:25-30
    f = f * 1024.0f;
    if (f < 0.0f) f = 0.0f;
    if (f > 1023.0f) f = 1023.0f;
    return (uint32_t)f;
}
uint32_t morton3(float x, float y, float z) {  // src/bvh.cpp:23-32
    return (spread3(quant10(x)) << 2) | (spread3(quant10(y)) << 1) | spread3(quant10(z));
}
void fit(float mn[3], float mx[3], const float p[3]) {  // src/bvh.cpp:34-42
    for (int k = 0; k < 3; ++k) {
        if (p[k] < mn[k]) mn[k] = p[k];
        if (p[k] > mx[k]) mx[k] = p[k];
    }
}
int find_split(const uint32_t* m, int first, int last, uint32_t mask) {  // src/bvh.cpp:55-65
    int lo = first, hi = last;
    while (lo < hi) {
        int mid = lo + ((hi - lo) >> 1);
        if (m[mid] & mask) hi = mid; else lo = mid + 1;
    }
    return lo;
}

struct Builder {
    std::vector<ONode>& nodes;
    const std::vector<OTri>& tris;  // sorted
    const uint32_t* morton;
    int recurse(int first, int last, int bit) {  // src/bvh.cpp:75-120
        int idx = (int)nodes.size();
        nodes.emplace_back();
        int count = last - first;
        int split = first;
        while (bit >= 0) {
            split = find_split(morton, first, last, 1u << bit);
            if (split != first && split != last) break;
            --bit;
        }
        if (count <= kLeafMaxTris || bit < 0) {
            ONode& n = nodes[idx];
            n.left = -1; n.right = -1; n.firstTri = first; n.triCount = count;
            float mn[3] = {INFINITY, INFINITY, INFINITY};
            float mx[3] = {-INFINITY, -INFINITY, -INFINITY};
            for (int i = first; i < last; ++i)
                for (int k = 0; k < 3; ++k) fit(mn, mx, tris[i].v[k]);
            std::memcpy(n.bbMin, mn, 12);
            std::memcpy(n.bbMax, mx, 12);
            return idx;
        }
        int L = recurse(first, split, bit - 1);
        int R = recurse(split, last, bit - 1);
        ONode& self = nodes[idx];
        self.left = L; self.right = R; self.firstTri = -1; self.triCount = 0;
        std::memcpy(self.bbMin, nodes[L].bbMin, 12);
        std::memcpy(self.bbMax, nodes[L].bbMax, 12);
        for (int k = 0; k < 3; ++k) {  // fit_union, src/bvh.cpp:44-52
            if (nodes[R].bbMin[k] < self.bbMin[k]) self.bbMin[k] = nodes[R].bbMin[k];
            if (nodes[R].bbMax[k] > self.bbMax[k]) self.bbMax[k] = nodes[R].bbMax[k];
        }
        return idx;
    }
};

void relayout_bfs(std::vector<ONode>& nodes) {  // src/bvh.cpp:124-162
    if (nodes.empty()) return;
    const int n = (int)nodes.size();
    std::vector<int> oldToNew(n, -1), order;
    order.reserve(n);
    // BFS with an explicit index queue; left enqueued before right.
    std::vector<int> q;
    q.reserve(n);
    q.push_back(0);
    oldToNew[0] = 0;
    for (size_t head = 0; head < q.size(); ++head) {
        int oi = q[head];
        order.push_back(oi);
        const ONode& nd = nodes[oi];
        if (nd.triCount == 0) {
            oldToNew[nd.left] = (int)q.size(); q.push_back(nd.left);
            oldToNew[nd.right] = (int)q.size(); q.push_back(nd.right);
        }
    }
    std::vector<ONode> out(n);
    for (int i = 0; i < n; ++i) {
        ONode nn = nodes[order[i]];
        if (nn.triCount == 0) { nn.left = oldToNew[nn.left]; nn.right = oldToNew[nn.right]; }
        out[i] = nn;
    }
    nodes.swap(out);
}

}  // namespace

// ======================================================================= C ABI
extern "C" {

uint32_t orc_morton3(float x, float y, float z) { return morton3(x, y, z); }
uint32_t orc_spread3(uint32_t v) { return spread3(v); }

void orc_sincos(const float* x, float* s, float* c, int64_t n) {
    for (int64_t i = 0; i < n; ++i) own_sincos(x[i], &s[i], &c[i]);
}

void orc_pcg32_stream(uint64_t initstate, uint64_t initseq, uint32_t* out, int n) {
    Pcg g;
    pcg_seed(g, initstate, initseq);
    for (int i = 0; i < n; ++i) out[i] = pcg_next(g);
}
void orc_xorshift_stream(uint32_t seed, uint32_t* out, int n) {
    uint32_t s = seed;
    for (int i = 0; i < n; ++i) out[i] = xorshift32(s);
}
// n cosine-weighted directions around N from consecutive draws of one PCG32 stream.
void orc_sample_cos_hemi(const float N[3], uint64_t initstate, uint64_t initseq, float* dirs, int n) {
    Rng r{};
    r.pcg = 1;
    pcg_seed(r.g, initstate, initseq);
    for (int i = 0; i < n; ++i) {
        V3 d = sample_cos_hemi(vfrom(N), r);
        dirs[3 * i] = d.x; dirs[3 * i + 1] = d.y; dirs[3 * i + 2] = d.z;
    }
}
// First n uniform floats of the (pixel, sample) stream used by the renderer.
void orc_pixel_randf(uint32_t seed, uint32_t pixel, uint32_t sample, float* out, int n) {
    Rng r;
    r.pcg = 1;
    pcg_for_sample(r.g, seed, pixel, sample);
    for (int i = 0; i < n; ++i) out[i] = r.next();
}

// Scene Morton codes (unsorted, original order), src/bvh.cpp:170-192.
void orc_morton_codes(const OTri* tris, int64_t n, uint32_t* out) {
    float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int64_t i = 0; i < n; ++i)
        for (int c = 0; c < 3; ++c) fit(mn, mx, tris[i].v[c]);
    float ext[3] = {mx[0] - mn[0], mx[1] - mn[1], mx[2] - mn[2]};
    for (int a = 0; a < 3; ++a) if (ext[a] <= 0.0f) ext[a] = 1.0f;
    for (int64_t i = 0; i < n; ++i) {
        const OTri& t = tris[i];
        float cx = (t.v[0][0] + t.v[1][0] + t.v[2][0]) * (1.0f / 3.0f);
        float cy = (t.v[0][1] + t.v[1][1] + t.v[2][1]) * (1.0f / 3.0f);
        float cz = (t.v[0][2] + t.v[1][2] + t.v[2][2]) * (1.0f / 3.0f);
        out[i] = morton3((cx - mn[0]) / ext[0], (cy - mn[1]) / ext[1], (cz - mn[2]) / ext[2]);
    }
}

// BuildBVH restated (src/bvh.cpp:164-211). The Morton sort is STABLE (ties keep the
// original triangle order); the reference's std::sort leaves tie order unspecified, which
// only permutes triangles inside a leaf. Returns the node count; nodes must hold 2*n.
// perm[i] = original index of the triangle at sorted position i; keys = sorted codes.
int64_t orc_bvh_build(const OTri* tris, int64_t n, ONode* nodes_out, int32_t* perm,
                      uint32_t* keys) {
    if (n <= 0) return 0;
    std::vector<uint32_t> code(n);
    orc_morton_codes(tris, n, code.data());
    std::vector<int32_t> order(n);
    for (int64_t i = 0; i < n; ++i) order[i] = (int32_t)i;
    std::stable_sort(order.begin(), order.end(),
                     [&](int32_t a, int32_t b) { return code[a] < code[b]; });
    std::vector<OTri> sorted(n);
    std::vector<uint32_t> msorted(n);
    for (int64_t i = 0; i < n; ++i) {
        sorted[i] = tris[order[i]];
        msorted[i] = code[order[i]];
        if (perm) perm[i] = order[i];
        if (keys) keys[i] = msorted[i];
    }
    std::vector<ONode> nodes;
    nodes.reserve(2 * n);
    Builder b{nodes, sorted, msorted.data()};
    b.recurse(0, (int)n, 29);
    relayout_bfs(nodes);
    std::memcpy(nodes_out, nodes.data(), nodes.size() * sizeof(ONode));
    return (int64_t)nodes.size();
}

// tanHalfFov as the reference host computes it (src/pathtracer.cu:660).
float orc_tan_half_fov(float fovYDeg) {
    return std::tan(fovYDeg * 0.5f * 3.14159265358979f / 180.0f);
}

// overwrite_camera, src/main.cpp:29-59
void orc_camera_look_at(OCam* cam, const float pos[3], const float tgt[3]) {
    float fx = tgt[0] - pos[0], fy = tgt[1] - pos[1], fz = tgt[2] - pos[2];
    float fl = std::sqrt(fx * fx + fy * fy + fz * fz);
    if (fl < 1e-8f) return;
    fx /= fl; fy /= fl; fz /= fl;
    float lx = -fy, ly = fx, lz = 0.0f;
    float ll = std::sqrt(lx * lx + ly * ly + lz * lz);
    if (ll < 1e-6f) { lx = 0; ly = 1; lz = 0; ll = 1; }
    lx /= ll; ly /= ll; lz /= ll;
    float ux = fy * lz - fz * ly, uy = fz * lx - fx * lz, uz = fx * ly - fy * lx;
    float* m = cam->c2w;
    m[0] = fx; m[1] = fy; m[2] = fz; m[3] = 0;
    m[4] = lx; m[5] = ly; m[6] = lz; m[7] = 0;
    m[8] = ux; m[9] = uy; m[10] = uz; m[11] = 0;
    m[12] = pos[0]; m[13] = pos[1]; m[14] = pos[2]; m[15] = 1;
    cam->pos[0] = pos[0]; cam->pos[1] = pos[1]; cam->pos[2] = pos[2];
}

struct OrcRenderArgs {
    const ONode* nodes;
    const OTri* tris;          // permuted (BVH leaf) order
    const float* mat_albedo;   // 3 per material
    const OPoint* points;
    int32_t n_points;
    OSun sun;
    OCam cam;
    int32_t spp, bounces;
    uint32_t flags;            // bit0: xorshift32-ref RNG; bit2: no Russian roulette
    uint32_t seed;
    int32_t sample0;
    int32_t row_step;          // render rows y with y % row_step == 0 (1 = full frame)
    int32_t threads;
    const uint32_t* tex_texels;  // nullable: diffuse textures (see TexView)
    const int32_t* tex_desc;     // 4 per material
    const float* srgb_lut;       // 256
};

// Full render. out_sum/out_comp: W*H*3 Kahan FP32 (only sampled rows written);
// out_sum64 (nullable): FP64 sums; out_sumsq (nullable): per-pixel sum of squared
// luminance (0.2126, 0.7152, 0.0722); counters: node_visits, tri_tests, closest_hits,
// rays, shadow_rays (nullable, uint64[5]).
void orc_render(const OrcRenderArgs* a, float* out_sum, float* out_comp, double* out_sum64,
                float* out_sumsq, uint64_t* counters) {
    SceneView s{a->nodes, a->tris, a->mat_albedo, a->points, a->n_points, a->sun,
                TexView{a->tex_texels, a->tex_desc, a->srgb_lut}};
    CamView cam = make_cam(a->cam, orc_tan_half_fov(a->cam.fovYDeg));
    const int W = a->cam.width, H = a->cam.height;
    const bool xs_ref = (a->flags & 1u) != 0;
    const bool no_rr = (a->flags & 4u) != 0;
    const int rs = a->row_step > 0 ? a->row_step : 1;
    int nthreads = a->threads > 0 ? a->threads : (int)std::max(1u, std::thread::hardware_concurrency());
    std::atomic<int> nextRow{0};
    std::vector<Counters> perThread(nthreads);
    auto worker = [&](int tid) {
        Counters& c = perThread[tid];
        for (;;) {
            int yi = nextRow.fetch_add(1);
            int py = yi * rs;
            if (py >= H) break;
            for (int px = 0; px < W; ++px) {
                const size_t pi = (size_t)py * W + px;
                float sum[3], comp[3];
                double s64[3] = {0, 0, 0};
                for (int k = 0; k < 3; ++k) { sum[k] = out_sum[3 * pi + k]; comp[k] = out_comp[3 * pi + k]; }
                float sq = out_sumsq ? out_sumsq[pi] : 0.0f;
                Rng rng{};
                rng.pcg = xs_ref ? 0 : 1;
                for (int k = 0; k < a->spp; ++k) {
                    const int sample = a->sample0 + k;
                    if (xs_ref) {
                        // src/pathtracer.cu:349 + chunking :699-702: reseed at each 8-sample chunk
                        if (k == 0 || (sample % 8) == 0) {
                            uint32_t base = (uint32_t)(sample - (sample % 8));
                            rng.xs = ((uint32_t)px * 1973u + (uint32_t)py * 9277u + base * 26699u) | 1u;
                        }
                    } else {
                        pcg_for_sample(rng.g, a->seed, (uint32_t)pi, (uint32_t)sample);
                    }
                    float jx = rng.next();
                    float jy = rng.next();
                    Ray r = primary_ray(cam, px, py, jx, jy);
                    V3 L = trace_path(s, r, a->bounces, rng, no_rr, c);
                    kahan_add(sum[0], comp[0], L.x);
                    kahan_add(sum[1], comp[1], L.y);
                    kahan_add(sum[2], comp[2], L.z);
                    s64[0] += L.x; s64[1] += L.y; s64[2] += L.z;
                    float lum = 0.2126f * L.x + 0.7152f * L.y + 0.0722f * L.z;
                    sq += lum * lum;
                }
                for (int k = 0; k < 3; ++k) { out_sum[3 * pi + k] = sum[k]; out_comp[3 * pi + k] = comp[k]; }
                if (out_sum64) for (int k = 0; k < 3; ++k) out_sum64[3 * pi + k] += s64[k];
                if (out_sumsq) out_sumsq[pi] = sq;
            }
        }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < nthreads; ++t) pool.emplace_back(worker, t);
    worker(0);
    for (auto& th : pool) th.join();
    if (counters) {
        Counters tot;
        for (auto& c : perThread) tot.add(c);
        counters[0] = tot.node_visits; counters[1] = tot.tri_tests; counters[2] = tot.closest_hits;
        counters[3] = tot.rays; counters[4] = tot.shadow_rays;
    }
}

// Primary hit at the pixel centre (src/pathtracer.cu:381-384): triangle index in the
// permuted order (-1 = miss) and hit distance.
void orc_primary_hits(const OrcRenderArgs* a, int32_t* tri_out, float* t_out) {
    SceneView s{a->nodes, a->tris, a->mat_albedo, a->points, a->n_points, a->sun,
                TexView{a->tex_texels, a->tex_desc, a->srgb_lut}};
    CamView cam = make_cam(a->cam, orc_tan_half_fov(a->cam.fovYDeg));
    const int W = a->cam.width, H = a->cam.height;
    int nthreads = a->threads > 0 ? a->threads : (int)std::max(1u, std::thread::hardware_concurrency());
    std::atomic<int> nextRow{0};
    auto worker = [&]() {
        Counters c;
        for (;;) {
            int py = nextRow.fetch_add(1);
            if (py >= H) break;
            for (int px = 0; px < W; ++px) {
                Ray r = primary_ray(cam, px, py, 0.5f, 0.5f);
                float t = 0.0f;
                V3 N;
                float tu, tv;
                int tri = traverse_closest(s, r, 1e30f, t, N, tu, tv, c);
                tri_out[(size_t)py * W + px] = tri;
                t_out[(size_t)py * W + px] = tri < 0 ? -1.0f : t;
            }
        }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < nthreads; ++t) pool.emplace_back(worker);
    worker();
    for (auto& th : pool) th.join();
}

// Brute-force closest hit over ALL triangles (no BVH), same ray_tri and tie rule.
int32_t orc_brute_closest(const OTri* tris, int64_t n, const float o[3], const float d[3], float* t_out) {
    Ray r = make_ray(vfrom(o), vfrom(d));
    float best = 1e30f;
    int32_t bi = -1;
    for (int64_t i = 0; i < n; ++i) {
        float tt, uu, vv;
        if (ray_tri(r, tris[i], best, tt, uu, vv)) { best = tt; bi = (int32_t)i; }
    }
    *t_out = best;
    return bi;
}
// BVH closest hit for one ray (for brute-force comparisons).
int32_t orc_bvh_closest(const ONode* nodes, const OTri* tris, const float o[3], const float d[3], float* t_out) {
    SceneView s{nodes, tris, nullptr, nullptr, 0, {}};
    Ray r = make_ray(vfrom(o), vfrom(d));
    Counters c;
    V3 N;
    float t = 1e30f;
    float tu, tv;
    int32_t tri = traverse_closest(s, r, 1e30f, t, N, tu, tv, c);
    *t_out = t;
    return tri;
}

// sample_albedo of material `mat` at texture coordinates (u, v), for direct tests.
void orc_tex_sample(const uint32_t* texels, const int32_t* desc, const float* srgb, const float* mat_albedo,
                    int32_t mat, float u, float v, float* out3) {
    SceneView s{nullptr, nullptr, mat_albedo, nullptr, 0, {}, TexView{texels, desc, srgb}};
    const V3 a = sample_albedo(s, mat, u, v);
    out3[0] = a.x; out3[1] = a.y; out3[2] = a.z;
}

// gbuffer_kernel, src/pathtracer.cu:366-411 and host extents :728-734.
void orc_gbuffer(const OrcRenderArgs* a, uint8_t* alb, uint8_t* nrm, uint8_t* wp) {
    SceneView s{a->nodes, a->tris, a->mat_albedo, a->points, a->n_points, a->sun,
                TexView{a->tex_texels, a->tex_desc, a->srgb_lut}};
    CamView cam = make_cam(a->cam, orc_tan_half_fov(a->cam.fovYDeg));
    const ONode& root = a->nodes[0];
    V3 bbMin = vfrom(root.bbMin), bbMax = vfrom(root.bbMax);
    V3 ext = sub(bbMax, bbMin);
    V3 invExt = v3(1.0f / std::fmax(ext.x, 1e-6f), 1.0f / std::fmax(ext.y, 1e-6f), 1.0f / std::fmax(ext.z, 1e-6f));
    auto enc = [](float v) { return (uint8_t)(std::fmin(1.0f, std::fmax(0.0f, v)) * 255.0f); };
    Counters c;
    for (int py = 0; py < cam.height; ++py)
        for (int px = 0; px < cam.width; ++px) {
            Ray r = primary_ray(cam, px, py, 0.5f, 0.5f);
            float t;
            V3 N;
            float tu, tv;
            int tri = traverse_closest(s, r, 1e30f, t, N, tu, tv, c);
            size_t i = ((size_t)py * cam.width + px) * 3;
            if (tri < 0) {
                alb[i] = alb[i + 1] = alb[i + 2] = 0;
                nrm[i] = 128; nrm[i + 1] = 128; nrm[i + 2] = 255;
                wp[i] = wp[i + 1] = wp[i + 2] = 0;
                continue;
            }
            if (dot(N, r.d) > 0.0f && ORC_MUTANT != 4) N = mul(N, -1.0f);
            const V3 A = sample_albedo(s, s.tris[tri].material, tu, tv);
            alb[i] = enc(A.x); alb[i + 1] = enc(A.y); alb[i + 2] = enc(A.z);
            nrm[i] = enc(N.x * 0.5f + 0.5f); nrm[i + 1] = enc(N.y * 0.5f + 0.5f); nrm[i + 2] = enc(N.z * 0.5f + 0.5f);
            V3 P = add(r.o, mul(r.d, t));
            V3 Pn = had(sub(P, bbMin), invExt);
            wp[i] = enc(Pn.x); wp[i + 1] = enc(Pn.y); wp[i + 2] = enc(Pn.z);
        }
}

// tonemap_kernel, src/pathtracer.cu:413-429
void orc_tonemap(const float* accum, int64_t n_px, float invSpp, uint8_t* out) {
    for (int64_t i = 0; i < 3 * n_px; ++i) {
        float v = std::sqrt(std::fmax(0.0f, std::fmin(accum[i] * invSpp, 1.0f)));
        out[i] = (uint8_t)(v * 255.0f);
    }
}

}  // extern "C"
