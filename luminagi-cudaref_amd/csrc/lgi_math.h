// lgi_math.h — device math for the tracer, written for gfx950 and built with
// -ffp-contract=off so each expression rounds exactly as written. Every function
// keeps the operation order of the reference semantics it implements
// (include/vecmath.h:9-44; src/pathtracer.cu:61-107, 193-293) so that FP32 results
// are bit-identical to a CPU evaluation of the same expressions.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define LGI_D __device__ __forceinline__

// Correctly rounded FP32 square root and division. HIP's __fsqrt_rn maps to the
// approximate native sqrt unless OCML_BASIC_ROUNDED_OPERATIONS is defined; the builtin
// lowers to the IEEE-exact expansion (the default -fhip-fp32-correctly-rounded-divide-sqrt).
LGI_D float lgi_sqrt(float x) { return __builtin_sqrtf(x); }
LGI_D float lgi_div(float a, float b) { return a / b; }

namespace lgi {

struct V3 { float x, y, z; };
LGI_D V3 v3(float x, float y, float z) { return {x, y, z}; }
LGI_D V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
LGI_D V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
LGI_D V3 mul(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
LGI_D V3 had(V3 a, V3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
LGI_D float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
LGI_D V3 cross(V3 a, V3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
// normalize with the 1e-20 fallback to +Z (include/vecmath.h:31-35): reciprocal, then scale.
LGI_D V3 normalize(V3 a) {
    float l = lgi_sqrt(dot(a, a));
    return (l > 1e-20f) ? mul(a, lgi_div(1.0f, l)) : v3(0.f, 0.f, 1.f);
}
struct Ray { V3 o, d, inv; };
LGI_D Ray make_ray(V3 o, V3 d) {
    Ray r;
    r.o = o;
    r.d = normalize(d);
    r.inv = {lgi_div(1.0f, r.d.x), lgi_div(1.0f, r.d.y), lgi_div(1.0f, r.d.z)};
    return r;
}

// sin/cos of phi in [0, 2*pi): Cody-Waite reduction by pi/2 with a three-part constant,
// then the Cephes single-precision minimax polynomials on [-pi/4, pi/4] (S. Moshier,
// Cephes Math Library, sinf.c). Separately rounded ops; identical on CPU.
LGI_D void own_sincos(float x, float* s, float* c) {
    const float kPio2Hi = 1.5703125f;
    const float kPio2Mid = 4.83751296997070312e-04f;
    const float kPio2Lo = 7.54978995489188216e-08f;
    float kf = __builtin_rintf(x * 0.636619772367581343f);
    int k = (int)kf;
    float r = ((x - kf * kPio2Hi) - kf * kPio2Mid) - kf * kPio2Lo;
    float z = r * r;
    float ps = ((-1.9515295891e-4f * z + 8.3321608736e-3f) * z - 1.6666654611e-1f) * z * r + r;
    float pc = ((2.443315711809948e-5f * z - 1.388731625493765e-3f) * z + 4.166664568298827e-2f) * z * z
               - 0.5f * z + 1.0f;
    int q = k & 3;
    float sv = (q & 1) ? pc : ps;
    float cv = (q & 1) ? ps : pc;
    *s = (q & 2) ? -sv : sv;
    *c = ((q + 1) & 2) ? -cv : cv;
}

// ------------------------------------------------------------------ RNG
// PCG32 XSH-RR 64/32 (M. O'Neill, pcg_basic.c). One stream per (pixel, sample):
// increment from the pixel stream id, initial state from a SplitMix64 hash of
// (stream, sample), so any sample can be generated independently of the others.
struct Pcg { uint64_t state, inc; };
LGI_D uint32_t pcg_next(Pcg& g) {
    uint64_t old = g.state;
    g.state = old * 6364136223846793005ULL + g.inc;
    uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xs >> rot) | (xs << ((0u - rot) & 31u));
}
LGI_D uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
LGI_D void pcg_for_sample(Pcg& g, uint32_t seed, uint32_t pixel, uint32_t sample) {
    uint64_t stream = ((uint64_t)seed << 32) | pixel;
    uint64_t initstate = mix64(stream ^ ((uint64_t)sample << 40) ^ 0x9E3779B97F4A7C15ULL);
    g.state = 0u;
    g.inc = (stream << 1u) | 1u;
    pcg_next(g);
    g.state += initstate;
    pcg_next(g);
}
// xorshift32 (src/pathtracer.cu:193-199)
LGI_D uint32_t xorshift32(uint32_t& s) {
    s ^= s << 13;
    s ^= s >> 17;
    s ^= s << 5;
    return s;
}
struct Rng {
    Pcg g;
    uint32_t xs;
    bool pcg;
    LGI_D float next() {
        uint32_t x = pcg ? pcg_next(g) : xorshift32(xs);
        return (float)(x >> 8) * (1.0f / 16777216.0f);
    }
};

// Frisvad ONB (src/pathtracer.cu:206-214)
LGI_D void onb(V3 N, V3& T, V3& B) {
    if (N.z < -0.99999f) { T = v3(0.f, -1.f, 0.f); B = v3(-1.f, 0.f, 0.f); return; }
    float a = lgi_div(1.0f, 1.0f + N.z);
    float b = -N.x * N.y * a;
    T = v3(1.0f - N.x * N.x * a, b, -N.x);
    B = v3(b, 1.0f - N.y * N.y * a, -N.y);
}

// Cosine-weighted hemisphere (src/pathtracer.cu:216-228)
LGI_D V3 sample_cos_hemi(V3 N, Rng& rng) {
    float u1 = rng.next();
    float u2 = rng.next();
    float r = lgi_sqrt(u1);
    float phi = 2.0f * 3.14159265358979f * u2;
    float sn, cs;
    own_sincos(phi, &sn, &cs);
    float x = r * cs;
    float y = r * sn;
    float z = lgi_sqrt(fmaxf(0.0f, 1.0f - u1));
    V3 T, B;
    onb(N, T, B);
    return normalize(add(add(mul(T, x), mul(B, y)), mul(N, z)));
}

LGI_D void kahan_add(float& sum, float& comp, float x) {
    float y = x - comp;
    float t = sum + y;
    comp = (t - sum) - y;
    sum = t;
}

}  // namespace lgi
