// trace_device.h — device-side ray/box, ray/triangle, camera and shading helpers shared by
// the megakernel (trace.hip) and the wavefront renderer (wavefront.hip). Expression trees
// follow the reference (src/pathtracer.cu:61-107, 149-155, 230-293) with no FP contraction.
#pragma once
#include "lgi_internal.h"
#include "lgi_math.h"
#include "trace_kernels.h"

namespace lgi {

LGI_D bool slab(const Ray& r, float lox, float hix, float loy, float hiy, float loz, float hiz, float tmax,
                float& tenter) {
    // src/pathtracer.cu:61-80, same expression tree
    float t0 = (lox - r.o.x) * r.inv.x;
    float t1 = (hix - r.o.x) * r.inv.x;
    float lo = fminf(t0, t1), hi = fmaxf(t0, t1);
    t0 = (loy - r.o.y) * r.inv.y;
    t1 = (hiy - r.o.y) * r.inv.y;
    lo = fmaxf(lo, fminf(t0, t1));
    hi = fminf(hi, fmaxf(t0, t1));
    t0 = (loz - r.o.z) * r.inv.z;
    t1 = (hiz - r.o.z) * r.inv.z;
    lo = fmaxf(lo, fminf(t0, t1));
    hi = fminf(hi, fmaxf(t0, t1));
    tenter = fmaxf(lo, 0.0f);
    return !(hi < tenter || lo > tmax);
}

// Moller-Trumbore on the precomputed edges (src/pathtracer.cu:82-107).
LGI_D bool tri_hit(const Ray& r, const TriV& t, float tmax, float& outT, float& outU, float& outV) {
    const float kEps = 1e-6f;
    V3 v0 = v3(t.v0.x, t.v0.y, t.v0.z);
    V3 e1 = v3(t.e1.x, t.e1.y, t.e1.z);
    V3 e2 = v3(t.e2.x, t.e2.y, t.e2.z);
    V3 p = cross(r.d, e2);
    float det = dot(e1, p);
    if (det > -kEps && det < kEps) return false;
    float invDet = lgi_div(1.0f, det);
    V3 s = sub(r.o, v0);
    float u = dot(s, p) * invDet;
    if (u < 0.0f || u > 1.0f) return false;
    V3 q = cross(s, e1);
    float v = dot(r.d, q) * invDet;
    if (v < 0.0f || u + v > 1.0f) return false;
    float tt = dot(e2, q) * invDet;
    if (tt <= kEps || tt >= tmax) return false;
    outT = tt;
    outU = u;
    outV = v;
    return true;
}

struct Hit {
    int tri;   // leaf-order index, -1 = miss
    float t, u, v;
};

// Traversal stack: the top kLds entries live in LDS as [depth][thread] (lanes at different
// depths still hit distinct banks), deeper entries in a per-thread global slot laid out as
// [depth - kLds][thread] (coalesced). The reference tree has at most 30 internal levels (one
// split per Morton bit), so kStackMax = 32 entries in total can never overflow; 16 LDS entries
// (16 KB per 256-thread block) keep 8 waves per SIMD resident.
template <int kLds>
struct TravStackT {
    int* lds;          // s_stack + threadIdx.x, stride kBlockThreads
    int* spill;        // spill + global thread id, stride spill_stride
    int spill_stride;
    static constexpr int kBlockThreads = 256;
    // The empty asm ends the LDS branch with an instruction the global branch lacks, so the
    // compiler cannot sink the two accesses into one generic-pointer (FLAT) access: a FLAT
    // access waits on both the vector-memory and the LDS counters, an LDS access only on the
    // latter.
    LGI_D void push(int sp, int v) const {
        if (sp < kLds) {
            lds[sp * kBlockThreads] = v;
            __asm__ volatile("" ::: "memory");
        } else {
            spill[(sp - kLds) * spill_stride] = v;
        }
    }
    LGI_D int pop(int sp) const {
        int v;
        if (sp < kLds) {
            v = lds[sp * kBlockThreads];
            __asm__ volatile("" ::: "memory");
        } else {
            v = spill[(sp - kLds) * spill_stride];
        }
        return v;
    }
};
using TravStack = TravStackT<kStackLds>;

// Shading normal and texture coordinates at the closest hit (src/pathtracer.cu:149-157).
LGI_D V3 hit_normal(const TriS* tris, const Hit& h, int& mat, float& tu, float& tv) {
    const TriS s = tris[h.tri];
    float w = 1.0f - h.u - h.v;
    mat = s.mat;
    tu = w * s.uv[0] + h.u * s.uv[2] + h.v * s.uv[4];
    tv = w * s.uv[1] + h.u * s.uv[3] + h.v * s.uv[5];
    return normalize(v3(w * s.n[0] + h.u * s.n[3] + h.v * s.n[6], w * s.n[1] + h.u * s.n[4] + h.v * s.n[7],
                        w * s.n[2] + h.u * s.n[5] + h.v * s.n[8]));
}

// Primary ray through (px + jx, py + jy) (src/pathtracer.cu:284-293).
LGI_D Ray primary_ray(const DevCamera& c, int px, int py, float jx, float jy) {
    float ndcx = lgi_div(2.0f * ((float)px + jx), (float)c.width) - 1.0f;
    float ndcy = 1.0f - lgi_div(2.0f * ((float)py + jy), (float)c.height);
    float rx = ndcx * c.aspect * c.tan_half_fov;
    float ry = ndcy * c.tan_half_fov;
    V3 fwd = v3(c.fwd[0], c.fwd[1], c.fwd[2]);
    V3 left = v3(c.left[0], c.left[1], c.left[2]);
    V3 up = v3(c.up[0], c.up[1], c.up[2]);
    V3 dw = add(add(mul(fwd, 1.0f), mul(left, -rx)), mul(up, ry));
    return make_ray(v3(c.pos[0], c.pos[1], c.pos[2]), dw);
}

// Diffuse texture lookup with the semantics of the reference's CUDA texture object
// (src/pathtracer.cu:274-282, 563-576): normalised coordinates (u, 1 - v), wrap addressing,
// RGBA8 texels converted from sRGB to linear before filtering, bilinear filtering with the
// texture unit's 8-bit fractional weights (1/256 steps, rounded to nearest).
LGI_D float tex_weight(float f) { return floorf(f * 256.0f + 0.5f) * (1.0f / 256.0f); }

LGI_D V3 tex_sample(const uint32_t* tex, int W, int H, const float* lut, float tu, float tv) {
    float u = tu - floorf(tu);
    float v = (1.0f - tv) - floorf(1.0f - tv);
    if (!(u >= 0.0f && u < 1.0f)) u = 0.0f;  // NaN or infinite coordinates
    if (!(v >= 0.0f && v < 1.0f)) v = 0.0f;
    const float x = u * (float)W - 0.5f, y = v * (float)H - 0.5f;
    const float xf = floorf(x), yf = floorf(y);
    const float a = tex_weight(x - xf), b = tex_weight(y - yf);
    int x0 = (int)xf, y0 = (int)yf;
    x0 = x0 < 0 ? x0 + W : x0;
    y0 = y0 < 0 ? y0 + H : y0;
    x0 = min(max(x0, 0), W - 1);
    y0 = min(max(y0, 0), H - 1);
    const int x1 = x0 + 1 == W ? 0 : x0 + 1, y1 = y0 + 1 == H ? 0 : y0 + 1;
    auto texel = [&](int xx, int yy) {
        const uint32_t p = tex[(size_t)yy * W + xx];
        return v3(lut[p & 255u], lut[(p >> 8) & 255u], lut[(p >> 16) & 255u]);
    };
    const V3 t00 = texel(x0, y0), t10 = texel(x1, y0), t01 = texel(x0, y1), t11 = texel(x1, y1);
    const V3 top = add(mul(t00, 1.0f - a), mul(t10, a));
    const V3 bot = add(mul(t01, 1.0f - a), mul(t11, a));
    return add(mul(top, 1.0f - b), mul(bot, b));
}

// sample_albedo (src/pathtracer.cu:274-282): the texture when the material has one, else the
// constant albedo.
LGI_D V3 sample_albedo(const TraceParams& P, int mat, float tu, float tv) {
    if (P.tex_desc) {
        const int4 d = P.tex_desc[mat];
        if (d.w) return tex_sample(P.tex_texels + d.x, d.y, d.z, P.srgb_lut, tu, tv);
    }
    return v3(P.mat_albedo[3 * mat], P.mat_albedo[3 * mat + 1], P.mat_albedo[3 * mat + 2]);
}

// Direction to light li from the shading point (li 0 = sun, then the point lights) and, for a
// point light, its distance (src/pathtracer.cu:236-262).
LGI_D void light_dir(const TraceParams& P, int li, V3 Pt, V3& toLight, float& dist) {
    if (li == 0) {
        toLight = normalize(mul(v3(P.sun_dir[0], P.sun_dir[1], P.sun_dir[2]), -1.0f));
        dist = 0.0f;
        return;
    }
    const lgi_point_light L = P.points[li - 1];
    V3 toLightVec = sub(v3(L.pos[0], L.pos[1], L.pos[2]), Pt);
    dist = lgi_sqrt(dot(toLightVec, toLightVec));
    toLight = mul(toLightVec, lgi_div(1.0f, dist));
}

// Shadow ray of a NEE candidate: from Pt + N * 1e-3 toward the light, tmax 1e30 for the sun and
// dist - 1e-3 for a point light.
LGI_D void shadow_ray(int li, V3 Pt, V3 N, V3 toLight, float dist, Ray& shadow, float& tmax) {
    const float kShadowBias = 1e-3f;
    shadow = make_ray(add(Pt, mul(N, kShadowBias)), toLight);
    tmax = li == 0 ? 1e30f : dist - kShadowBias;
}

// Whether light li is a NEE candidate at (Pt, N) in shade_direct (src/pathtracer.cu:230-272):
// light index 0 is the sun, 1 + i the i-th point light. Also returns the direction and distance
// to the light and the clamped cosine. A point light at distance >= radius has attenuation
// exactly 0, so it is not a candidate (its shadow ray would only ever add +0).
LGI_D bool nee_light(const TraceParams& P, int li, V3 Pt, V3 N, V3& toLight, float& dist, float& ndotl) {
    light_dir(P, li, Pt, toLight, dist);
    ndotl = fmaxf(0.0f, dot(N, toLight));
    if (li == 0) return ndotl > 0.0f;
    const float r = P.points[li - 1].radius;
    if (ndotl <= 0.0f || dist >= r * 3.0f) return false;
    return !(dist >= r);  // atten == 0 exactly
}

// One NEE candidate: false when the light cannot contribute; otherwise the shadow ray, its tmax
// and the contribution albedo * (color * I) to add if unoccluded.
LGI_D bool nee_candidate(const TraceParams& P, int li, V3 Pt, V3 N, V3 albedo, Ray& shadow, float& tmax,
                         V3& contrib) {
    V3 toLight;
    float dist, ndotl;
    if (!nee_light(P, li, Pt, N, toLight, dist, ndotl)) return false;
    shadow_ray(li, Pt, N, toLight, dist, shadow, tmax);
    if (li == 0) {
        float I = P.sun_intensity * ndotl;
        contrib = had(albedo, mul(v3(P.sun_color[0], P.sun_color[1], P.sun_color[2]), I));
        return true;
    }
    const lgi_point_light L = P.points[li - 1];
    float r = L.radius;
    float atten = fmaxf(0.0f, 1.0f - lgi_div(dist, r));
    atten = atten * atten;
    float I = L.intensity * ndotl * atten;
    contrib = had(albedo, mul(v3(L.color[0], L.color[1], L.color[2]), I));
    return true;
}

}  // namespace lgi
