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
    LGI_D void push(int sp, int v) const {
        if (sp < kLds) lds[sp * kBlockThreads] = v;
        else spill[(sp - kLds) * spill_stride] = v;
    }
    LGI_D int pop(int sp) const {
        return sp < kLds ? lds[sp * kBlockThreads] : spill[(sp - kLds) * spill_stride];
    }
};
using TravStack = TravStackT<kStackLds>;

// Shading normal at the closest hit (src/pathtracer.cu:149-155).
LGI_D V3 hit_normal(const TriS* tris, const Hit& h, int& mat) {
    const TriS s = tris[h.tri];
    float w = 1.0f - h.u - h.v;
    mat = s.mat;
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

// One NEE candidate of shade_direct (src/pathtracer.cu:230-272): light index 0 is the sun,
// 1 + i the i-th point light. Returns false when the light cannot contribute; otherwise
// the shadow ray, its tmax and the contribution albedo * (color * I) to add if unoccluded.
// A point light at distance >= radius has attenuation exactly 0, so it is reported as not
// contributing (its shadow ray would only ever add +0).
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

LGI_D bool nee_candidate(const TraceParams& P, int li, V3 Pt, V3 N, V3 albedo, Ray& shadow, float& tmax,
                         V3& contrib) {
    V3 toLight;
    float dist;
    light_dir(P, li, Pt, toLight, dist);
    float ndotl = fmaxf(0.0f, dot(N, toLight));
    if (li == 0) {
        if (!(ndotl > 0.0f)) return false;
        shadow_ray(li, Pt, N, toLight, dist, shadow, tmax);
        float I = P.sun_intensity * ndotl;
        contrib = had(albedo, mul(v3(P.sun_color[0], P.sun_color[1], P.sun_color[2]), I));
        return true;
    }
    const lgi_point_light L = P.points[li - 1];
    if (ndotl <= 0.0f || dist >= L.radius * 3.0f) return false;
    if (dist >= L.radius) return false;  // atten == 0 exactly
    shadow_ray(li, Pt, N, toLight, dist, shadow, tmax);
    float r = L.radius;
    float atten = fmaxf(0.0f, 1.0f - lgi_div(dist, r));
    atten = atten * atten;
    float I = L.intensity * ndotl * atten;
    contrib = had(albedo, mul(v3(L.color[0], L.color[1], L.color[2]), I));
    return true;
}

}  // namespace lgi
