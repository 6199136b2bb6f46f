// wavefront.hip — sorted ("wavefront") path tracer for gfx950.
//
// The reference's --sort path (src/pathtracer.cu:431-547, 761-909) splits a sample into a
// primary-ray kernel, a trace-and-bounce kernel per bounce and a Thrust sort of the
// surviving rays by a direction key. This renderer keeps that structure and its purpose
// (coherent secondary rays) but is organised for MI355X:
//   * all spp samples of a frame are in flight at once (one path per (pixel, sample): up to
//     tens of millions of paths, a few GB of queues in HBM);
//   * closest-hit and shadow rays are traced by one persistent kernel that pulls 64-ray
//     chunks from a queue with one atomic per chunk and refills each lane the moment its ray
//     finishes, so no lane idles while its wave still has work;
//   * secondary rays are sorted between bounces by a 30-bit key (24-bit origin Morton code in
//     the scene box, 6-bit direction Morton code) with the project's LSD radix sort;
//   * results are written per path, never with atomics: each path keeps its own radiance, and
//     samples are added to the pixel in sample order with Kahan compensation at the end.
// Unlike the reference sorted path, Russian roulette and the per-path RNG stream are kept,
// so the image is bit-identical to the inline megakernel (same PCG32 stream per
// (pixel, sample), same expression trees, same light order).
#include "trace_device.h"

#include <climits>

using namespace lgi;

namespace {

constexpr int kBlock = 256;
constexpr int kTraceGridMax = 256 * 8;
static_assert(TravStack::kBlockThreads == 256, "stack stride");

struct alignas(16) QRay {
    float4 o;   // origin, w = path id (int bits)
    float4 d;   // direction, w = tmax (shadow rays) / unused
};

struct WfParams {
    TraceParams P;
    int32_t npx;          // W*H
    int32_t nl;           // lights: sun + points
    float3 bb_min;        // scene box for the sort key
    float3 bb_inv;        // 1 / extent
    float* T;             // throughput per path (3)
    float4* vT;           // per queue slot: throughput before the bounce update (xyz)
    float* L;             // radiance per path (3)
    float* contrib;       // per (queue slot, light): contribution if unoccluded (3)
    uint8_t* nee;         // per (queue slot, light): 0 none, 1 candidate, 2 occluded
    float4* vP;           // per queue slot: shading point (xyz)
    float4* vN;           // per queue slot: facing shading normal (xyz)
    uint8_t* occl;        // per (queue slot, light) of the last bounce when its shadow rays are
                          // traced in the launch before (kFold): for its candidates, 2 occluded,
                          // else 0; or null
    const int32_t* tiles; // 8x8 tiles to render (row-major tile ids), or null for the whole image
    int32_t set_sums;     // k_wf_accum starts from zero (LGI_RENDER_SET_SUMS, first pass)
};

// Shadow rays are not stored: the any-hit kernel reads the id j = slot * nl + light of each
// candidate and rebuilds the ray from the slot's shading point and normal with the same
// expressions the shading kernel evaluated (light_dir, shadow_ray).
struct ShadowSrc {
    const uint32_t* ids;  // compacted candidate ids
    const float4* vP;
    const float4* vN;
    int nl;
    TraceParams P;        // light parameters
    uint8_t* occl;        // kFold: occlusion of the last bounce's candidates (2 = occluded)
};

__device__ __forceinline__ uint32_t spread3(uint32_t v) {
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}
__device__ __forceinline__ uint32_t q01(float f, float scale, uint32_t maxv) {
    f = f * scale;
    f = fminf(fmaxf(f, 0.0f), (float)maxv);
    return (uint32_t)f;
}
// Sort key: origin Morton code (kKeyOriginBits per axis) above direction Morton code
// (kKeyDirBits per axis); 30 bits. On the 1.05M-triangle scene at 64 spp, 8 + 2 bits render
// 1.2% faster than 6 + 4 (7 + 3: +0.8%, 5 + 5: -0.8%, 9 + 1: same as 8 + 2, origin only:
// +0.9%, direction bits above origin bits: +0.3%). The key orders work only: each path keeps
// its own RNG stream and radiance, so the image does not depend on it.
constexpr int kKeyOriginBits = 8, kKeyDirBits = 2;
constexpr int kKeySortBits = (3 * (kKeyOriginBits + kKeyDirBits) + 7) / 8 * 8;  // radix passes of 8 bits
static_assert(3 * (kKeyOriginBits + kKeyDirBits) <= 32, "sort key fits 32 bits");
__device__ __forceinline__ uint32_t ray_key(const WfParams& W, V3 o, V3 d) {
    constexpr float so = (float)(1 << kKeyOriginBits), sd = (float)(1 << kKeyDirBits);
    constexpr uint32_t mo = (1u << kKeyOriginBits) - 1u, md = (1u << kKeyDirBits) - 1u;
    uint32_t ox = q01((o.x - W.bb_min.x) * W.bb_inv.x, so, mo), oy = q01((o.y - W.bb_min.y) * W.bb_inv.y, so, mo),
             oz = q01((o.z - W.bb_min.z) * W.bb_inv.z, so, mo);
    uint32_t dx = q01(d.x * 0.5f + 0.5f, sd, md), dy = q01(d.y * 0.5f + 0.5f, sd, md), dz = q01(d.z * 0.5f + 0.5f, sd, md);
    return (((spread3(ox) << 2) | (spread3(oy) << 1) | spread3(oz)) << (3 * kKeyDirBits)) |
           ((spread3(dx) << 2) | (spread3(dy) << 1) | spread3(dz));
}

// ------------------------------------------------------------------ primary rays
// The bounce-0 queue is never stored: entry i is a function of i, evaluated where it is read
// (the closest-hit trace of bounce 0 and the shading of bounce 0). Queue index i -> (sample s,
// 8x8 pixel tile, lane), so each wave starts on one tile.
struct PrimarySrc {
    DevCamera cam;
    uint32_t seed;
    int32_t sample0;      // first sample of this pass
    int32_t npx;          // W*H
    const int32_t* tiles; // 8x8 tiles to render, or null for the whole image
    int32_t tiles_x, padded;
};

// Sample s and pixel (px, py) of entry i; false for a lane outside the image.
__device__ __forceinline__ bool primary_pixel(const PrimarySrc& ps, int64_t i, int& s, int& px, int& py) {
    s = (int)(i / ps.padded);
    const int j = (int)(i % ps.padded);
    const int tile = ps.tiles ? ps.tiles[j >> 6] : j >> 6, lane = j & 63;
    px = (tile % ps.tiles_x) * 8 + (lane & 7);
    py = (tile / ps.tiles_x) * 8 + (lane >> 3);
    return px < ps.cam.width && py < ps.cam.height;
}

// Path id of entry i (-1 for a lane outside the image).
__device__ __forceinline__ int64_t primary_path(const PrimarySrc& ps, int64_t i) {
    int s, px, py;
    if (!primary_pixel(ps, i, s, px, py)) return -1;
    return (int64_t)s * ps.npx + (uint32_t)py * ps.cam.width + px;
}

// Entry i's ray (o.w = path id, or -1 for a lane outside the image; d.w = tmax) and the path's
// RNG after the two jitter draws.
__device__ __forceinline__ QRay primary_entry(const PrimarySrc& ps, int64_t i, Pcg& g) {
    int s, px, py;
    QRay r;
    if (!primary_pixel(ps, i, s, px, py)) {
        r.o = make_float4(0.f, 0.f, 0.f, __int_as_float(-1));
        r.d = make_float4(0.f, 0.f, 1.f, 0.f);
        return r;
    }
    const uint32_t pix = (uint32_t)py * ps.cam.width + px;
    const int64_t p = (int64_t)s * ps.npx + pix;
    Rng rng;
    rng.pcg = true;
    pcg_for_sample(rng.g, ps.seed, pix, (uint32_t)(ps.sample0 + s));
    const float jx = rng.next();
    const float jy = rng.next();
    const Ray ray = primary_ray(ps.cam, px, py, jx, jy);
    g = rng.g;
    r.o = make_float4(ray.o.x, ray.o.y, ray.o.z, __int_as_float((int)p));
    r.d = make_float4(ray.d.x, ray.d.y, ray.d.z, 1e30f);
    return r;
}

// ------------------------------------------------------------------ persistent traversal
// Closest hit (hits[i] = t, u, v, tri) or any hit (nee[slot] = 2 when occluded) for queue
// entries [0, n). Lanes whose ray finishes take the next ray from the wave's 64-entry chunk;
// a wave grabs a new chunk with one atomicAdd when the current one runs out.
// The queue is split into G = min(8, grid) contiguous segments, one per XCD (workgroups are
// dispatched to the 8 XCDs round-robin, so block b runs on XCD b % 8): each segment has its
// own cursor (cursors[16 * g], one 64-byte line each), so the rays one XCD traces at a time
// are neighbours in queue order (ray-sorted, or tile order) and share its L2. A wave whose
// segment is drained takes chunks from the following segments.
constexpr int kCursorStride = 16;  // u32 entries between segment cursors
constexpr int kSegments = 8;
// Queue entries per cursor grab: one wave's worth. An XCD's waves then trace a window of about
// 1024 waves x 64 rays of its segment at a time; 256-entry chunks widen that window and measured
// 10% slower on the 1.05M-triangle frame.
constexpr uint32_t kChunk = 64;
// kTop > 0: the first kTop traversal nodes (the top levels: nodes are in breadth-first order)
// are copied into LDS once per block and read from there; kLds stack entries live in LDS.
// kMinLanes: phase 1 hands over to the leaf tests once fewer than kMinLanes of the lanes still
// traversing have no parked leaf (1: only when every such lane holds one).
// kMode: kTraceClosest (queue q, n entries), kTraceShadow (shadow rays, n_sh entries) or
// kTraceBoth: the shadow rays of one bounce and the closest-hit rays of the next in one launch
// (queue positions [0, n_sh) are shadow rays, [n_sh, n_sh + n) closest-hit rays). Both depend
// only on the shading of the bounce before, and one launch instead of two saves a persistent
// kernel's tail, the time its last rays take after the queue has run dry.
// kFold (kTraceBoth whose closest-hit rays are the path's last bounce): a lane that finds such a
// ray's closest hit computes its shading point and normal as the shading kernel does, then
// traces that hit's NEE shadow rays itself, one light after another, and marks occluded ones in
// shadow.occl. The last bounce's shading then reads those marks instead of a shadow-ray launch
// of its own: one persistent launch, with its tail, fewer per pass.
enum TraceMode { kTraceClosest = 0, kTraceShadow = 1, kTraceBoth = 2 };

// Lane of the k-th set bit of m (k < popcount(m)).
__device__ __forceinline__ int nth_lane(uint64_t m, int k) {
    for (int j = 0; j < k; ++j) m &= m - 1;
    return __ffsll((unsigned long long)m) - 1;
}

template <int kMode, int kLds, int kTop, int kMinLanes, bool kFold>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(kFold ? 7 : 1))) void k_wf_trace(const TNode* __restrict__ tnodes, int n_tnodes,
                                                     const TriV* __restrict__ triv, const QRay* __restrict__ q, int n,
                                                     const uint32_t* n_dev, int n_sh, const uint32_t* n_sh_dev,
                                                     float4* hits, uint8_t* nee, uint32_t* cursors, int* spill,
                                                     ShadowSrc shadow, PrimarySrc prim, const int32_t* qord) {
    if (kMode == kTraceShadow) n = 0;
    if (kMode == kTraceClosest) n_sh = 0;
    if (n_dev) n = min(n, (int)*n_dev);
    if (n_sh_dev) n_sh = min(n_sh, (int)*n_sh_dev);
    const int n_all = n_sh + n;
    __shared__ int s_stack[kLds * kBlock];
    __shared__ TNode s_top[kTop > 0 ? kTop : 1];
    // kFold: the shading point and normal of the lane's last-bounce hit, [component][thread]
    __shared__ float s_fold[kFold ? 6 * kBlock : 1];
    static_assert(!kFold || kMode == kTraceBoth, "the fold traces closest-hit and shadow rays");
    if (kTop > 0) {
        const int cnt = min(kTop, n_tnodes);
        for (int i = threadIdx.x; i < cnt * 4; i += kBlock)
            reinterpret_cast<float4*>(s_top)[i] = reinterpret_cast<const float4*>(tnodes)[i];
        __syncthreads();
    }
    const TravStackT<kLds> stk{s_stack + threadIdx.x, spill + blockIdx.x * kBlock + threadIdx.x,
                               (int)gridDim.x * kBlock};
    const int lane = threadIdx.x & 63;
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    // wave-uniform chunk cursor and segment
    uint32_t cbase = 0, cend = 0;
    bool exhausted = false;
    const int n_seg = min(kSegments, (int)gridDim.x);
    int seg = (int)blockIdx.x % n_seg, tried = 0;
    auto seg_lo = [&](int g) { return (uint32_t)((uint64_t)(uint32_t)n_all * (uint32_t)g / (uint32_t)n_seg); };
    bool active = false;
    bool any = kMode == kTraceShadow;  // this lane's ray is a shadow ray (any hit ends it)
    // node: >= 0 internal node to visit, < 0 leaf (~first triangle), kNone when the stack is
    // exhausted; leaf: a parked leaf (0 = none)
    constexpr int kNone = INT_MIN;
    // kFold: qi = -1 while the lane traces a shadow ray of its own last-bounce hit (-2 when it
    // helps another lane with one); while inactive, slot = ~j marks candidate j = qi * nl + light
    // of that hit as the next one to test
    int qi = 0, node = 0, sp = 0, leaf = 0, best_tri = -1, slot = 0;
    float best_t = 0.f, best_u = 0.f, best_v = 0.f;
    Ray r;
    for (;;) {
        if (kFold && !active && slot < 0) {
            // the next light of this lane's last-bounce hit that is a NEE candidate, if any
            const int nl = shadow.nl, j0 = ~slot, i = j0 / nl;
            const float* f = s_fold + threadIdx.x;
            const V3 Pt = v3(f[0], f[kBlock], f[2 * kBlock]), N = v3(f[3 * kBlock], f[4 * kBlock], f[5 * kBlock]);
            V3 toLight = v3(0.f, 0.f, 0.f);
            float dist = 0.f, ndotl = 0.f;
            int li = j0 - i * nl;
            while (li < nl && !nee_light(shadow.P, li, Pt, N, toLight, dist, ndotl)) ++li;
            slot = 0;
            if (li < nl) {
                Ray sr;
                float tmax;
                shadow_ray(li, Pt, N, toLight, dist, sr, tmax);
                r.o = sr.o;
                r.d = sr.d;
                r.inv = v3(lgi_div(1.0f, r.d.x), lgi_div(1.0f, r.d.y), lgi_div(1.0f, r.d.z));
                best_t = tmax;
                best_tri = -1;
                node = 0;
                leaf = 0;
                sp = 0;
                any = true;
                qi = -1;
                slot = i * nl + li;
                active = true;
                // written before any lane can help with this ray (helpers take subtrees of
                // a ray in a later iteration of this wave), so a helper's 2 lands after it
                shadow.occl[slot] = 0;
            }
        }
        const uint64_t need = __ballot(!active);
        if (need && !exhausted) {
            const uint32_t cnt = (uint32_t)__popcll(need);
            const uint32_t rank = (uint32_t)__popcll(need & lt);
            uint32_t avail = cend - cbase;
            uint32_t nb = 0xFFFFFFFFu, nend = 0;
            if (avail < cnt) {
                for (;;) {  // ends: every failed grab moves to the next segment, G segments at most
                    uint32_t old = 0;
                    if (lane == 0) old = atomicAdd(cursors + kCursorStride * seg, kChunk);
                    old = __shfl(old, 0, 64);
                    const uint32_t s0 = seg_lo(seg), s1 = seg_lo(seg + 1);
                    if (old < s1 - s0) {
                        nb = s0 + old;
                        nend = min(nb + kChunk, s1);
                        break;
                    }
                    if (++tried == n_seg) {
                        exhausted = true;
                        nb = 0xFFFFFFFFu;
                        break;
                    }
                    seg = seg + 1 == n_seg ? 0 : seg + 1;
                }
            }
            if (!active) {
                uint32_t idx = 0xFFFFFFFFu;
                if (rank < avail) idx = cbase + rank;
                else if (nb != 0xFFFFFFFFu && rank - avail < nend - nb) idx = nb + (rank - avail);
                const bool sh = kMode == kTraceShadow || (kMode == kTraceBoth && idx < (uint32_t)n_sh);
                if (kMode == kTraceBoth) any = sh;
                if (idx != 0xFFFFFFFFu && sh) {
                    const uint32_t j = shadow.ids[idx];
                    const int si = (int)(j / (uint32_t)shadow.nl), li = (int)(j % (uint32_t)shadow.nl);
                    const float4 p4 = shadow.vP[si], n4 = shadow.vN[si];
                    const V3 Pt = v3(p4.x, p4.y, p4.z), N = v3(n4.x, n4.y, n4.z);
                    V3 toLight;
                    float dist, tmax;
                    light_dir(shadow.P, li, Pt, toLight, dist);
                    Ray sr;
                    shadow_ray(li, Pt, N, toLight, dist, sr, tmax);
                    qi = (int)idx;
                    slot = (int)j;
                    r.o = sr.o;
                    r.d = sr.d;
                    r.inv = v3(lgi_div(1.0f, r.d.x), lgi_div(1.0f, r.d.y), lgi_div(1.0f, r.d.z));
                    best_t = tmax;
                    best_tri = -1;
                    node = 0;
                    leaf = 0;
                    sp = 0;
                    active = true;
                } else if (idx != 0xFFFFFFFFu) {
                    if (kMode == kTraceBoth) idx -= (uint32_t)n_sh;
                    Pcg unused;
                    const QRay e = kMode == kTraceClosest ? primary_entry(prim, idx, unused)
                                                          : q[qord ? qord[idx] : (int)idx];
                    qi = (int)idx;
                    slot = __float_as_int(e.o.w);
                    const float tmax = e.d.w;
                    if (slot >= 0) {
                        r.o = v3(e.o.x, e.o.y, e.o.z);
                        r.d = v3(e.d.x, e.d.y, e.d.z);
                        r.inv = v3(lgi_div(1.0f, r.d.x), lgi_div(1.0f, r.d.y), lgi_div(1.0f, r.d.z));
                        best_t = tmax;
                        best_tri = -1;
                        node = 0;
                        leaf = 0;
                        sp = 0;
                        active = true;
                    } else {
                        hits[idx] = make_float4(-1.f, 0.f, 0.f, __int_as_float(-1));
                        if (kFold) slot = 0;  // no pending candidate
                    }
                }
            }
            // advance the cursor past what was consumed
            const uint32_t from_old = min(cnt, avail);
            if (nb != 0xFFFFFFFFu) {
                const uint32_t from_new = min(cnt - from_old, nend - nb);
                cbase = nb + from_new;
                cend = nend;
            } else {
                cbase += from_old;
            }
        }
        if (kFold && exhausted) {
            // Once the queue is dry, a lane that has just started a light of its own hit hands
            // the lights after it to an idle lane (with the hit's shading point and normal), so
            // a hit's shadow rays run side by side in the launch's tail instead of one after
            // another. The taker tests them from the next iteration on, as its own.
            const bool can_hand = active && qi == -1 && (slot + 1) % shadow.nl != 0;
            const uint64_t idle = __ballot(!active), handers = __ballot(can_hand);
            if (idle != 0 && handers != 0) {
                const int npair = min(__popcll(idle), __popcll(handers));
                const int hrank = __popcll(handers & lt), irank = __popcll(idle & lt);
                const bool take = !active && irank < npair;
                const int src = take ? nth_lane(handers, irank) : lane;
                const int sl = __shfl(slot, src, 64);
                if (can_hand && hrank < npair) qi = -2;  // its later lights are the taker's
                if (take) {
                    const int from = (int)threadIdx.x - lane + src;
                    for (int k = 0; k < 6; ++k) s_fold[k * kBlock + threadIdx.x] = s_fold[k * kBlock + from];
                    slot = ~(sl + 1);
                }
            }
        }
        if (kMode != kTraceClosest && exhausted) {
            // Work sharing once the queue is dry: an idle lane takes the next pending subtree of
            // a lane whose any-hit (shadow) ray still has one (its stack top, with the ray), so
            // the long rays that end a launch are traversed by several lanes. A hit in any
            // subtree marks the ray occluded (nee[slot] = 2), which no order of visits changes.
            // Closest-hit rays are not shared: their tie order at equal distances would follow
            // the schedule.
            const bool can_give = active && any && sp > 0;
            // (a lane with a pending light of its own hit is not idle)
            const uint64_t idle = __ballot(!active && !(kFold && slot < 0)), donors = __ballot(can_give);
            if (idle != 0 && donors != 0) {
                const int npair = min(__popcll(idle), __popcll(donors));
                const int drank = __popcll(donors & lt), irank = __popcll(idle & lt);
                int give = 0;
                if (can_give && drank < npair) give = stk.pop(--sp);
                const bool take = !active && !(kFold && slot < 0) && irank < npair;
                const int src = take ? nth_lane(donors, irank) : lane;
                const int g = __shfl(give, src, 64), sl = __shfl(slot, src, 64);
                const int dq = kFold ? __shfl(qi, src, 64) : 0;
                const float bt = __shfl(best_t, src, 64);
                const V3 o = v3(__shfl(r.o.x, src, 64), __shfl(r.o.y, src, 64), __shfl(r.o.z, src, 64));
                const V3 d = v3(__shfl(r.d.x, src, 64), __shfl(r.d.y, src, 64), __shfl(r.d.z, src, 64));
                const V3 iv = v3(__shfl(r.inv.x, src, 64), __shfl(r.inv.y, src, 64), __shfl(r.inv.z, src, 64));
                if (take) {
                    any = true;
                    if (kFold) qi = dq < 0 ? -2 : 0;  // a helper marks the array its donor marks
                    r.o = o;
                    r.d = d;
                    r.inv = iv;
                    best_t = bt;
                    slot = sl;
                    node = g;
                    leaf = 0;
                    sp = 0;
                    best_tri = -1;
                    active = true;
                }
            }
        }
        if (__ballot(active || (kFold && slot < 0)) == 0) {
            if (exhausted) break;
            continue;
        }
        if (!active) continue;
        // Phase 1: internal nodes. A lane that reaches a leaf parks it and keeps traversing
        // (speculatively) until every lane still in this loop holds a parked leaf, so the
        // wave then tests triangles with most lanes busy instead of alternating one node step
        // with one whole leaf per iteration.
        while (node >= 0) {
            TNode nd;
            if (kTop > 0 && node < kTop) {
                nd = s_top[node];
                __asm__ volatile("" ::: "memory");  // keeps this an LDS read (see TravStackT)
            } else {
                nd = tnodes[node];
            }
            float ta, tb;
            const bool ha = slab(r, nd.b0xy.x, nd.b0xy.y, nd.b0xy.z, nd.b0xy.w, nd.bz.x, nd.bz.y, best_t, ta);
            const bool hb = slab(r, nd.b1xy.x, nd.b1xy.y, nd.b1xy.z, nd.b1xy.w, nd.bz.z, nd.bz.w, best_t, tb);
            if (ha && hb) {
                int nearc = nd.ch.x, farc = nd.ch.y;
                if (tb < ta) { nearc = nd.ch.y; farc = nd.ch.x; }
                stk.push(sp, farc);
                ++sp;
                node = nearc;
            } else if (ha) {
                node = nd.ch.x;
            } else if (hb) {
                node = nd.ch.y;
            } else {
                node = sp > 0 ? stk.pop(--sp) : kNone;
            }
            if (node < 0 && node != kNone && leaf == 0) {
                leaf = node;
                node = sp > 0 ? stk.pop(--sp) : kNone;
            }
            if (__popcll(__ballot(leaf == 0)) < kMinLanes) break;
        }
        // Phase 2: the parked leaf, then the current node for as long as it is a leaf.
        while (leaf != 0 || (node < 0 && node != kNone)) {
            if (leaf == 0) {
                leaf = node;
                node = sp > 0 ? stk.pop(--sp) : kNone;
            }
            int ti = ~leaf;
            leaf = 0;
            for (;;) {
                const TriV t = triv[ti];
                float tt, uu, vv;
                if (tri_hit(r, t, best_t, tt, uu, vv)) {
                    best_tri = ti; best_t = tt; best_u = uu; best_v = vv;
                    if (any) break;
                }
                if (__float_as_int(t.v0.w) != 0) break;
                ++ti;
            }
            if (any && best_tri >= 0) {
                node = kNone;
                sp = 0;
                break;
            }
        }
        if (node == kNone && leaf == 0) {
            if (any) {
                if (kFold && qi < 0) {
                    if (best_tri >= 0) shadow.occl[slot] = 2;
                    if (qi == -1) {  // the lane's own hit: its next light, if any
                        const int nx = slot + 1;
                        slot = nx % shadow.nl != 0 ? ~nx : 0;
                    }
                } else if (best_tri >= 0) {
                    nee[slot] = 2;
                }
            } else {
                hits[qi] = make_float4(best_t, best_u, best_v, __int_as_float(best_tri));
                if (kFold && best_tri >= 0) {
                    // the last bounce's shading point and facing normal, with the shading
                    // kernel's expressions (k_wf_shade), and its first light as pending
                    int mat;
                    float tu, tv;
                    const Hit h{best_tri, best_t, best_u, best_v};
                    V3 N = hit_normal(shadow.P.tris, h, mat, tu, tv);
                    if (dot(N, r.d) > 0.0f) N = mul(N, -1.0f);
                    const V3 Pt = add(r.o, mul(r.d, best_t));
                    float* f = s_fold + threadIdx.x;
                    f[0] = Pt.x; f[kBlock] = Pt.y; f[2 * kBlock] = Pt.z;
                    f[3 * kBlock] = N.x; f[4 * kBlock] = N.y; f[5 * kBlock] = N.z;
                    slot = ~(qi * shadow.nl);
                }
            }
            active = false;
        }
    }
}

// ------------------------------------------------------------------ shading
// For each traced path: NEE candidates -> shadow queue, Russian roulette + cosine bounce ->
// next queue with its sort key (src/pathtracer.cu:295-331, same order of RNG draws).
// A path's RNG at bounce b >= 1, recomputed from its id rather than stored: its PCG32 stream
// starts at the (pixel, sample) seed, and a path reaches bounce b only by continuing through
// every earlier bounce, each of which drew a fixed number of values (2 for the camera jitter;
// per earlier bounce j, 1 Russian-roulette draw when j >= 1 and roulette is on, then 2 for the
// cosine-weighted direction). The state is advanced past those draws.
__device__ __forceinline__ Pcg path_rng(const WfParams& W, int64_t p, int bounce) {
    Pcg g;
    const uint32_t pix = (uint32_t)(p % W.npx), s = (uint32_t)(p / W.npx);
    pcg_for_sample(g, W.P.seed, pix, (uint32_t)W.P.sample0 + s);
    const bool rr = (W.P.flags & LGI_RENDER_NO_RR) == 0;
    int draws = 2;
    for (int j = 0; j < bounce; ++j) draws += 2 + (j >= 1 && rr ? 1 : 0);
    for (int k = 0; k < draws; ++k) pcg_next(g);
    return g;
}

// Block-wide exclusive scan of one value per thread (256 threads; every thread must call).
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* s_w) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t x = v;
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) s_w[wave] = x;
    __syncthreads();
    uint32_t pre = 0;
    for (int w = 0; w < wave; ++w) pre += s_w[w];
    return pre + x - v;
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t x) {
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    return x;
}

// Shading also counts, per 256-slot block, the NEE candidates and continuations it produced
// (blk_s, blk_n); a scan of those counts places each block's output, and the compaction
// kernels rank entries within their block, so no scan runs over the per-slot flags.
__global__ __launch_bounds__(kBlock) void k_wf_shade(WfParams W, const QRay* q, const float4* hits, int n,
                                                     const uint32_t* n_dev, int bounce, QRay* nq, uint32_t* nkeys,
                                                     uint8_t* ncont, uint32_t* blk_s, uint32_t* blk_n,
                                                     PrimarySrc prim, const int32_t* qord) {
    if (n_dev) n = min(n, (int)*n_dev);
    if ((int)blockIdx.x * kBlock >= n) {  // whole block past the count: zero counts for the scans
        if (threadIdx.x == 0) {
            blk_s[blockIdx.x] = 0;
            blk_n[blockIdx.x] = 0;
        }
        return;
    }
    const int i = blockIdx.x * kBlock + threadIdx.x;
    int p = -1;
    Hit h{-1, 0.f, 0.f, 0.f};
    QRay e;
    Pcg g0{0u, 0u};  // bounce 0: the path's RNG after its jitter draws
    if (i < n) {
        e = bounce == 0 ? primary_entry(prim, i, g0) : q[qord ? qord[i] : i];
        p = __float_as_int(e.o.w);
        if (p >= 0) {
            const float4 hv = hits[i];
            h.tri = __float_as_int(hv.w);
            h.t = hv.x; h.u = hv.y; h.v = hv.z;
        }
    }
    const bool live = p >= 0 && h.tri >= 0;
    V3 N = v3(0.f, 0.f, 1.f), Pt = v3(0.f, 0.f, 0.f), albedo = v3(0.f, 0.f, 0.f);
    V3 T0 = v3(1.f, 1.f, 1.f);  // throughput before this bounce: exactly 1 until the first bounce
    if (live) {
        const V3 d = v3(e.d.x, e.d.y, e.d.z);
        int mat;
        float tu, tv;
        N = hit_normal(W.P.tris, h, mat, tu, tv);
        if (dot(N, d) > 0.0f) N = mul(N, -1.0f);
        Pt = add(v3(e.o.x, e.o.y, e.o.z), mul(d, h.t));
        albedo = sample_albedo(W.P, mat, tu, tv);
        if (bounce > 0) {
            const float* T = W.T + 3 * (int64_t)p;
            T0 = v3(T[0], T[1], T[2]);
        }
    }
    // NEE candidates: flags and contributions in fixed slots i*nl + li (compacted afterwards);
    // the shading point and normal for the shadow rays
    uint32_t ncand = 0;
    bool cont = false;
    if (i < n) {
        for (int li = 0; li < W.nl; ++li) {
            Ray shadow;
            float tmax = 0.f;
            V3 c = v3(0.f, 0.f, 0.f);
            const bool cand = live && nee_candidate(W.P, li, Pt, N, albedo, shadow, tmax, c);
            const int64_t j = (int64_t)i * W.nl + li;
            // the last bounce's shadow rays were traced in the launch before (kFold): their
            // occlusion marks (written for every candidate)
            const bool occluded = W.occl && bounce == W.P.bounces && cand && W.occl[j] == 2;
            W.nee[j] = occluded ? 2 : cand ? 1 : 0;
            if (cand) { W.contrib[3 * j] = c.x; W.contrib[3 * j + 1] = c.y; W.contrib[3 * j + 2] = c.z; }
            ncand += cand ? 1u : 0u;
        }
        if (ncand) {  // w: the path id, for k_wf_resolve; the throughput it scales the light by
            W.vP[i] = make_float4(Pt.x, Pt.y, Pt.z, __int_as_float(p));
            W.vN[i] = make_float4(N.x, N.y, N.z, 0.f);
            W.vT[i] = make_float4(T0.x, T0.y, T0.z, 0.f);
        }
        // continuation
        cont = live && bounce < W.P.bounces;
        Ray next;
        if (cont) {
            float* T = W.T + 3 * (int64_t)p;
            V3 Tv = T0;
            Rng rng;
            rng.pcg = true;
            rng.g = bounce == 0 ? g0 : path_rng(W, p, bounce);
            if (bounce >= 1 && (W.P.flags & LGI_RENDER_NO_RR) == 0) {
                float qv = fminf(0.95f, fmaxf(Tv.x, fmaxf(Tv.y, Tv.z)));
                if (rng.next() > qv) cont = false;
                else Tv = mul(Tv, lgi_div(1.0f, qv));
            }
            if (cont) {
                V3 nd = sample_cos_hemi(N, rng);
                next = make_ray(add(Pt, mul(N, 1e-3f)), nd);
                Tv = had(Tv, albedo);
                T[0] = Tv.x; T[1] = Tv.y; T[2] = Tv.z;
            }
        }
        if (cont) {
            QRay nr;
            nr.o = make_float4(next.o.x, next.o.y, next.o.z, __int_as_float(p));
            nr.d = make_float4(next.d.x, next.d.y, next.d.z, 1e30f);
            nq[i] = nr;
            if (nkeys) nkeys[i] = ray_key(W, next.o, next.d);
        }
        ncont[i] = cont ? 1 : 0;
    }
    // per-block totals
    __shared__ uint32_t s_c[2][kBlock / 64];
    const uint32_t ws = wave_sum(ncand), wn = wave_sum(cont ? 1u : 0u);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) { s_c[0][wave] = ws; s_c[1][wave] = wn; }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t a0 = 0, a1 = 0;
        for (int w = 0; w < kBlock / 64; ++w) { a0 += s_c[0][w]; a1 += s_c[1][w]; }
        blk_s[blockIdx.x] = a0;
        blk_n[blockIdx.x] = a1;
    }
}

// Slots of the continuations of block b (flag ncont), and their sort keys when src_key is set,
// to the block's offset + their rank in the block (order-preserving). The block's offset is
// blk_off[b] + blk_chunk[b / kScanChunk] (lgi_exclusive_scan_chunks). The rays themselves stay where
// shading wrote them: the next bounce reads them through these slots, sorted by key or, without
// the ray sort, in slot order, so a continuation ray is never copied.
__global__ __launch_bounds__(kBlock) void k_wf_compact_keys(const uint32_t* src_key, const uint8_t* flag, int n,
                                                            const uint32_t* n_dev, const uint32_t* blk_off,
                                                            const uint32_t* blk_chunk, uint32_t* dst_key,
                                                            int32_t* dst_slot) {
    __shared__ uint32_t s_w[kBlock / 64];
    if (n_dev) n = min(n, (int)*n_dev);
    if ((int)blockIdx.x * kBlock >= n) return;  // whole block past the count (block-uniform)
    const int i = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t f = (i < n && flag[i]) ? 1u : 0u;
    const uint32_t r = block_excl_scan(f, s_w);
    if (!f) return;
    const uint32_t o = blk_off[blockIdx.x] + blk_chunk[blockIdx.x / kScanChunk] + r;
    if (src_key) dst_key[o] = src_key[i];
    dst_slot[o] = i;
}

// NEE candidate ids j = slot * nl + light of block b to the block's offset (as above) + their rank
// in the block (slot-major, lights in order).
__global__ __launch_bounds__(kBlock) void k_wf_compact_ids(const uint8_t* nee, int nl, int n, const uint32_t* n_dev,
                                                           const uint32_t* blk_off, const uint32_t* blk_chunk,
                                                           uint32_t* ids) {
    __shared__ uint32_t s_w[kBlock / 64];
    if (n_dev) n = min(n, (int)*n_dev);
    if ((int)blockIdx.x * kBlock >= n) return;  // whole block past the count (block-uniform)
    const int i = blockIdx.x * kBlock + threadIdx.x;
    uint32_t c = 0;
    if (i < n)
        for (int li = 0; li < nl; ++li) c += nee[(int64_t)i * nl + li] == 1 ? 1u : 0u;
    uint32_t o = blk_off[blockIdx.x] + blk_chunk[blockIdx.x / kScanChunk] + block_excl_scan(c, s_w);
    if (!c) return;
    for (int li = 0; li < nl; ++li) {
        const int64_t j = (int64_t)i * nl + li;
        if (nee[j] == 1) ids[o++] = (uint32_t)j;
    }
}

// L += Tb * sum of unoccluded contributions, lights in order (src/pathtracer.cu:313). A slot
// without an unoccluded candidate would add Tb * 0 = +0 to a non-negative L, which changes no
// bit, so it is skipped; the others find their path id in vP.w, written by the shading kernel
// for exactly the slots that have candidates.
__global__ __launch_bounds__(kBlock) void k_wf_resolve(WfParams W, int n, const uint32_t* n_dev, int bounce,
                                                       PrimarySrc prim) {
    if (n_dev) n = min(n, (int)*n_dev);
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    // bounce 0's slots are every path of the pass: their radiance sums start here, at 0 + the
    // unoccluded light (the same sum as adding to a zeroed L), or at 0
    const int64_t p0 = bounce == 0 ? primary_path(prim, i) : -1;
    V3 color = v3(0.f, 0.f, 0.f);
    bool any = false;
    for (int li = 0; li < W.nl; ++li) {
        const int64_t k = (int64_t)i * W.nl + li;
        if (W.nee[k] == 1) {
            color = add(color, v3(W.contrib[3 * k], W.contrib[3 * k + 1], W.contrib[3 * k + 2]));
            any = true;
        }
    }
    if (!any) {
        if (p0 >= 0) {
            float* L = W.L + 3 * p0;
            L[0] = 0.f; L[1] = 0.f; L[2] = 0.f;
        }
        return;
    }
    const int p = __float_as_int(W.vP[i].w);
    float* L = W.L + 3 * (int64_t)p;
    const float4 Tb = W.vT[i];  // the path's throughput before this bounce, stored by k_wf_shade
    const V3 L0 = bounce == 0 ? v3(0.f, 0.f, 0.f) : v3(L[0], L[1], L[2]);
    V3 Lv = add(L0, had(v3(Tb.x, Tb.y, Tb.z), color));
    L[0] = Lv.x; L[1] = Lv.y; L[2] = Lv.z;
}

__global__ __launch_bounds__(kBlock) void k_wf_fill(uint32_t* p, int n, uint32_t v) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) p[i] = v;
}

// Kahan-add the S samples of every pixel in sample order (same as the inline kernel).
__global__ __launch_bounds__(kBlock) void k_wf_accum(WfParams W, int S, int n_tiles, int tiles_x) {
    int pix;
    if (W.tiles) {  // pixels of the listed tiles
        const int k = blockIdx.x * kBlock + threadIdx.x;
        if (k >= n_tiles * 64) return;
        const int tile = W.tiles[k >> 6], lane = k & 63;
        const int px = (tile % tiles_x) * 8 + (lane & 7), py = (tile / tiles_x) * 8 + (lane >> 3);
        if (px >= W.P.cam.width || py >= W.P.cam.height) return;
        pix = py * W.P.cam.width + px;
    } else {
        pix = blockIdx.x * kBlock + threadIdx.x;
        if (pix >= W.npx) return;
    }
    float* sum = W.P.sum + 3 * (int64_t)pix;
    float* comp = W.P.comp + 3 * (int64_t)pix;
    const bool add = !W.set_sums;
    float s0 = add ? sum[0] : 0.f, s1 = add ? sum[1] : 0.f, s2 = add ? sum[2] : 0.f;
    float c0 = add ? comp[0] : 0.f, c1 = add ? comp[1] : 0.f, c2 = add ? comp[2] : 0.f;
    float sq = (add && W.P.sumsq) ? W.P.sumsq[pix] : 0.f;
    for (int s = 0; s < S; ++s) {
        const float* L = W.L + 3 * ((int64_t)s * W.npx + pix);
        kahan_add(s0, c0, L[0]);
        kahan_add(s1, c1, L[1]);
        kahan_add(s2, c2, L[2]);
        float lum = 0.2126f * L[0] + 0.7152f * L[1] + 0.0722f * L[2];
        sq += lum * lum;
    }
    sum[0] = s0; sum[1] = s1; sum[2] = s2;
    comp[0] = c0; comp[1] = c1; comp[2] = c2;
    if (W.P.sumsq) W.P.sumsq[pix] = sq;
}

inline int cdiv(int64_t a, int64_t b) { return (int)((a + b - 1) / b); }

// Traversal kernel configuration: 8 stack entries and the top 127 nodes (7 levels) in LDS,
// 16 KB per block as with 16 stack entries alone. Measured on the 1.05M-triangle scene at
// 1728x864, 64 spp: 3.4% less traversal time than 16 LDS stack entries and no node cache,
// with identical images; 255 nodes (32 KB) cost occupancy and were slower.
constexpr int kTraceLdsStack = 8, kTraceTopNodes = 127;
// Phase-1 hand-over threshold of the closest-hit and any-hit traversals. Going to the leaf
// tests once fewer than 8 lanes still look for their first leaf (instead of waiting for the
// last one) cut the traversal time by 4.5% on the same workload (thresholds 4 to 12 measure
// within 0.5% of each other); the lanes' leaf-test order, and so every result, is unchanged.
constexpr int kTraceMinLanes = 8, kTraceMinLanesAny = 8;
static_assert(kTraceLdsStack >= kStackLdsMin, "spill area sized for kStackMax - kStackLdsMin entries");

int trace_grid(int n) {
    // persistent: one resident round (16 KB of LDS per block -> 8 blocks = 32 waves per
    // CU on 256 CUs), never more blocks than the work needs
    return std::max(1, std::min(cdiv(n, kBlock), kTraceGridMax));
}

template <int kMode, bool kFold = false>
void launch_trace(hipStream_t st, const TNode* tn, int n_tn, const TriV* tv, const QRay* q, int n,
                  const uint32_t* n_dev, int n_sh, const uint32_t* n_sh_dev, float4* hits, uint8_t* nee,
                  uint32_t* cursors, int* spill, const ShadowSrc& shadow, const PrimarySrc& prim,
                  const int32_t* qord) {
    constexpr int kMin = kMode == kTraceShadow ? kTraceMinLanesAny : kTraceMinLanes;
    const int grid = trace_grid((kMode == kTraceShadow ? 0 : n) + (kMode == kTraceClosest ? 0 : n_sh));
    hipLaunchKernelGGL((k_wf_trace<kMode, kTraceLdsStack, kTraceTopNodes, kMin, kFold>), dim3(grid), dim3(kBlock), 0, st, tn,
                       n_tn, tv, q, n, n_dev, n_sh, n_sh_dev, hits, nee, cursors, spill, shadow, prim, qord);
}


}  // namespace

extern "C" lgi_status lgi_wavefront_samples_per_pass(int64_t n_slots, int32_t n_points, int32_t spp,
                                                     int32_t* samples) {
    if (!samples || n_slots <= 0 || n_points < 0 || spp <= 0) return LGI_ERR_INVALID_ARG;
    *samples = 0;
    // Paths in flight per pass: up to ~100M (about 25 GB of queues and path state, of the 288 GB
    // of HBM), so a 64-spp 1728x864 frame is one pass. More rays per launch make each XCD's
    // queue segment and the ray sort denser: on the 1.05M-triangle frame, one 64-spp pass traces
    // in 19.2 ms what four 16-spp passes traced in 26.2 ms.
    constexpr int64_t kMaxPaths = int64_t(96) << 20;
    // Shadow-ray candidates (slot x light) per pass: as many as the paths of a sun + 2-point-light
    // scene, so a scene with more lights runs fewer samples per pass instead of more memory. The
    // candidate ids slot * lights + light, the queue lengths and the launch sizes are 32-bit, so a
    // pass never exceeds 2^31 - 1 candidates; when one sample does, the image is refused.
    constexpr int64_t kMaxCandidates = 3 * kMaxPaths;
    const int64_t nl = 1 + (int64_t)n_points;
    const int64_t per_sample = n_slots * nl;
    if (n_slots > INT32_MAX || per_sample > INT32_MAX) return LGI_ERR_UNSUPPORTED;
    const int64_t s = std::min<int64_t>({spp, kMaxPaths / n_slots, kMaxCandidates / per_sample});
    *samples = (int32_t)std::max<int64_t>(1, s);
    return LGI_OK;
}

// Host driver of the wavefront renderer. Adds spp samples into P.sum/P.comp/P.sumsq.
lgi_status lgi_render_wavefront(lgi_ctx* ctx, const TraceParams& P, float3 bb_min, float3 bb_max,
                                const int32_t* d_tiles, int n_tiles) {
    hipStream_t st = ctx->stream;
    const int Wd = P.cam.width, Hd = P.cam.height;
    const int npx = Wd * Hd;
    const int tiles_x = (Wd + 7) / 8, tiles_y = (Hd + 7) / 8;
    const int padded = (d_tiles ? n_tiles : tiles_x * tiles_y) * 64;
    const int nl = 1 + P.n_points;
    int32_t S_max = 0;
    if (lgi_wavefront_samples_per_pass(padded, P.n_points, std::max(1, P.spp), &S_max) != LGI_OK)
        return lgi_set_err(ctx, LGI_ERR_UNSUPPORTED,
                           "wavefront renderer: one sample of this image times its lights exceeds 2^31 "
                           "shadow-ray candidates");
    const int64_t cap = (int64_t)S_max * padded;
    auto S_ = [&](int slot, size_t bytes) { return ctx->scratch_get(slot, bytes); };
    QRay* qa = (QRay*)S_(50, cap * sizeof(QRay));
    QRay* qb = (QRay*)S_(51, cap * sizeof(QRay));
    float4* hits = (float4*)S_(52, cap * sizeof(float4));
    uint32_t* sids = (uint32_t*)S_(53, cap * nl * 4);
    const int64_t nblk = cdiv(cap, kBlock);
    uint32_t* blk_s = (uint32_t*)S_(65, nblk * 4 + 64);
    uint32_t* blk_n = (uint32_t*)S_(66, nblk * 4 + 64);
    uint8_t* ncont = (uint8_t*)S_(71, cap + 64);
    uint32_t* keys_tmp = (uint32_t*)S_(67, cap * 4);
    uint32_t* keys = (uint32_t*)S_(54, cap * 4);
    int32_t* idx = (int32_t*)S_(55, cap * 4);
    void* sort_scratch = S_(56, lgi_sort_scratch_bytes(cap));
    // chunk sums of the two block-count scans (candidates, continuations), kept until the
    // compaction kernels add them
    const size_t scan_words = (lgi_scan_scratch_bytes(nblk) + 255) / 256 * 64;
    uint32_t* scan_scratch = (uint32_t*)S_(68, 8 * scan_words);
    uint32_t* sums_s = scan_scratch;
    uint32_t* sums_n = scan_scratch + scan_words;
    // per bounce: [1] shadow rays, [2] continuations, then the closest-hit and the shadow-ray
    // traversal's segment cursors, kCursorStride apart
    constexpr int kCounters = kCursorStride * (1 + 2 * kSegments);
    if (P.bounces > 64) return lgi_set_err(ctx, LGI_ERR_INVALID_ARG, "too many bounces");
    uint32_t* counters = (uint32_t*)S_(57, (size_t)kCounters * 4 * 65);
    int* spill = (int*)S_(69, (size_t)kStackSpill * kTraceGridMax * kBlock * 4);
    if (!spill) return lgi_set_err(ctx, LGI_ERR_HIP, "stack spill allocation failed");
    WfParams W;
    W.P = P;
    W.tiles = d_tiles;
    W.npx = npx;
    W.nl = nl;
    W.bb_min = bb_min;
    W.bb_inv = make_float3(1.0f / std::max(bb_max.x - bb_min.x, 1e-6f), 1.0f / std::max(bb_max.y - bb_min.y, 1e-6f),
                           1.0f / std::max(bb_max.z - bb_min.z, 1e-6f));
    const int64_t paths = (int64_t)S_max * npx;
    W.T = (float*)S_(59, paths * 12);
    W.vT = (float4*)S_(60, cap * 16);
    W.L = (float*)S_(61, paths * 12);
    W.contrib = (float*)S_(62, cap * nl * 12);
    W.nee = (uint8_t*)S_(63, cap * nl);
    W.vP = (float4*)S_(64, cap * 16);
    W.vN = (float4*)S_(70, cap * 16);
    // With a bounce after the first hit, the last bounce's shadow rays are traced in the launch
    // that finds its closest hits (kFold).
    const bool fold = P.bounces >= 1;
    W.occl = fold ? (uint8_t*)S_(75, cap * nl) : nullptr;
    if (fold && !W.occl) return lgi_set_err(ctx, LGI_ERR_HIP, "wavefront buffer allocation failed");
    ShadowSrc shadow{sids, W.vP, W.vN, nl, P, W.occl};
    for (void* p : {(void*)qa, (void*)qb, (void*)hits, (void*)sids, (void*)blk_s, (void*)blk_n, (void*)ncont,
                    (void*)keys_tmp,
                    (void*)keys, (void*)idx, sort_scratch, (void*)scan_scratch, (void*)counters, (void*)W.T,
                    (void*)W.vT, (void*)W.L, (void*)W.contrib, (void*)W.nee, (void*)W.vP, (void*)W.vN})
        if (!p) return lgi_set_err(ctx, LGI_ERR_HIP, "wavefront buffer allocation failed");
    const int n_tn = (int)ctx->dev.n_tnodes;
    const bool sort = !(P.flags & LGI_RENDER_NO_RAY_SORT);
    // No host round trip inside a frame: queue lengths stay on the device (each kernel reads its
    // count and the launches are sized for the capacity), so a frame can be captured in a graph.
    // Timing events are skipped while the stream is being captured.
    hipStreamCaptureStatus cap_status = hipStreamCaptureStatusNone;
    LGI_HIP(ctx, hipStreamIsCapturing(st, &cap_status));
    const bool timed = cap_status == hipStreamCaptureStatusNone;
    if (timed) LGI_HIP(ctx, hipEventRecord(ctx->ev0, st));
    ctx->kev_used = 0;
    ctx->kev_overflow = false;
    lgi_status ms;
    auto mark = [&](bool begin) { return timed ? lgi_mark_trace(ctx, begin) : LGI_OK; };
    for (int s0 = 0; s0 < P.spp; s0 += S_max) {
        const int S = std::min(S_max, P.spp - s0);
        W.P.sample0 = P.sample0 + s0;
        W.set_sums = (P.flags & LGI_RENDER_SET_SUMS) != 0 && s0 == 0;
        const int n_max = S * padded;
        const int nb = cdiv(n_max, kBlock);
        hipLaunchKernelGGL(k_wf_fill, dim3(cdiv(kCounters * (P.bounces + 1), kBlock)), dim3(kBlock), 0, st, counters,
                           kCounters * (P.bounces + 1), 0u);
        // bounce 0's queue is computed where it is read (primary_entry), never stored
        const PrimarySrc prim{W.P.cam, W.P.seed, W.P.sample0, W.npx, W.tiles, tiles_x, padded};
        QRay* cur = qa;
        QRay* other = qb;
        // queue order of bounces after the first: entry i is cur[qord[i]]
        const int32_t* qord = nullptr;
        const uint32_t* n_dev = nullptr;  // live entries of this bounce's queue (bounce 0: all n_max)
        if ((ms = mark(true)) != LGI_OK) return ms;
        launch_trace<kTraceClosest>(st, ctx->dev.tnodes, n_tn, ctx->dev.triv, cur, n_max, n_dev, 0, nullptr, hits,
                                    nullptr, counters + kCursorStride, spill, shadow, prim, nullptr);
        if ((ms = mark(false)) != LGI_OK) return ms;
        for (int b = 0; b <= P.bounces; ++b) {
            uint32_t* cnt = counters + kCounters * b;
            const bool last = b == P.bounces;
            hipLaunchKernelGGL(k_wf_shade, dim3(nb), dim3(kBlock), 0, st, W, cur, hits, n_max, n_dev, b, other,
                               sort ? keys_tmp : nullptr, ncont, blk_s, blk_n, prim, qord);
            if (last && fold) {  // its shadow rays are traced: add the unoccluded light
                hipLaunchKernelGGL(k_wf_resolve, dim3(nb), dim3(kBlock), 0, st, W, n_max, n_dev, b, prim);
                break;
            }
            // block offsets of the shadow-ray candidates and (but for the last bounce) of the
            // continuations, in two parts that the compaction kernels add
            lgi_exclusive_scan_chunks(st, blk_s, nb, cnt + 1, sums_s);
            if (!last) lgi_exclusive_scan_chunks(st, blk_n, nb, cnt + 2, sums_n);
            hipLaunchKernelGGL(k_wf_compact_ids, dim3(nb), dim3(kBlock), 0, st, (const uint8_t*)W.nee, nl, n_max, n_dev,
                               (const uint32_t*)blk_s, (const uint32_t*)sums_s, sids);
            if (!last) {
                // The next bounce's queue is shading's output read through a slot order (qord):
                // the continuations' slots compacted in slot order, and with the ray sort, sorted
                // by their keys. The rays stay where shading wrote them (this bounce's queue is no
                // longer read, so the next shading writes there); qord is read before the next
                // compaction or sort refills its buffer.
                hipLaunchKernelGGL(k_wf_compact_keys, dim3(nb), dim3(kBlock), 0, st, sort ? keys_tmp : nullptr,
                                   (const uint8_t*)ncont, n_max, n_dev, (const uint32_t*)blk_n,
                                   (const uint32_t*)sums_n, sort ? keys : nullptr, idx);
                qord = sort ? lgi_sort_pairs(st, keys, idx, n_max, kKeySortBits, sort_scratch, cnt + 2) : idx;
                std::swap(cur, other);
            }
            // this bounce's shadow rays and, but for the last bounce, the next bounce's closest
            // hits in the same persistent launch
            uint32_t* cur_sh = cnt + kCursorStride * (1 + kSegments);
            if ((ms = mark(true)) != LGI_OK) return ms;
            if (last)
                launch_trace<kTraceShadow>(st, ctx->dev.tnodes, n_tn, ctx->dev.triv, nullptr, 0, nullptr, n_max * nl,
                                           cnt + 1, nullptr, W.nee, cur_sh, spill, shadow, prim, nullptr);
            else if (fold && b + 1 == P.bounces)
                launch_trace<kTraceBoth, true>(st, ctx->dev.tnodes, n_tn, ctx->dev.triv, cur, n_max, cnt + 2,
                                               n_max * nl, cnt + 1, hits, W.nee, cur_sh, spill, shadow, prim, qord);
            else
                launch_trace<kTraceBoth>(st, ctx->dev.tnodes, n_tn, ctx->dev.triv, cur, n_max, cnt + 2, n_max * nl,
                                         cnt + 1, hits, W.nee, cur_sh, spill, shadow, prim, qord);
            if ((ms = mark(false)) != LGI_OK) return ms;
            hipLaunchKernelGGL(k_wf_resolve, dim3(nb), dim3(kBlock), 0, st, W, n_max, n_dev, b, prim);
            n_dev = cnt + 2;
        }
        const int n_acc = d_tiles ? n_tiles * 64 : npx;
        hipLaunchKernelGGL(k_wf_accum, dim3(cdiv(n_acc, kBlock)), dim3(kBlock), 0, st, W, S, n_tiles, tiles_x);
    }
    LGI_HIP(ctx, hipGetLastError());
    if (timed) LGI_HIP(ctx, hipEventRecord(ctx->ev1, st));
    return LGI_OK;
}
