// trace.hip — inline ("megakernel") path tracer for gfx950 (wave64), plus the primary-hit,
// G-buffer and tonemap kernels.
//
// Semantics follow the reference renderer (src/pathtracer.cu:61-429): Moller-Trumbore with
// the (1e-6, bestT) acceptance window, slab test with the `lo > tmax` cull, sun + point-light
// NEE with shadow rays, cosine bounces, Russian roulette from bounce 1, sqrt tonemap,
// G-buffer encodings. Every floating-point expression keeps the reference operation order
// and the file is built with -ffp-contract=off, so FP32 results are bit-identical to the CPU
// oracle; closest hits can differ only where two triangles tie in distance, because the
// traversal visits the nearer child first.
//
// Layout: TNode (both child boxes in the parent, 64 B), TriV (v0,e1,e2 = 48 B) in leaf order,
// TriS (normals + material) read once per closest hit. One lane per pixel, 8x8 pixel tile
// per wave, 4 waves per block. The traversal stack lives in LDS as [depth][thread], so lanes
// at different depths still hit distinct banks; the tree depth is at most 30 internal levels
// (one per Morton bit), so 32 entries never overflow.
#include "trace_device.h"

#include <algorithm>
#include <climits>
#include <mutex>

using namespace lgi;

namespace {

constexpr int kBlock = 256;
static_assert(TravStack::kBlockThreads == kBlock, "stack stride");

__device__ __forceinline__ TravStack make_stack(const TraceParams& P, int* s_stack) {
    const int gtid = (blockIdx.y * gridDim.x + blockIdx.x) * kBlock + threadIdx.x;
    return TravStack{s_stack + threadIdx.x, P.spill + gtid, P.spill_stride};
}

// Closest hit (kAny = false) or any hit within (1e-6, tmax) (kAny = true). While-while loop:
// a lane that reaches a leaf parks it and keeps descending until every lane still walking
// internal nodes holds a parked leaf, then the parked leaves are tested together. Leaves are
// still tested in the order they are reached, so the result equals the plain loop's.
template <bool kAny>
__device__ Hit traverse(const TraceParams& P, const Ray& r, float tmax, const TravStack& stk) {
    constexpr int kNone = INT_MIN;  // stack exhausted
    Hit h{-1, tmax, 0.f, 0.f};
    int node = 0, leaf = 0, sp = 0;
    while (node != kNone || leaf != 0) {
        while (node >= 0) {
            const TNode n = P.tnodes[node];
            float ta, tb;
            const bool ha = slab(r, n.b0xy.x, n.b0xy.y, n.b0xy.z, n.b0xy.w, n.bz.x, n.bz.y, h.t, ta);
            const bool hb = slab(r, n.b1xy.x, n.b1xy.y, n.b1xy.z, n.b1xy.w, n.bz.z, n.bz.w, h.t, tb);
            if (ha && hb) {
                int nearc = n.ch.x, farc = n.ch.y;
                if (tb < ta) { nearc = n.ch.y; farc = n.ch.x; }
                stk.push(sp, farc);
                ++sp;
                node = nearc;
            } else if (ha) {
                node = n.ch.x;
            } else if (hb) {
                node = n.ch.y;
            } else {
                node = sp > 0 ? stk.pop(--sp) : kNone;
            }
            if (node < 0 && node != kNone && leaf == 0) {
                leaf = node;
                node = sp > 0 ? stk.pop(--sp) : kNone;
            }
            if (__ballot(leaf == 0) == 0) break;
        }
        while (leaf != 0 || (node < 0 && node != kNone)) {
            if (leaf == 0) {
                leaf = node;
                node = sp > 0 ? stk.pop(--sp) : kNone;
            }
            int ti = ~leaf;
            leaf = 0;
            for (;;) {
                const TriV t = P.triv[ti];
                float tt, uu, vv;
                if (tri_hit(r, t, h.t, tt, uu, vv)) {
                    h.tri = ti; h.t = tt; h.u = uu; h.v = vv;
                    if (kAny) return h;
                }
                if (__float_as_int(t.v0.w) != 0) break;
                ++ti;
            }
        }
    }
    return h;
}

// shade_direct (src/pathtracer.cu:230-272): lights in order sun, points; the colour is
// accumulated only for unoccluded lights.
__device__ V3 shade_direct(const TraceParams& P, V3 Pt, V3 N, V3 albedo, const TravStack& stk) {
    V3 color = v3(0.f, 0.f, 0.f);
    for (int li = 0; li <= P.n_points; ++li) {
        Ray shadow;
        float tmax;
        V3 contrib;
        if (!nee_candidate(P, li, Pt, N, albedo, shadow, tmax, contrib)) continue;
        if (traverse<true>(P, shadow, tmax, stk).tri >= 0) continue;
        color = add(color, contrib);
    }
    return color;
}

// trace_path (src/pathtracer.cu:295-331).
__device__ V3 trace_path(const TraceParams& P, Ray ray, Rng& rng, const TravStack& stk) {
    V3 L = v3(0.f, 0.f, 0.f);
    V3 T = v3(1.f, 1.f, 1.f);
    const bool rr = (P.flags & LGI_RENDER_NO_RR) == 0;
    for (int bounce = 0; bounce <= P.bounces; ++bounce) {
        Hit h = traverse<false>(P, ray, 1e30f, stk);
        if (h.tri < 0) break;
        int mat;
        float tu, tv;
        V3 N = hit_normal(P.tris, h, mat, tu, tv);
        if (dot(N, ray.d) > 0.0f) N = mul(N, -1.0f);
        V3 Pt = add(ray.o, mul(ray.d, h.t));
        V3 albedo = sample_albedo(P, mat, tu, tv);
        L = add(L, had(T, shade_direct(P, Pt, N, albedo, stk)));
        if (bounce == P.bounces) break;
        if (bounce >= 1 && rr) {
            float q = fminf(0.95f, fmaxf(T.x, fmaxf(T.y, T.z)));
            if (rng.next() > q) break;
            T = mul(T, lgi_div(1.0f, q));
        }
        V3 nd = sample_cos_hemi(N, rng);
        ray = make_ray(add(Pt, mul(N, 1e-3f)), nd);
        T = had(T, albedo);
    }
    return L;
}

// Pixel of lane `tid` in block (bx, by): 16x16 pixels per block, 8x8 per wave.
__device__ __forceinline__ void tile_pixel(int bx, int by, int tid, int& px, int& py) {
    const int lane = tid & 63, wave = tid >> 6;
    px = bx * 16 + (wave & 1) * 8 + (lane & 7);
    py = by * 16 + (wave >> 1) * 8 + (lane >> 3);
}

// accumulate_kernel (src/pathtracer.cu:333-364): spp samples per pixel, Kahan-compensated.
// Samples [sample0, sample0 + spp) of pixel (px, py) through `cam`, Kahan-added to the
// pixel's sum/comp (and squared luminance to sumsq) in sample order.
__device__ void render_pixel(const TraceParams& P, const DevCamera& cam, int px, int py, int spp, int sample0,
                             float* sum, float* comp, float* sumsq, const TravStack& stk) {
    const uint32_t pix = (uint32_t)py * (uint32_t)cam.width + (uint32_t)px;
    // LGI_RENDER_SET_SUMS: start from zero instead of the buffers' contents
    const bool add = (P.flags & LGI_RENDER_SET_SUMS) == 0;
    float s0 = add ? sum[3 * pix] : 0.f, s1 = add ? sum[3 * pix + 1] : 0.f, s2 = add ? sum[3 * pix + 2] : 0.f;
    float c0 = add ? comp[3 * pix] : 0.f, c1 = add ? comp[3 * pix + 1] : 0.f, c2 = add ? comp[3 * pix + 2] : 0.f;
    float sq = (add && sumsq) ? sumsq[pix] : 0.0f;
    const bool xs = (P.flags & LGI_RNG_XORSHIFT_REF) != 0;
    Rng rng;
    rng.pcg = !xs;
    for (int k = 0; k < spp; ++k) {
        const int sample = sample0 + k;
        if (xs) {
            // src/pathtracer.cu:349 with 8-sample launch chunks (:699-702)
            if (k == 0 || (sample & 7) == 0) {
                const uint32_t base = (uint32_t)(sample - (sample & 7));
                rng.xs = ((uint32_t)px * 1973u + (uint32_t)py * 9277u + base * 26699u) | 1u;
            }
        } else {
            pcg_for_sample(rng.g, P.seed, pix, (uint32_t)sample);
        }
        float jx = rng.next();
        float jy = rng.next();
        Ray r = primary_ray(cam, px, py, jx, jy);
        V3 L = trace_path(P, r, rng, stk);
        kahan_add(s0, c0, L.x);
        kahan_add(s1, c1, L.y);
        kahan_add(s2, c2, L.z);
        float lum = 0.2126f * L.x + 0.7152f * L.y + 0.0722f * L.z;
        sq += lum * lum;
    }
    sum[3 * pix] = s0; sum[3 * pix + 1] = s1; sum[3 * pix + 2] = s2;
    comp[3 * pix] = c0; comp[3 * pix + 1] = c1; comp[3 * pix + 2] = c2;
    if (sumsq) sumsq[pix] = sq;
}

// The megakernels ask for 5 waves per SIMD (at most 96 VGPRs, against 111-121 unconstrained):
// a few spills outside the traversal loop buy a fifth wave to hide its load latency. Measured
// +2% on the 1.05M-triangle frame and +4.6% on the 50-pose queue; 6 waves spill more and lose.
constexpr int kRenderWavesPerSimd = 5;

__global__ __launch_bounds__(kBlock, kRenderWavesPerSimd) void k_render(TraceParams P) {
    __shared__ int s_stack[kStackLds * kBlock];
    int px, py;
    tile_pixel(blockIdx.x, blockIdx.y, threadIdx.x, px, py);
    if (px >= P.cam.width || py >= P.cam.height) return;
    render_pixel(P, P.cam, px, py, P.spp, P.sample0, P.sum, P.comp, P.sumsq, make_stack(P, s_stack));
}

// Persistent multi-pose renderer: every wave takes (pose, 8x8 tile, spp, first sample) work
// items from a device queue (one atomic per item) until the queue is empty, so poses with
// uneven spp budgets share one launch and no wave idles while items remain. Per-pose
// accumulators are [pose][W*H*3] (sum, comp) and [pose][W*H] (sumsq).
// kAllTiles: no item list; item i is every pixel of tile i of the one image P.cam describes,
// with P.spp samples from P.sample0 (the whole-image render of lgi_render_accumulate_noalloc).
template <bool kAllTiles>
__global__ __launch_bounds__(kBlock, kRenderWavesPerSimd) void k_render_queue(TraceParams P, const DevCamera* cams,
                                                         const lgi_work_item* items, int n_items, uint32_t* counter) {
    __shared__ int s_stack[kStackLds * kBlock];
    const TravStack stk{s_stack + threadIdx.x, P.spill + blockIdx.x * kBlock + threadIdx.x, P.spill_stride};
    const int lane = threadIdx.x & 63;
    const int W = P.cam.width, H = P.cam.height;
    const int tiles_x = (W + 7) / 8;
    const int64_t npx = (int64_t)W * H;
    for (;;) {
        uint32_t idx = 0;
        if (lane == 0) idx = atomicAdd(counter, 1u);
        idx = __shfl(idx, 0, 64);
        if (idx >= (uint32_t)n_items) break;
        const lgi_work_item it = kAllTiles ? lgi_work_item{0, (int32_t)idx, P.spp, P.sample0} : items[idx];
        const int px = (it.tile % tiles_x) * 8 + (lane & 7), py = (it.tile / tiles_x) * 8 + (lane >> 3);
        if (px < W && py < H)
            render_pixel(P, kAllTiles ? P.cam : cams[it.pose], px, py, it.spp, it.sample0, P.sum + it.pose * npx * 3,
                         P.comp + it.pose * npx * 3, P.sumsq ? P.sumsq + it.pose * npx : nullptr, stk);
    }
}

// Zeroes the queue counter inside a captured graph (no memset nodes in replayed graphs).
__global__ void k_reset_counter(uint32_t* c) { *c = 0u; }

// Pixel-centre primary hits: original triangle index and t (src/pathtracer.cu:381-384).
__global__ __launch_bounds__(kBlock) void k_primary(TraceParams P, int32_t* tri_out, float* t_out) {
    __shared__ int s_stack[kStackLds * kBlock];
    int px, py;
    tile_pixel(blockIdx.x, blockIdx.y, threadIdx.x, px, py);
    if (px >= P.cam.width || py >= P.cam.height) return;
    Ray r = primary_ray(P.cam, px, py, 0.5f, 0.5f);
    Hit h = traverse<false>(P, r, 1e30f, make_stack(P, s_stack));
    const size_t i = (size_t)py * P.cam.width + px;
    tri_out[i] = h.tri < 0 ? -1 : __float_as_int(P.triv[h.tri].e1.w);
    t_out[i] = h.tri < 0 ? -1.0f : h.t;
}

// gbuffer_kernel (src/pathtracer.cu:366-411).
__global__ __launch_bounds__(kBlock) void k_gbuffer(TraceParams P, float3 bbMin, float3 invExt, uint8_t* alb,
                                                    uint8_t* nrm, uint8_t* wp) {
    __shared__ int s_stack[kStackLds * kBlock];
    int px, py;
    tile_pixel(blockIdx.x, blockIdx.y, threadIdx.x, px, py);
    if (px >= P.cam.width || py >= P.cam.height) return;
    Ray r = primary_ray(P.cam, px, py, 0.5f, 0.5f);
    Hit h = traverse<false>(P, r, 1e30f, make_stack(P, s_stack));
    const size_t i = ((size_t)py * P.cam.width + px) * 3;
    if (h.tri < 0) {
        alb[i] = alb[i + 1] = alb[i + 2] = 0;
        nrm[i] = 128; nrm[i + 1] = 128; nrm[i + 2] = 255;
        wp[i] = wp[i + 1] = wp[i + 2] = 0;
        return;
    }
    int mat;
    float tu, tv;
    V3 N = hit_normal(P.tris, h, mat, tu, tv);
    if (dot(N, r.d) > 0.0f) N = mul(N, -1.0f);
    auto enc = [](float v) { return (uint8_t)(fminf(1.0f, fmaxf(0.0f, v)) * 255.0f); };
    const V3 a = sample_albedo(P, mat, tu, tv);
    alb[i] = enc(a.x);
    alb[i + 1] = enc(a.y);
    alb[i + 2] = enc(a.z);
    nrm[i] = enc(N.x * 0.5f + 0.5f);
    nrm[i + 1] = enc(N.y * 0.5f + 0.5f);
    nrm[i + 2] = enc(N.z * 0.5f + 0.5f);
    V3 Pt = add(r.o, mul(r.d, h.t));
    wp[i] = enc((Pt.x - bbMin.x) * invExt.x);
    wp[i + 1] = enc((Pt.y - bbMin.y) * invExt.y);
    wp[i + 2] = enc((Pt.z - bbMin.z) * invExt.z);
}

// tonemap_kernel (src/pathtracer.cu:413-429), one byte per thread.
__global__ __launch_bounds__(kBlock) void k_tonemap(const float* sum, int64_t n, float inv_spp, uint8_t* out) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    float v = lgi_sqrt(fmaxf(0.0f, fminf(sum[i] * inv_spp, 1.0f)));
    out[i] = (uint8_t)(v * 255.0f);
}

}  // namespace

// ------------------------------------------------------------------ host launchers
void lgi_launch_render(const TraceParams& P, hipStream_t st) {
    dim3 grid((P.cam.width + 15) / 16, (P.cam.height + 15) / 16);
    hipLaunchKernelGGL(k_render, grid, dim3(kBlock), 0, st, P);
}
void lgi_launch_render_queue(const TraceParams& P, const DevCamera* cams, const lgi_work_item* items, int n_items,
                             uint32_t* counter, int grid, hipStream_t st) {
    hipLaunchKernelGGL(k_reset_counter, dim3(1), dim3(1), 0, st, counter);
    hipLaunchKernelGGL(k_render_queue<false>, dim3(grid), dim3(kBlock), 0, st, P, cams, items, n_items, counter);
}

// ------------------------------------------------------------- allocation-free whole-image render
// Scratch of lgi_render_accumulate_noalloc, one copy per device for the life of the process:
// the queue counter (the first 64 ints, 256 B) and the traversal-stack spill slots of a
// persistent grid of kNoAllocGrid blocks (kStackSpill entries per thread beyond the LDS part).
// Host code never passes `scratch` itself to a kernel: its name is a host-side stand-in, so
// lgi_noalloc_scratch resolves each device's address with hipGetSymbolAddress once.
constexpr int kNoAllocGrid = 1024;  // 4 blocks of 256 threads per CU
constexpr int64_t kNoAllocCounterInts = 64;
constexpr int64_t kNoAllocScratchInts = kNoAllocCounterInts + (int64_t)kStackSpill * kNoAllocGrid * kBlock;
static __device__ __align__(256) int scratch[kNoAllocScratchInts];
static_assert(kNoAllocScratchInts * 4 == LGI_NOALLOC_SCRATCH_BYTES, "documented scratch size");

lgi_status lgi_noalloc_scratch(int device, int** out) {
    constexpr int kMaxDevices = 64;
    static std::mutex m;
    static int* ptrs[kMaxDevices] = {};
    if (!out || device < 0 || device >= kMaxDevices) return LGI_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lock(m);
    if (!ptrs[device]) {
        int prev = 0;
        if (hipGetDevice(&prev) != hipSuccess) return LGI_ERR_HIP;
        if (hipSetDevice(device) != hipSuccess) return LGI_ERR_HIP;
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        (void)hipSetDevice(prev);
        if (e != hipSuccess || !p) return LGI_ERR_HIP;
        ptrs[device] = (int*)p;
    }
    *out = ptrs[device];
    return LGI_OK;
}

void lgi_launch_render_all_tiles_noalloc(TraceParams P, int* scratch_dev, hipStream_t st) {
    const int n_tiles = ((P.cam.width + 7) / 8) * ((P.cam.height + 7) / 8);
    const int grid = std::max(1, std::min(kNoAllocGrid, n_tiles));
    uint32_t* counter = (uint32_t*)scratch_dev;
    P.spill = scratch_dev + kNoAllocCounterInts;
    P.spill_stride = grid * kBlock;
    hipLaunchKernelGGL(k_reset_counter, dim3(1), dim3(1), 0, st, counter);
    hipLaunchKernelGGL(k_render_queue<true>, dim3(grid), dim3(kBlock), 0, st, P, nullptr, nullptr, n_tiles, counter);
}
void lgi_launch_primary(const TraceParams& P, int32_t* tri, float* t, hipStream_t st) {
    dim3 grid((P.cam.width + 15) / 16, (P.cam.height + 15) / 16);
    hipLaunchKernelGGL(k_primary, grid, dim3(kBlock), 0, st, P, tri, t);
}
void lgi_launch_gbuffer(const TraceParams& P, float3 bbMin, float3 invExt, uint8_t* alb, uint8_t* nrm, uint8_t* wp,
                        hipStream_t st) {
    dim3 grid((P.cam.width + 15) / 16, (P.cam.height + 15) / 16);
    hipLaunchKernelGGL(k_gbuffer, grid, dim3(kBlock), 0, st, P, bbMin, invExt, alb, nrm, wp);
}
void lgi_launch_tonemap(const float* sum, int64_t n, float inv_spp, uint8_t* out, hipStream_t st) {
    hipLaunchKernelGGL(k_tonemap, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, st, sum, n, inv_spp,
                       out);
}
