// lgi_internal.h — host-side context and device data layouts shared by the HIP
// translation units of liblgi.so.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "lgi.h"
#include "scene_host.h"

// ------------------------------------------------------------------ device layouts
// Traversal node: both children's boxes live in the parent (64 B, one 64-B line), so a
// wave fetches one node per step and tests two boxes. Child refs: >= 0 internal node
// index; < 0 leaf starting at triangle ~ref (leaf triangles end at the TriV end flag).
struct alignas(16) TNode {
    float4 b0xy;   // child0 lo.x, hi.x, lo.y, hi.y
    float4 b1xy;   // child1 lo.x, hi.x, lo.y, hi.y
    float4 bz;     // child0 lo.z, hi.z, child1 lo.z, hi.z
    int4 ch;       // child0 ref, child1 ref, unused, unused
};
static_assert(sizeof(TNode) == 64, "TNode is one 64-byte line");

// Intersection-only triangle record in leaf order: v0, e1 = v1 - v0, e2 = v2 - v0 (the
// exact FP32 differences ray_tri forms, src/pathtracer.cu:86-88). v0.w holds the
// end-of-leaf flag bit, e1.w the original scene index (as int bits).
struct alignas(16) TriV {
    float4 v0;
    float4 e1;
    float4 e2;
};
static_assert(sizeof(TriV) == 48, "TriV");

// Shading record in leaf order: vertex normals, material index and vertex texture
// coordinates (one 64-byte line, read once per closest hit).
struct TriS {
    float n[9];
    int32_t mat;
    float uv[6];
};
static_assert(sizeof(TriS) == 64, "TriS");

struct DevLights {
    float sun_dir[3];
    float sun_color[3];
    float sun_intensity;
    int32_t n_points;
    const lgi_point_light* points;   // device array
};

struct DevCamera {
    float pos[3], fwd[3], left[3], up[3];
    float tan_half_fov, aspect;
    int32_t width, height;
};

// Device-resident scene and acceleration structure (kept across renders and batch jobs).
struct DevScene {
    int64_t n_tris = 0;
    int64_t n_nodes = 0;        // BvhNode count (reference format)
    int64_t n_tnodes = 0;       // traversal nodes
    lgi_triangle* tri_in = nullptr;   // original order, AoS 100 B
    lgi_bvh_node* nodes = nullptr;    // BFS order, reference format
    TNode* tnodes = nullptr;
    const lgi_bvh_node* trav_nodes = nullptr;  // tree the traversal layout was made from (not owned)
    TriV* triv = nullptr;
    TriS* tris = nullptr;
    float* mat_albedo = nullptr;      // 3 per material
    uint32_t* tex_texels = nullptr;   // RGBA8 texels of every loaded texture, one after another
    int4* tex_desc = nullptr;         // per material: texel offset, width, height, 1 if textured
    float* srgb_lut = nullptr;        // 8-bit sRGB -> linear
    int32_t tex_loaded = 0, tex_missing = 0;
    lgi_point_light* points = nullptr;
    int32_t n_points = 0;
    int32_t* perm = nullptr;          // sorted position -> original index
    uint32_t* keys = nullptr;         // sorted Morton keys
    int32_t root_is_leaf = 0;
    int32_t n_mats = 0;
};

struct lgi_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    HostScene scene;             // host copy (original order until a build permutes it)
    bool gpu_ready = false;      // device selected, events created
    bool uploaded = false;       // dev mirrors `scene`
    DevScene dev;
    lgi_bvh_node root{};         // host copy of the BVH root (scene box)
    bool bvh_ready = false;
    float last_build_ms = 0.f;
    int32_t bvh_treelet_passes = 2;  // restructuring passes of the traversal tree (0: reference tree)
    float last_render_ms = 0.f;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipEvent_t ev_tree0 = nullptr, ev_tree1 = nullptr;  // around the traversal-tree restructuring
    float last_treelet_ms = 0.f;
    // event pairs around each traversal-kernel launch of the last wavefront render
    std::vector<hipEvent_t> kev;
    int kev_used = 0;           // pairs recorded by the last render (0: inline renderer)
    bool kev_overflow = false;  // more launches than pairs: the sum covers the first kev.size()/2
    int* noalloc_scratch = nullptr;  // device address of this device's lgi_render_accumulate_noalloc scratch
    void* cap_stream = nullptr;  // hipStream_t used for graph capture
    void* graph_exec = nullptr;  // hipGraphExec_t of the last captured batch launch
    std::string graph_key;       // inputs the captured graph was built for
    std::string asset_root;      // texture paths are relative to this directory
    bool asset_root_set = false; // set explicitly (lgi_set_asset_root)
    std::string scene_root;      // default root: the loaded scene file's grandparent directory
    bool tex_dirty = false;      // asset root changed after the upload
    // scratch buffers (grow-only)
    std::vector<std::pair<void*, size_t>> scratch;
    void* scratch_get(int slot, size_t bytes);
};

// ------------------------------------------------------------------ helpers
lgi_status lgi_set_err(lgi_ctx* ctx, lgi_status st, const std::string& msg);

#define LGI_HIP(ctx, call)                                                                \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess)                                                             \
            return lgi_set_err((ctx), LGI_ERR_HIP,                                        \
                               std::string(#call) + ": " + hipGetErrorString(e_));        \
    } while (0)

// Kernel-side entry points implemented in the .hip files.
lgi_status lgi_gpu_build_bvh(lgi_ctx* ctx);
// Records the start (begin=true) or end event of the next traversal launch on ctx->stream.
lgi_status lgi_mark_trace(lgi_ctx* ctx, bool begin);
// Stable LSD radix sort of (keys, vals) on `bits` (multiple of 8) low key bits; in place when
// bits is a multiple of 16, else the result is left in the scratch buffers (the returned
// pointer tells). scratch must hold lgi_sort_scratch_bytes(n). With n_dev set, the key count
// is read on the device (at most n) and n only sizes the launches.
size_t lgi_sort_scratch_bytes(int64_t n);
// Returns the sorted values: vals itself when bits / 8 is even, else the scratch copy.
const int32_t* lgi_sort_pairs(hipStream_t st, uint32_t* keys, int32_t* vals, int n, int bits, void* scratch,
                              const uint32_t* n_dev = nullptr);
// In-place exclusive scan of n u32 values; *total (device) receives the sum.
size_t lgi_scan_scratch_bytes(int64_t n);
void lgi_exclusive_scan(hipStream_t st, uint32_t* data, int n, uint32_t* total, void* scratch);
// The same scan in two parts that a consumer adds itself: data[i] becomes the exclusive prefix
// within its kScanChunk-value chunk and chunk_sums[i / kScanChunk] the prefix of the chunks, so
// the full prefix is data[i] + chunk_sums[i / kScanChunk] (chunk_sums: lgi_scan_scratch_bytes).
// data is 16-byte aligned.
constexpr int kScanChunk = 4096;
void lgi_exclusive_scan_chunks(hipStream_t st, uint32_t* data, int n, uint32_t* total, uint32_t* chunk_sums);
void lgi_free_dev_scene(DevScene& d);
DevCamera lgi_make_dev_camera(const lgi_camera& c);
lgi_status lgi_upload_lights(lgi_ctx* ctx, DevLights* out);
