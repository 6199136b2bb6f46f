// trace_kernels.h — parameter block and launchers of the tracer kernels (trace.hip).
#pragma once
#include "lgi_internal.h"

struct TraceParams {
    const TNode* tnodes;
    const TriV* triv;
    const TriS* tris;
    const float* mat_albedo;
    const uint32_t* tex_texels;  // RGBA8 texels (null when no material is textured)
    const int4* tex_desc;        // per material: texel offset, width, height, textured
    const float* srgb_lut;       // 256 entries
    const lgi_point_light* points;
    int32_t n_points;
    float sun_dir[3];
    float sun_color[3];
    float sun_intensity;
    DevCamera cam;
    int32_t spp;
    int32_t bounces;
    int32_t sample0;
    uint32_t flags;
    uint32_t seed;
    float* sum;     // W*H*3
    float* comp;    // W*H*3 Kahan compensation
    float* sumsq;   // W*H (nullable)
    int* spill;     // traversal-stack spill slots, [depth][thread]
    int32_t spill_stride;
};

// Traversal stack split: entries [0, kStackLds) in LDS, [kStackLds, 32) in global spill slots.
// Spill areas are sized for kStackMax - kStackLdsMin entries so that kernels holding fewer
// stack entries in LDS (to make room for the top BVH levels) fit the same allocation.
constexpr int kStackMax = 32;
constexpr int kStackLds = 16;
constexpr int kStackLdsMin = 8;
constexpr int kStackSpill = kStackMax - kStackLdsMin;

// Threads of the pixel-grid kernels (16x16-pixel blocks) for a W x H image.
inline int64_t lgi_grid_threads(int w, int h) { return (int64_t)((w + 15) / 16) * ((h + 15) / 16) * 256; }

void lgi_launch_render(const TraceParams& P, hipStream_t st);
void lgi_launch_primary(const TraceParams& P, int32_t* tri, float* t, hipStream_t st);
void lgi_launch_gbuffer(const TraceParams& P, float3 bbMin, float3 invExt, uint8_t* alb, uint8_t* nrm, uint8_t* wp,
                        hipStream_t st);
void lgi_launch_tonemap(const float* sum, int64_t n, float inv_spp, uint8_t* out, hipStream_t st);
// Persistent queue renderer over (pose, 8x8 tile, spp, sample0) items; resets the counter first.
void lgi_launch_render_queue(const TraceParams& P, const DevCamera* cams, const lgi_work_item* items, int n_items,
                             uint32_t* counter, int grid, hipStream_t st);
// Allocation-free whole-image render (lgi_render_accumulate_noalloc): the device address of this
// device's static scratch array (resolved once per device), and the persistent queue launch over
// every 8x8 tile of P.cam's image that uses it for its counter and stack spill slots.
lgi_status lgi_noalloc_scratch(int device, int** out);
void lgi_launch_render_all_tiles_noalloc(TraceParams P, int* scratch_dev, hipStream_t st);
// Sorted wavefront renderer (wavefront.hip); bb = scene box used by the ray sort key.
// d_tiles (device, n_tiles distinct row-major 8x8 tile ids) restricts the render to those tiles.
lgi_status lgi_render_wavefront(lgi_ctx* ctx, const TraceParams& P, float3 bb_min, float3 bb_max,
                                const int32_t* d_tiles = nullptr, int n_tiles = 0);
