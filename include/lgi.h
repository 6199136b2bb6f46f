/*
 * lgi.h — C ABI of the MI355X-native LuminaGI reference tracer + denoiser (liblgi.so).
 *
 * Plain C: POD structs, pointers and sizes; no C++ or torch types cross this line.
 * Every entry point returns an lgi_status; nothing calls exit(). Use one context
 * per host thread. Device work is ordered on the context's HIP stream (the null
 * stream unless lgi_create/lgi_set_stream was given one).
 *
 * Reference interfaces replaced (paths relative to the reference repository):
 *   LoadSceneJSON            include/scene.h:58,  src/scene_loader.cpp:26-138
 *   BuildBVH                 include/bvh.h:26,    src/bvh.cpp:164-211
 *   RenderSceneCUDA          src/main.cpp:19-24,  src/pathtracer.cu:581-757
 *   RenderSceneCUDASorted    src/main.cpp:25-27,  src/pathtracer.cu:761-909
 *   overwrite_camera         src/main.cpp:29-59
 *   read_batch_file          src/main.cpp:61-91
 *   tonemap_kernel           src/pathtracer.cu:413-429
 *   gbuffer_kernel           src/pathtracer.cu:366-411
 *   UNet layers (PyTorch)    ml/train.py:37-75 (conv3x3/1x1, pool, upsample, concat)
 */
#ifndef LGI_H
#define LGI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LGI_API __attribute__((visibility("default")))

typedef enum lgi_status {
    LGI_OK              = 0,
    LGI_ERR_INVALID_ARG = 1,  /* null pointer, size out of range, bad flag */
    LGI_ERR_IO          = 2,  /* file cannot be opened / written */
    LGI_ERR_PARSE       = 3,  /* malformed JSON or missing key (reference throws, we return) */
    LGI_ERR_VERSION     = 4,  /* scene.json "version" != 1 */
    LGI_ERR_HIP         = 5,  /* HIP runtime error (message in lgi_last_error) */
    LGI_ERR_NO_SCENE    = 6,  /* render/build before a scene was set */
    LGI_ERR_NO_BVH      = 7,  /* render before lgi_bvh_build */
    LGI_ERR_UNSUPPORTED = 8   /* request outside what the library implements */
} lgi_status;

typedef struct lgi_ctx lgi_ctx;

/* ---- POD mirrors of the reference data model (include/scene.h:6-56, include/bvh.h:8-16) ---- */

typedef struct lgi_camera {          /* CameraDesc, include/scene.h:6-16 */
    float   pos[3];
    float   camera_to_world[16];     /* rows 0/1/2 = forward/left/up (src/pathtracer.cu:656-659) */
    float   fov_y_deg;
    float   aspect;
    float   near_z;
    float   far_z;
    int32_t width;
    int32_t height;
} lgi_camera;

typedef struct lgi_dir_light {       /* DirectionalLight, 28 B */
    float dir[3];                    /* travel direction; direction to light = -dir */
    float color[3];
    float intensity;
} lgi_dir_light;

typedef struct lgi_point_light {     /* PointLight, 32 B */
    float pos[3];
    float color[3];
    float intensity;
    float radius;
} lgi_point_light;

typedef struct lgi_material {        /* Material without the texture path */
    float albedo[3];
    float emissive[3];
} lgi_material;

typedef struct lgi_triangle {        /* Triangle, 100 B, byte-identical layout */
    float   v[3][3];
    float   n[3][3];
    float   uv[3][2];
    int32_t material;
} lgi_triangle;

typedef struct lgi_bvh_node {        /* BvhNode, 40 B, byte-identical layout */
    float   bb_min[3];
    float   bb_max[3];
    int32_t left;                    /* -1 for a leaf */
    int32_t right;                   /* -1 for a leaf */
    int32_t first_tri;               /* leaf: first triangle in the permuted order; internal: -1 */
    int32_t tri_count;               /* 0 internal, >0 leaf */
} lgi_bvh_node;

typedef struct lgi_batch_job {       /* BatchJob, src/main.cpp:63-69 */
    char    out_png[512];
    int32_t spp;
    int32_t bounces;
    float   cam_pos[3];
    float   cam_tgt[3];
} lgi_batch_job;

/* ---- render flags ---- */
#define LGI_RNG_PCG32          0u   /* PCG32 per (pixel, sample) stream (default) */
#define LGI_RNG_XORSHIFT_REF   1u   /* reference xorshift32 chunked seeding (src/pathtracer.cu:349) */
#define LGI_RENDER_SORTED      2u   /* wavefront path: radix-sort rays between bounces */
#define LGI_RENDER_NO_RR       4u   /* disable Russian roulette (reference sorted path semantics) */
#define LGI_RENDER_NO_RAY_SORT 8u   /* with LGI_RENDER_SORTED: no key sort between bounces, the
                                       continuations keep their compaction order (same image bits;
                                       no reference counterpart) */
#define LGI_RENDER_SET_SUMS   16u   /* lgi_render_accumulate(_noalloc): the call's sums replace the
                                       contents of d_sum, d_comp and d_sumsq (as if they held
                                       zeros) instead of adding to them, so that a frame needs no
                                       clearing pass; same bits as zeroed buffers. The batch and
                                       tile renders reject it (no reference counterpart) */

/* ---- context ---- */
LGI_API lgi_status  lgi_create(lgi_ctx** out, int device, void* hip_stream);
LGI_API void        lgi_destroy(lgi_ctx* ctx);
LGI_API const char* lgi_last_error(const lgi_ctx* ctx);
LGI_API lgi_status  lgi_set_stream(lgi_ctx* ctx, void* hip_stream);
LGI_API const char* lgi_version(void);

/* ---- scene (LoadSceneJSON, include/scene.h:58) ---- */
LGI_API lgi_status lgi_scene_load_json(lgi_ctx* ctx, const char* path, int verbose);
LGI_API lgi_status lgi_scene_save_json(lgi_ctx* ctx, const char* path);
/* Binary scene file: the same scene model as bulk little-endian records (64-byte header,
 * camera, sun, lights, materials with texture paths, 100-byte triangles). No reference
 * counterpart: the reference reads only scene.json, whose DOM parse takes 12 s for 300 MB
 * (README.md:46; binary format listed as a TODO at README.md:196). Errors: LGI_ERR_IO,
 * LGI_ERR_PARSE (bad magic, truncated or inconsistent sizes), LGI_ERR_VERSION. */
LGI_API lgi_status lgi_scene_load_bin(lgi_ctx* ctx, const char* path, int verbose);
LGI_API lgi_status lgi_scene_save_bin(lgi_ctx* ctx, const char* path);
LGI_API lgi_status lgi_scene_set(lgi_ctx* ctx, const lgi_camera* cam, const lgi_dir_light* sun,
                                 const lgi_dir_light* dirs, int32_t n_dirs,
                                 const lgi_point_light* points, int32_t n_points,
                                 const lgi_material* mats, int32_t n_mats,
                                 const lgi_triangle* tris, int64_t n_tris);
/* Procedural scenes: kind 1 = ~65k-triangle test scene, kind 2 = 62 meshes / ~1.05M triangles. */
LGI_API lgi_status lgi_scene_generate(lgi_ctx* ctx, int32_t kind, uint64_t seed,
                                      int32_t width, int32_t height);
LGI_API lgi_status lgi_scene_counts(const lgi_ctx* ctx, int64_t* n_tris, int32_t* n_mats,
                                    int32_t* n_points, int32_t* n_dirs);
LGI_API lgi_status lgi_scene_get(const lgi_ctx* ctx, lgi_camera* cam, lgi_dir_light* sun,
                                 lgi_point_light* points, lgi_material* mats,
                                 lgi_triangle* tris /* current order (permuted after a build) */);
LGI_API lgi_status lgi_scene_set_camera(lgi_ctx* ctx, const lgi_camera* cam);
/* Directory the materials' diffuse_tex paths are relative to (the reference's --asset-root,
 * src/main.cpp:118-141; default: the loaded scene file's grandparent directory). Null restores
 * the default. Textures (PNG) load with the scene upload and reload before the next render
 * after a change; unreadable ones are reported and leave the material with its constant albedo
 * (src/pathtracer.cu:594-612). */
LGI_API lgi_status lgi_set_asset_root(lgi_ctx* ctx, const char* dir);
/* overwrite_camera, src/main.cpp:29-59 (no-op when |target-pos| < 1e-8). */
LGI_API lgi_status lgi_camera_look_at(lgi_camera* cam, const float pos[3], const float target[3]);

/* ---- BVH (BuildBVH, include/bvh.h:26) ---- */
/* GPU LBVH. Like BuildBVH it permutes the context's triangles into Morton order. */
LGI_API lgi_status lgi_bvh_build(lgi_ctx* ctx);
LGI_API lgi_status lgi_bvh_node_count(const lgi_ctx* ctx, int64_t* n_nodes);
/* nodes: n_nodes entries in BFS order (root 0); tri_perm[i] = original index of the
 * triangle now at position i; keys: sorted 30-bit Morton codes. Any output may be null. */
LGI_API lgi_status lgi_bvh_get(lgi_ctx* ctx, lgi_bvh_node* nodes, int32_t* tri_perm, uint32_t* keys);
/* Traversal tree (no reference counterpart). The tracers traverse a copy of the reference tree
 * whose topology above the leaves is restructured to lower its surface-area cost (5-leaf
 * treelets, Karras & Aila 2013): same leaves and triangle order, every box the exact union of
 * its children's, so the closest hit equals the reference tree's except where visit order
 * decides it (exact distance ties, boxes grazed within rounding). passes 0..8 (default 2; 0 =
 * traverse the reference tree itself) takes effect at the next lgi_bvh_build. */
LGI_API lgi_status lgi_bvh_set_treelet_passes(lgi_ctx* ctx, int32_t passes);
/* nodes: n_nodes entries, the traversal tree in the slots of lgi_bvh_get's array (root 0). */
LGI_API lgi_status lgi_bvh_get_traversal(lgi_ctx* ctx, lgi_bvh_node* nodes);
/* Milliseconds of the last GPU build (device time, HIP events), restructuring included. */
LGI_API lgi_status lgi_bvh_last_build_ms(const lgi_ctx* ctx, float* ms);
/* Milliseconds of the last build's traversal-tree restructuring (0 when it did not run). */
LGI_API lgi_status lgi_bvh_last_treelet_ms(const lgi_ctx* ctx, float* ms);

/* ---- rendering (RenderSceneCUDA / RenderSceneCUDASorted) ---- */
/* Host-buffer render of the scene camera (or *cam when non-null): rgb = W*H*3 u8
 * sqrt-tonemapped; hdr (optional) = W*H*3 f32 mean radiance; alb/nrm/wp (all three or
 * none) = G-buffer encodings. */
LGI_API lgi_status lgi_render(lgi_ctx* ctx, const lgi_camera* cam, int32_t spp, int32_t bounces,
                              uint32_t flags, uint8_t* rgb, float* hdr,
                              uint8_t* alb, uint8_t* nrm, uint8_t* wp);
/* Device-buffer render: adds spp samples (sample indices [sample0, sample0+spp)) into
 * d_sum (W*H*3 f32) with Kahan compensation d_comp (W*H*3 f32) and per-pixel sum of
 * squared luminance d_sumsq (W*H f32, may be null). Buffers are caller-owned device memory. */
LGI_API lgi_status lgi_render_accumulate(lgi_ctx* ctx, const lgi_camera* cam, int32_t spp,
                                         int32_t bounces, uint32_t flags, int32_t sample0,
                                         float* d_sum, float* d_comp, float* d_sumsq);
/* lgi_render_accumulate without any device allocation. No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc). Same arguments and sums as lgi_render_accumulate for
 * flags without LGI_RENDER_SORTED (bit for bit): the whole image runs as one persistent queue
 * of 8x8 tiles whose counter and traversal-stack spill slots live in a static device array of
 * LGI_NOALLOC_SCRATCH_BYTES that each GPU the library runs on holds until the process ends; its
 * size does not depend on the image. The first call on a device resolves that device's array
 * (hipGetSymbolAddress); no call allocates, frees or copies host memory. Calls that overlap on
 * one GPU (other streams or threads) share the array and race; calls ordered on one stream do
 * not. LGI_ERR_UNSUPPORTED for LGI_RENDER_SORTED (the wavefront renderer's queues take GBs) and
 * when textures must be reloaded first (lgi_set_asset_root since the last render). */
#define LGI_NOALLOC_SCRATCH_BYTES 25166080
LGI_API lgi_status lgi_render_accumulate_noalloc(lgi_ctx* ctx, const lgi_camera* cam, int32_t spp,
                                                 int32_t bounces, uint32_t flags, int32_t sample0,
                                                 float* d_sum, float* d_comp, float* d_sumsq);
/* Primary-hit triangle ids (original scene index, -1 on miss) and t at pixel centres. */
LGI_API lgi_status lgi_primary_hits(lgi_ctx* ctx, const lgi_camera* cam, int32_t* d_tri, float* d_t);
/* Device tonemap: u8 = (u8)(sqrt(clamp(sum*inv_spp,0,1))*255). */
LGI_API lgi_status lgi_tonemap(lgi_ctx* ctx, const float* d_sum, int32_t n_px, float inv_spp,
                               uint8_t* d_rgb);
/* Device G-buffer (albedo, normal, worldpos encodings) for camera cam. */
LGI_API lgi_status lgi_gbuffer(lgi_ctx* ctx, const lgi_camera* cam, uint8_t* d_alb,
                               uint8_t* d_nrm, uint8_t* d_wp);
/* Kernel time of the last render call (ms, HIP events). */
LGI_API lgi_status lgi_last_render_ms(const lgi_ctx* ctx, float* ms);
/* Measurement hook, no reference counterpart: summed HIP-event time of the BVH traversal
 * kernel launches of the last render call (for the sorted renderer every closest-hit and
 * shadow-ray launch; for the inline renderer the one megakernel launch) and their count. */
LGI_API lgi_status lgi_last_trace_ms(const lgi_ctx* ctx, float* ms, int32_t* launches);

/* Pass plan of the sorted (wavefront) renderer, host only; no reference counterpart (the
 * reference renders 8-sample chunks, src/pathtracer.cu:673-704). For n_slots queue slots per
 * sample (64 per 8x8 tile) and n_points point lights, *samples = the samples each pass of
 * LGI_RENDER_SORTED holds (<= spp): about 100M paths in flight, at most 3 x that many
 * shadow-ray candidates (slot x light, sun included) and never more than 2^31 - 1 of them.
 * LGI_ERR_UNSUPPORTED when one sample alone exceeds 2^31 - 1 candidates (the sorted renderer
 * then refuses the image with the same status). */
LGI_API lgi_status lgi_wavefront_samples_per_pass(int64_t n_slots, int32_t n_points, int32_t spp, int32_t* samples);

/* ---- many poses / selected tiles in one persistent launch ----
 * No reference counterpart beyond the --batch loop (src/main.cpp:158-189), which re-uploads the
 * scene and launches once per job. Work is split into (pose, 8x8-pixel tile, spp, first
 * sample) items on a device queue; one persistent kernel drains it. */
typedef struct lgi_work_item {
    int32_t pose;     /* index into the camera list */
    int32_t tile;     /* 8x8 tile, row-major over ceil(W/8) x ceil(H/8) */
    int32_t spp;
    int32_t sample0;  /* first sample index (PCG32 stream per (pixel, sample)) */
} lgi_work_item;
typedef struct lgi_pose_job {
    lgi_camera cam;   /* all poses share the first pose's width and height */
    int32_t spp;
    int32_t sample0;
} lgi_pose_job;
/* Renders every pose into d_sum/d_comp ([n][W*H*3] f32) and d_sumsq ([n][W*H], nullable).
 * use_graph: capture the launch in a hipGraph on first use and replay it for repeated calls
 * with the same jobs. */
LGI_API lgi_status lgi_render_batch(lgi_ctx* ctx, const lgi_pose_job* jobs, int32_t n, int32_t bounces,
                                    uint32_t flags, float* d_sum, float* d_comp, float* d_sumsq, int32_t use_graph);
/* Adds spp samples (from sample0) to the listed, distinct 8x8 tiles of one image (host tile
 * list). With LGI_RENDER_SORTED all samples of all listed tiles run in the wavefront renderer
 * at once; otherwise one persistent-queue item per tile (one wave per tile). */
LGI_API lgi_status lgi_render_tiles(lgi_ctx* ctx, const lgi_camera* cam, const int32_t* tiles, int32_t n_tiles,
                                    int32_t spp, int32_t sample0, int32_t bounces, uint32_t flags, float* d_sum,
                                    float* d_comp, float* d_sumsq);

/* ---- image metrics (tools/diff.py:25-29, 100-101), FP64 reductions on the GPU ----
 * d_a, d_b: h*w*3 u8 RGB; d_scratch: >= 3*1024 doubles; d_out (device, 3 doubles) receives
 * {MSE, SSIM (7x7 uniform window, sample covariance, K1 .01, K2 .03, range 255), mean |a-b|}. */
LGI_API lgi_status lgi_image_metrics(const uint8_t* d_a, const uint8_t* d_b, int32_t h, int32_t w,
                                     double* d_scratch, double* d_out, void* stream);
/* Difference heatmap (heatmap, tools/diff.py:41-45): per pixel, the mean absolute channel
 * difference (FP64) times gain, clipped to [0, 255] and truncated, indexes the colour table of
 * lgi_heat_lut. d_a, d_b, d_out: h*w*3 u8 RGB on the device. Asynchronous on stream. */
LGI_API lgi_status lgi_diff_heatmap(const uint8_t* d_a, const uint8_t* d_b, int32_t h, int32_t w, double gain,
                                    uint8_t* d_out, void* stream);
/* The heatmap's 256-entry RGB colour table (viridis_lut, tools/diff.py:32-38), host only. */
LGI_API lgi_status lgi_heat_lut(uint8_t* rgb768);

/* ---- memory-hierarchy probe (no counterpart in the reference: it measures the L2, MALL and
 * HBM read bandwidth the tracer's roofline is compared with) ----
 * Reads the first bytes (rounded down to 16) of d_buf reps times with `blocks` workgroups of
 * 256 threads; each XCD re-reads the same 1/8 of the buffer on every pass. d_buf 16-byte
 * aligned; d_sink: one device u32 (written only on a data-dependent condition). With ms
 * non-null the launch is bracketed by HIP events on `stream` and waited for, and ms receives
 * its duration; otherwise the call is asynchronous. */
LGI_API lgi_status lgi_probe_read_bw(const void* d_buf, uint64_t bytes, int32_t reps, int32_t blocks,
                                     uint32_t* d_sink, void* stream, float* ms);

/* ---- batch (read_batch_file, src/main.cpp:61-91) ---- */
LGI_API lgi_status lgi_batch_parse(const char* path, lgi_batch_job* jobs, int32_t cap, int32_t* n_jobs);

/* ---- image output (stbi_write_png call sites, src/main.cpp:176,214) ---- */
LGI_API lgi_status lgi_write_png(const char* path, const uint8_t* rgb, int32_t w, int32_t h);
/* PNG -> RGBA8 rows (stbi_load(..., 4), src/pathtracer.cu:557): non-interlaced, depth 8 or 16,
 * grey / grey+alpha / RGB / RGBA / palette. rgba may be null to query the size. */
LGI_API lgi_status lgi_read_png(const char* path, int32_t* w, int32_t* h, uint8_t* rgba);

/* ---- denoiser layers (UNet of ml/train.py:37-75; NHWC BF16 activations on device) ----
 * All pointers are device pointers; `stream` is a hipStream_t (null = default stream).
 * Channel counts of BF16 tensors must be multiples of 32. Weights: [cout][9][cin] BF16 for
 * the kernels, FP32 gradients in the same layout. Replaces the cuDNN/PyTorch calls of
 * ml/train.py:41-44 (Conv2d + activation), :62-64 (MaxPool2d, Upsample), :71-75 (cat, tail). */
/* y = relu?(conv3x3(cat(src0, src1)) + bias); src1 may be null (c1 ignored). */
LGI_API lgi_status lgi_unet_conv3x3_fwd(const void* src0, int32_t c0, const void* src1, int32_t c1,
                                        const void* w, const float* bias, int32_t cout, int32_t n,
                                        int32_t h, int32_t w_, int32_t relu, void* out, void* stream);
/* lgi_unet_conv3x3_fwd that also writes pool_out (n, h/2, w/2, cout) = MaxPool2d(2)(y), taken
 * from the output tile while it is still on chip (ml/train.py:41-44 followed by :62, the
 * encoder's second convolution and the pool after it). Bit-identical to lgi_unet_conv3x3_fwd
 * then lgi_unet_maxpool_fwd; h and w must be even, as there. */
LGI_API lgi_status lgi_unet_conv3x3_pool_fwd(const void* src0, int32_t c0, const void* src1, int32_t c1,
                                             const void* w, const float* bias, int32_t cout, int32_t n,
                                             int32_t h, int32_t w_, int32_t relu, void* out, void* pool_out,
                                             void* stream);
/* The denoiser's last convolution and its tail in one kernel (inference): y = relu?(conv3x3(src0) +
 * bias) with 32 output channels, then out (n, h, w, 3) FP32 = max(expm1(log1p(max(hdr, 0)) +
 * tail_w . y + tail_b), 0), tail_w [3][32], tail_b [3] (ml/train.py's last ConvBlock, tail and
 * residual, in the radiance domain). y is never written: the tail reads it from the output tile
 * on chip. Bit-identical to lgi_unet_conv3x3_fwd followed by lgi_unet_tail_radiance_fwd with the
 * same tiling. */
LGI_API lgi_status lgi_unet_conv3x3_tail_radiance_fwd(const void* src0, int32_t c0, const void* w,
                                                      const float* bias, int32_t n, int32_t h, int32_t w_,
                                                      int32_t relu, const float* tail_w, const float* tail_b,
                                                      const float* hdr, float* out, void* stream);
/* lgi_unet_conv3x3_fwd (or _pool_fwd with pool_out non-null) on FP8 E4M3 operands: the BF16
 * inputs are quantized on chip with the per-tensor act_scale, q = e4m3(clamp(x / act_scale, +-448))
 * rounded to nearest even; w_e4m3 holds the weights [cout][9][c0 + c1] quantized per output
 * channel (w ~= q * w_scale[co]); FP32 accumulation, then * act_scale * w_scale[co] + bias,
 * ReLU and BF16 output. c0 and c1 multiples of 64. The reference UNet runs in FP32 (ml/train.py);
 * this is the denoiser's reduced-precision inference path. */
LGI_API lgi_status lgi_unet_conv3x3_fp8_fwd(const void* src0, int32_t c0, const void* src1, int32_t c1,
                                            const void* w_e4m3, const float* w_scale, float act_scale,
                                            const float* bias, int32_t cout, int32_t n, int32_t h, int32_t w_,
                                            int32_t relu, void* out, void* pool_out, void* stream);
/* Decoder convolution for inference on the upsample it follows (ml/train.py:62-64, 71-73:
 * Upsample(2x, bilinear) + cat + Conv2d): out (n, 2h, 2w, cout) = relu?(conv3x3(cat(up2(lo), skip))
 * + bias) without the upsampled tensor. lo (n, h, w, cu) is the upsample's input, skip
 * (n, 2h, 2w, cs) (may be null with cs 0), w [cout][9][cu + cs] BF16 the layer's weights and
 * w_phase [4][cout][9][cu] BF16 the weights lgi_unet_up_phase_weights derives from them: the
 * convolution of the upsampled image equals, at each of the four output phases (row, column
 * parity), a 3x3 convolution of lo with mixed weights. Pixels whose window reaches the image
 * border are computed from the upsampled values themselves. Equals lgi_unet_upsample_fwd then
 * lgi_unet_conv3x3_fwd up to BF16 rounding (of the mixed weights instead of the upsampled
 * activations) and FP32 summation order. cu, cs, cout multiples of 32. */
LGI_API lgi_status lgi_unet_conv3x3_up_fwd(const void* lo, int32_t cu, const void* skip, int32_t cs,
                                           const void* w_phase, const void* w, const float* bias, int32_t cout,
                                           int32_t n, int32_t h, int32_t w_, int32_t relu, void* out, void* stream);
/* w_phase [4][cout][9][cu] BF16 of lgi_unet_conv3x3_up_fwd from the FP32 weights w [cout][9][cin]
 * (their first cu input channels): for phase 2 py + px, tap 3 dy + dx,
 * w_phase = sum_{ky,kx} M_py[dy][ky] M_px[dx][kx] w[3 ky + kx], M_0 = [[.75,.25,0],[.25,.75,.75],
 * [0,0,.25]], M_1 = [[.25,0,0],[.75,.75,.25],[0,.25,.75]], summed in FP32. */
LGI_API lgi_status lgi_unet_up_phase_weights(const float* w, int32_t cout, int32_t cin, int32_t cu, void* w_phase,
                                             void* stream);
/* First layer for inference: y = relu?(conv3x3(x) + bias) with x the FP32 3-channel NHWC
 * input (read directly, no padded copy), w [32][9][cin_w] BF16 of which channels 0..2 are used,
 * y (n, h, w, 32) BF16. Equals lgi_unet_conv3x3_fwd on lgi_unet_pack_input(x) up to FP32
 * summation order (the 27 products of a pixel are one MFMA row). */
LGI_API lgi_status lgi_unet_conv3x3_first_fwd(const float* x, const void* w, int32_t cin_w, const float* bias,
                                              int32_t n, int32_t h, int32_t w_, int32_t relu, void* out,
                                              void* stream);
/* lgi_unet_conv3x3_first_fwd on log1p(max(hdr, 0)) of the radiance image hdr, the transform
 * applied inside the layer's input read: the denoiser's input conversion of tools/denoise.py
 * (radiance in) without a separate pass over the frame. */
LGI_API lgi_status lgi_unet_conv3x3_first_radiance_fwd(const float* hdr, const void* w, int32_t cin_w,
                                                       const float* bias, int32_t n, int32_t h, int32_t w_,
                                                       int32_t relu, void* out, void* stream);
/* dx = conv3x3^T(dy * (ymask > 0)); channels [0, cin_split) -> dx0, the rest -> dx1.
 * ymask may be null (no ReLU); wt_scratch holds cout*9*cin BF16. */
LGI_API lgi_status lgi_unet_conv3x3_bwd_data(const void* dy, const void* ymask, const void* w, int32_t cout,
                                             int32_t cin, int32_t cin_split, int32_t n, int32_t h, int32_t w_,
                                             void* dx0, void* dx1, void* wt_scratch, void* stream);
/* dw = sum_pixels (dy * (ymask > 0)) (x) patches(cat(src0, src1)) in FP32; db likewise (nullable).
 * scratch: FP32 workspace of scratch_floats (>= cout*9*cin, and >= cout when db is given);
 * lgi_unet_wgrad_scratch_floats gives the size at which the pixel splits fill the GPU, less
 * means fewer splits. Deterministic (fixed-order sums for a given scratch size). */
LGI_API int64_t lgi_unet_wgrad_scratch_floats(int32_t cout, int32_t cin, int32_t n, int32_t h, int32_t w_);
LGI_API lgi_status lgi_unet_conv3x3_bwd_weight(const void* dy, const void* ymask, int32_t cout, const void* src0,
                                               int32_t c0, const void* src1, int32_t c1, int32_t n, int32_t h,
                                               int32_t w_, float* dw, float* db, float* scratch,
                                               int64_t scratch_floats, void* stream);
/* Convolution tile-shape selection (no reference counterpart; the role of cuDNN's benchmark
 * mode under the reference's PyTorch UNet). mode 1 (default): the first call of each layer
 * shape times every tiling variant that fits it on the caller's buffers and stream (it
 * synchronizes that stream once) and later calls reuse the fastest; calls under stream
 * capture and mode 0 use a fixed choice. force_variant(v >= 0) makes every call use variant v
 * (LGI_ERR_UNSUPPORTED where it does not fit), -1 restores selection. tune_table copies up to
 * cap entries: rows[10*i..] = {n, h, w, c0, c1, cout, cout_split, kind, relu, variant} with kind
 * 0 = forward, 1 = ReLU-masked input (data gradient), 2 = forward with a pooled copy, 3 = forward
 * with the fused denoiser tail, each + 8 with FP8 operands,
 * us[LGI_UNET_CONV_VARIANTS*i..] = microseconds per launch of each variant (0 = does not fit);
 * *n = entries. */
#define LGI_UNET_CONV_VARIANTS 14
LGI_API lgi_status lgi_unet_conv_tuning(int32_t mode);
LGI_API lgi_status lgi_unet_conv_force_variant(int32_t variant);
LGI_API lgi_status lgi_unet_conv_tune_table(int32_t* rows, float* us, int32_t cap, int32_t* n);
LGI_API lgi_status lgi_unet_maxpool_fwd(const void* in, int32_t n, int32_t h, int32_t w_, int32_t c, void* out,
                                        void* stream);
LGI_API lgi_status lgi_unet_maxpool_bwd(const void* in, const void* dout, int32_t n, int32_t h, int32_t w_,
                                        int32_t c, void* din, void* stream);
/* bilinear 2x, align_corners = False: (n, h, w, c) -> (n, 2h, 2w, c) */
LGI_API lgi_status lgi_unet_upsample_fwd(const void* in, int32_t n, int32_t h, int32_t w_, int32_t c, void* out,
                                         void* stream);
LGI_API lgi_status lgi_unet_upsample_bwd(const void* dout, int32_t n, int32_t h, int32_t w_, int32_t c, void* din,
                                         void* stream);
/* FP32 3-channel NHWC image -> BF16 NHWC zero-padded to 32 channels */
LGI_API lgi_status lgi_unet_pack_input(const float* x, int64_t npx, void* out, void* stream);
/* out = x + W_t feat + b_t (1x1, 32 -> 3, FP32 out); backward gives dfeat, dW_t, db_t */
LGI_API lgi_status lgi_unet_tail_fwd(const void* feat, const float* x, const float* w, const float* b, int64_t npx,
                                     float* out, void* stream);
/* Inference tail in the radiance domain: out = max(expm1(log1p(max(hdr, 0)) + conv1x1(feat)), 0),
 * the denoiser's output conversion of tools/denoise.py (radiance out) inside the tail. */
LGI_API lgi_status lgi_unet_tail_radiance_fwd(const void* feat, const float* hdr, const float* w, const float* b,
                                              int64_t npx, float* out, void* stream);
/* scratch: 99 * min(1024, ceil(npx / 256)) floats at least (per-block partial sums). */
LGI_API lgi_status lgi_unet_tail_bwd(const void* feat, const float* dout, const float* w, int64_t npx, void* dfeat,
                                     float* dw, float* db, float* scratch, int64_t scratch_floats, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* LGI_H */
