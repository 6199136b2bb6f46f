// api.cpp — C ABI of liblgi.so (declared in include/lgi.h): context, scene upload, GPU
// LBVH, renders, G-buffer, tonemap. Device buffers live in the context and are reused
// across calls (the reference re-uploads the scene on every render, src/pathtracer.cu:635-644).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <new>

#include "lgi_internal.h"
#include "trace_kernels.h"

lgi_status lgi_set_err(lgi_ctx* ctx, lgi_status st, const std::string& msg) {
    if (ctx) ctx->err = msg;
    return st;
}

void* lgi_ctx::scratch_get(int slot, size_t bytes) {
    if ((int)scratch.size() <= slot) scratch.resize(slot + 1, {nullptr, 0});
    auto& e = scratch[slot];
    if (e.second >= bytes && e.first) return e.first;
    if (e.first) (void)hipFree(e.first);
    e = {nullptr, 0};
    void* p = nullptr;
    if (hipMalloc(&p, bytes ? bytes : 16) != hipSuccess) return nullptr;
    e = {p, bytes};
    return p;
}

DevCamera lgi_make_dev_camera(const lgi_camera& c) {
    DevCamera d;
    for (int k = 0; k < 3; ++k) {
        d.pos[k] = c.pos[k];
        d.fwd[k] = c.camera_to_world[k];      // row 0
        d.left[k] = c.camera_to_world[4 + k]; // row 1
        d.up[k] = c.camera_to_world[8 + k];   // row 2
    }
    // host tanf, as src/pathtracer.cu:660
    d.tan_half_fov = std::tan(c.fov_y_deg * 0.5f * 3.14159265358979f / 180.0f);
    d.aspect = c.aspect;
    d.width = c.width;
    d.height = c.height;
    return d;
}

namespace {

// Device initialisation is deferred to the first call that needs the GPU, so the host-side
// entry points (scene generation, JSON, batch files, PNG) also work on machines without one.
lgi_status ensure_gpu(lgi_ctx* ctx) {
    if (ctx->gpu_ready) return LGI_OK;
    LGI_HIP(ctx, hipSetDevice(ctx->device));
    LGI_HIP(ctx, hipEventCreate(&ctx->ev0));
    LGI_HIP(ctx, hipEventCreate(&ctx->ev1));
    LGI_HIP(ctx, hipEventCreate(&ctx->ev_tree0));
    LGI_HIP(ctx, hipEventCreate(&ctx->ev_tree1));
    // this device's static scratch of lgi_render_accumulate_noalloc (resolved once per device)
    if (lgi_noalloc_scratch(ctx->device, &ctx->noalloc_scratch) != LGI_OK)
        return lgi_set_err(ctx, LGI_ERR_HIP, "hipGetSymbolAddress of the allocation-free render's scratch failed");
    ctx->gpu_ready = true;
    return LGI_OK;
}

// A new host scene invalidates the device copy; the upload happens on first GPU use.
void scene_changed(lgi_ctx* ctx) {
    ctx->uploaded = false;
    ctx->bvh_ready = false;
}

// Diffuse textures (src/pathtracer.cu:594-612): every material's texture is loaded from
// asset_root + "/" + path, missing or unreadable files are reported and leave the material
// with its constant albedo. Loaded texels go into one device buffer with a descriptor per
// material, next to the 8-bit sRGB -> linear table the sampler uses.
lgi_status upload_textures(lgi_ctx* ctx) {
    DevScene& d = ctx->dev;
    const HostScene& s = ctx->scene;
    d.tex_loaded = d.tex_missing = 0;
    ctx->tex_dirty = false;
    const std::string& root = ctx->asset_root_set ? ctx->asset_root : ctx->scene_root;
    bool any = false;
    for (const HostMaterial& m : s.materials) any |= !m.diffuse_tex.empty();
    if (!any) return LGI_OK;
    std::vector<int4> desc(s.materials.size(), make_int4(0, 0, 0, 0));
    std::vector<uint32_t> texels;
    for (size_t i = 0; i < s.materials.size(); ++i) {
        const std::string& rel = s.materials[i].diffuse_tex;
        if (rel.empty()) continue;
        const std::string full = root + "/" + rel;
        int w = 0, h = 0;
        std::vector<uint8_t> rgba;
        std::string err;
        if (read_png_rgba(full.c_str(), w, h, rgba, err) != LGI_OK ||
            texels.size() + (size_t)w * h > (size_t)INT32_MAX) {
            std::fprintf(stderr, "  texture not found: %s\n", full.c_str());
            ++d.tex_missing;
            continue;
        }
        desc[i] = make_int4((int)texels.size(), w, h, 1);
        const size_t at = texels.size();
        texels.resize(at + (size_t)w * h);
        std::memcpy(&texels[at], rgba.data(), (size_t)w * h * 4);
        ++d.tex_loaded;
    }
    std::vector<float> lut(256);
    for (int i = 0; i < 256; ++i) {
        const double c = i / 255.0;
        lut[i] = (float)(c <= 0.04045 ? c / 12.92 : std::pow((c + 0.055) / 1.055, 2.4));
    }
    LGI_HIP(ctx, hipMalloc((void**)&d.srgb_lut, 256 * 4));
    LGI_HIP(ctx, hipMemcpy(d.srgb_lut, lut.data(), 256 * 4, hipMemcpyHostToDevice));
    LGI_HIP(ctx, hipMalloc((void**)&d.tex_desc, desc.size() * sizeof(int4)));
    LGI_HIP(ctx, hipMemcpy(d.tex_desc, desc.data(), desc.size() * sizeof(int4), hipMemcpyHostToDevice));
    LGI_HIP(ctx, hipMalloc((void**)&d.tex_texels, std::max<size_t>(1, texels.size()) * 4));
    if (!texels.empty())
        LGI_HIP(ctx, hipMemcpy(d.tex_texels, texels.data(), texels.size() * 4, hipMemcpyHostToDevice));
    return LGI_OK;
}

lgi_status upload_scene(lgi_ctx* ctx) {
    lgi_status st = ensure_gpu(ctx);
    if (st != LGI_OK) return st;
    if (ctx->uploaded) return LGI_OK;
    DevScene& d = ctx->dev;
    lgi_free_dev_scene(d);
    ctx->bvh_ready = false;
    const HostScene& s = ctx->scene;
    d.n_tris = (int64_t)s.tris.size();
    d.n_mats = (int32_t)s.materials.size();
    d.n_points = (int32_t)s.points.size();
    for (const lgi_triangle& t : s.tris)
        if (t.material < 0 || t.material >= d.n_mats)
            return lgi_set_err(ctx, LGI_ERR_INVALID_ARG, "triangle material index out of range");
    if (d.n_tris > 0) {
        LGI_HIP(ctx, hipMalloc((void**)&d.tri_in, s.tris.size() * sizeof(lgi_triangle)));
        LGI_HIP(ctx, hipMemcpy(d.tri_in, s.tris.data(), s.tris.size() * sizeof(lgi_triangle), hipMemcpyHostToDevice));
    }
    std::vector<float> alb(3 * (size_t)std::max(1, d.n_mats), 0.f);
    for (int i = 0; i < d.n_mats; ++i)
        for (int k = 0; k < 3; ++k) alb[3 * i + k] = s.materials[i].albedo[k];
    LGI_HIP(ctx, hipMalloc((void**)&d.mat_albedo, alb.size() * 4));
    LGI_HIP(ctx, hipMemcpy(d.mat_albedo, alb.data(), alb.size() * 4, hipMemcpyHostToDevice));
    LGI_HIP(ctx, hipMalloc((void**)&d.points, sizeof(lgi_point_light) * std::max(1, d.n_points)));
    if (d.n_points)
        LGI_HIP(ctx, hipMemcpy(d.points, s.points.data(), sizeof(lgi_point_light) * d.n_points, hipMemcpyHostToDevice));
    if ((st = upload_textures(ctx)) != LGI_OK) return st;
    ctx->uploaded = true;
    return LGI_OK;
}

bool has_textures(const HostScene& s) {
    for (auto& m : s.materials)
        if (!m.diffuse_tex.empty()) return true;
    return false;
}

TraceParams make_params(lgi_ctx* ctx, const lgi_camera& cam, int spp, int bounces, uint32_t flags, int sample0) {
    TraceParams P{};
    const DevScene& d = ctx->dev;
    P.tnodes = d.tnodes;
    P.triv = d.triv;
    P.tris = d.tris;
    P.mat_albedo = d.mat_albedo;
    P.tex_texels = d.tex_texels;
    P.tex_desc = d.tex_desc;
    P.srgb_lut = d.srgb_lut;
    P.points = d.points;
    P.n_points = d.n_points;
    for (int k = 0; k < 3; ++k) {
        P.sun_dir[k] = ctx->scene.sun.dir[k];
        P.sun_color[k] = ctx->scene.sun.color[k];
    }
    P.sun_intensity = ctx->scene.sun.intensity;
    P.cam = lgi_make_dev_camera(cam);
    P.spp = spp;
    P.bounces = bounces;
    P.sample0 = sample0;
    P.flags = flags;
    P.seed = 0;
    return P;
}

// Global slots for traversal-stack entries beyond the LDS part, one column per thread.
lgi_status attach_spill(lgi_ctx* ctx, TraceParams& P, int64_t threads) {
    P.spill = (int*)ctx->scratch_get(44, (size_t)kStackSpill * threads * 4);
    P.spill_stride = (int32_t)threads;
    if (!P.spill) return lgi_set_err(ctx, LGI_ERR_HIP, "stack spill allocation failed");
    return LGI_OK;
}

constexpr int kQueueGrid = 256 * 8;  // one resident round of 256-thread blocks on 256 CUs

// Uploads the work items and cameras and drains them with the persistent queue kernel,
// directly or through a cached hipGraph (captured on first use of identical inputs).
lgi_status run_queue(lgi_ctx* ctx, const std::vector<lgi_work_item>& items, const std::vector<DevCamera>& cams,
                     const lgi_camera& cam0, int bounces, uint32_t flags, float* sum, float* comp, float* sumsq,
                     bool use_graph) {
    lgi_status st;
    TraceParams P = make_params(ctx, cam0, 0, bounces, flags, 0);
    P.sum = sum;
    P.comp = comp;
    P.sumsq = sumsq;
    if ((st = attach_spill(ctx, P, (int64_t)kQueueGrid * 256)) != LGI_OK) return st;
    // The captured graph reads its items, cameras and counter from slots of its own (72-74), so
    // that direct launches (slots 45-47) never overwrite what a later replay reads.
    const int slot0 = use_graph ? 72 : 45;
    auto* d_items =
        (lgi_work_item*)ctx->scratch_get(slot0, std::max<size_t>(1, items.size()) * sizeof(lgi_work_item));
    auto* d_cams = (DevCamera*)ctx->scratch_get(slot0 + 1, std::max<size_t>(1, cams.size()) * sizeof(DevCamera));
    auto* counter = (uint32_t*)ctx->scratch_get(slot0 + 2, 64);
    if (!d_items || !d_cams || !counter) return lgi_set_err(ctx, LGI_ERR_HIP, "queue allocation failed");
    // key of the captured graph: everything the kernels read by value or address
    std::string key((const char*)items.data(), items.size() * sizeof(lgi_work_item));
    key.append((const char*)cams.data(), cams.size() * sizeof(DevCamera));
    key.append((const char*)&P, sizeof(P));
    for (const void* p : {(const void*)d_items, (const void*)d_cams, (const void*)counter})
        key.append((const char*)&p, sizeof(p));
    const int grid = std::max(1, std::min<int>(kQueueGrid, (int)items.size()));
    ctx->kev_used = 0;
    LGI_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    if (use_graph && ctx->graph_exec && ctx->graph_key == key) {
        LGI_HIP(ctx, hipGraphLaunch((hipGraphExec_t)ctx->graph_exec, ctx->stream));
    } else {
        // Ordered on the context's stream after any earlier launch that still reads these slots.
        // The host vectors are pageable, so each copy has been staged when the call returns.
        LGI_HIP(ctx, hipMemcpyAsync(d_items, items.data(), items.size() * sizeof(lgi_work_item),
                                    hipMemcpyHostToDevice, ctx->stream));
        LGI_HIP(ctx, hipMemcpyAsync(d_cams, cams.data(), cams.size() * sizeof(DevCamera), hipMemcpyHostToDevice,
                                    ctx->stream));
        if (!use_graph) {
            lgi_launch_render_queue(P, d_cams, d_items, (int)items.size(), counter, grid, ctx->stream);
        } else {
            if (!ctx->cap_stream) LGI_HIP(ctx, hipStreamCreateWithFlags((hipStream_t*)&ctx->cap_stream,
                                                                        hipStreamNonBlocking));
            hipStream_t cs = (hipStream_t)ctx->cap_stream;
            if (ctx->graph_exec) {
                (void)hipGraphExecDestroy((hipGraphExec_t)ctx->graph_exec);
                ctx->graph_exec = nullptr;
            }
            hipGraph_t g = nullptr;
            LGI_HIP(ctx, hipStreamBeginCapture(cs, hipStreamCaptureModeRelaxed));
            lgi_launch_render_queue(P, d_cams, d_items, (int)items.size(), counter, grid, cs);
            LGI_HIP(ctx, hipStreamEndCapture(cs, &g));
            hipGraphExec_t exec = nullptr;
            hipError_t e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess) return lgi_set_err(ctx, LGI_ERR_HIP, "hipGraphInstantiate failed");
            ctx->graph_exec = exec;
            ctx->graph_key = key;
            LGI_HIP(ctx, hipGraphLaunch(exec, ctx->stream));
        }
    }
    LGI_HIP(ctx, hipGetLastError());
    LGI_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    return LGI_OK;
}

lgi_status check_render_ready(lgi_ctx* ctx, const lgi_camera& cam) {
    if (!ctx) return LGI_ERR_INVALID_ARG;
    if (ctx->scene.tris.empty()) return lgi_set_err(ctx, LGI_ERR_NO_SCENE, "no scene");
    if (!ctx->bvh_ready) return lgi_set_err(ctx, LGI_ERR_NO_BVH, "lgi_bvh_build has not run");
    if (ctx->tex_dirty) {  // the asset root changed since the upload: reload the textures only
        DevScene& d = ctx->dev;
        for (void* p : {(void*)d.tex_texels, (void*)d.tex_desc, (void*)d.srgb_lut})
            if (p) LGI_HIP(ctx, hipFree(p));
        d.tex_texels = nullptr;
        d.tex_desc = nullptr;
        d.srgb_lut = nullptr;
        lgi_status st = upload_textures(ctx);
        if (st != LGI_OK) return st;
        ctx->tex_dirty = false;
    }
    if (cam.width <= 0 || cam.height <= 0 || (int64_t)cam.width * cam.height > (int64_t(1) << 30))
        return lgi_set_err(ctx, LGI_ERR_INVALID_ARG, "bad image size");
    return LGI_OK;
}

}  // namespace

lgi_status lgi_mark_trace(lgi_ctx* ctx, bool begin) {
    constexpr int kPairs = 64;
    if (ctx->kev.empty()) {
        ctx->kev.resize(2 * kPairs, nullptr);
        for (auto& e : ctx->kev) LGI_HIP(ctx, hipEventCreate(&e));
    }
    if (ctx->kev_used >= kPairs) {
        ctx->kev_overflow = true;
        return LGI_OK;
    }
    LGI_HIP(ctx, hipEventRecord(ctx->kev[2 * ctx->kev_used + (begin ? 0 : 1)], ctx->stream));
    if (!begin) ++ctx->kev_used;
    return LGI_OK;
}

extern "C" {

const char* lgi_version(void) { return "lgi 0.1 (gfx950)"; }

lgi_status lgi_create(lgi_ctx** out, int device, void* stream) {
    if (!out) return LGI_ERR_INVALID_ARG;
    *out = nullptr;
    if (device < 0) return LGI_ERR_INVALID_ARG;
    lgi_ctx* c = new (std::nothrow) lgi_ctx();
    if (!c) return LGI_ERR_INVALID_ARG;
    c->device = device;
    c->stream = (hipStream_t)stream;
    *out = c;
    return LGI_OK;
}

void lgi_destroy(lgi_ctx* ctx) {
    if (!ctx) return;
    if (ctx->gpu_ready) {
        (void)hipSetDevice(ctx->device);
        (void)hipStreamSynchronize(ctx->stream);
        lgi_free_dev_scene(ctx->dev);
        for (auto& e : ctx->scratch)
            if (e.first) (void)hipFree(e.first);
        if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
        if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
        if (ctx->ev_tree0) (void)hipEventDestroy(ctx->ev_tree0);
        if (ctx->ev_tree1) (void)hipEventDestroy(ctx->ev_tree1);
        for (auto e : ctx->kev)
            if (e) (void)hipEventDestroy(e);
        if (ctx->graph_exec) (void)hipGraphExecDestroy((hipGraphExec_t)ctx->graph_exec);
        if (ctx->cap_stream) (void)hipStreamDestroy((hipStream_t)ctx->cap_stream);
    }
    delete ctx;
}

const char* lgi_last_error(const lgi_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

lgi_status lgi_set_stream(lgi_ctx* ctx, void* stream) {
    if (!ctx) return LGI_ERR_INVALID_ARG;
    ctx->stream = (hipStream_t)stream;
    return LGI_OK;
}

lgi_status lgi_scene_load_json(lgi_ctx* ctx, const char* path, int verbose) {
    if (!ctx || !path) return LGI_ERR_INVALID_ARG;
    HostScene s;
    lgi_status st = scene_load_json(path, s, ctx->err, verbose != 0);
    if (st != LGI_OK) return st;
    ctx->scene = std::move(s);
    // texture paths resolve against the scene's grandparent directory unless an asset root was
    // set (the reference CLI's default, src/main.cpp:140-141)
    ctx->scene_root = std::filesystem::path(path).parent_path().parent_path().string();
    scene_changed(ctx);
    return LGI_OK;
}

lgi_status lgi_set_asset_root(lgi_ctx* ctx, const char* dir) {
    if (!ctx) return LGI_ERR_INVALID_ARG;
    ctx->asset_root = dir ? dir : "";
    ctx->asset_root_set = dir != nullptr;
    ctx->tex_dirty = ctx->uploaded;  // textures reload before the next render
    return LGI_OK;
}

lgi_status lgi_scene_save_json(lgi_ctx* ctx, const char* path) {
    if (!ctx || !path) return LGI_ERR_INVALID_ARG;
    return scene_save_json(path, ctx->scene, ctx->err);
}

lgi_status lgi_scene_load_bin(lgi_ctx* ctx, const char* path, int verbose) {
    if (!ctx || !path) return LGI_ERR_INVALID_ARG;
    HostScene s;
    lgi_status st = scene_load_bin(path, s, ctx->err, verbose != 0);
    if (st != LGI_OK) return st;
    ctx->scene = std::move(s);
    ctx->scene_root = std::filesystem::path(path).parent_path().parent_path().string();
    scene_changed(ctx);
    return LGI_OK;
}

lgi_status lgi_scene_save_bin(lgi_ctx* ctx, const char* path) {
    if (!ctx || !path) return LGI_ERR_INVALID_ARG;
    return scene_save_bin(path, ctx->scene, ctx->err);
}

lgi_status lgi_scene_set(lgi_ctx* ctx, const lgi_camera* cam, const lgi_dir_light* sun, const lgi_dir_light* dirs,
                         int32_t n_dirs, const lgi_point_light* points, int32_t n_points, const lgi_material* mats,
                         int32_t n_mats, const lgi_triangle* tris, int64_t n_tris) {
    if (!ctx || !cam || !sun || n_dirs < 0 || n_points < 0 || n_mats < 0 || n_tris < 0) return LGI_ERR_INVALID_ARG;
    if ((n_dirs && !dirs) || (n_points && !points) || (n_mats && !mats) || (n_tris && !tris))
        return LGI_ERR_INVALID_ARG;
    HostScene s;
    s.camera = *cam;
    s.sun = *sun;
    s.dirs.assign(dirs, dirs + n_dirs);
    s.points.assign(points, points + n_points);
    s.materials.resize(n_mats);
    for (int i = 0; i < n_mats; ++i) {
        std::memcpy(s.materials[i].albedo, mats[i].albedo, 12);
        std::memcpy(s.materials[i].emissive, mats[i].emissive, 12);
    }
    s.tris.assign(tris, tris + n_tris);
    ctx->scene = std::move(s);
    scene_changed(ctx);
    return LGI_OK;
}

lgi_status lgi_scene_generate(lgi_ctx* ctx, int32_t kind, uint64_t seed, int32_t width, int32_t height) {
    if (!ctx) return LGI_ERR_INVALID_ARG;
    HostScene s;
    lgi_status st = scene_generate(kind, seed, width, height, s, ctx->err);
    if (st != LGI_OK) return st;
    ctx->scene = std::move(s);
    scene_changed(ctx);
    return LGI_OK;
}

lgi_status lgi_scene_counts(const lgi_ctx* ctx, int64_t* n_tris, int32_t* n_mats, int32_t* n_points, int32_t* n_dirs) {
    if (!ctx) return LGI_ERR_INVALID_ARG;
    if (n_tris) *n_tris = (int64_t)ctx->scene.tris.size();
    if (n_mats) *n_mats = (int32_t)ctx->scene.materials.size();
    if (n_points) *n_points = (int32_t)ctx->scene.points.size();
    if (n_dirs) *n_dirs = (int32_t)ctx->scene.dirs.size();
    return LGI_OK;
}

lgi_status lgi_scene_get(const lgi_ctx* ctx_c, lgi_camera* cam, lgi_dir_light* sun, lgi_point_light* points,
                         lgi_material* mats, lgi_triangle* tris) {
    lgi_ctx* ctx = const_cast<lgi_ctx*>(ctx_c);
    if (!ctx) return LGI_ERR_INVALID_ARG;
    const HostScene& s = ctx->scene;
    if (cam) *cam = s.camera;
    if (sun) *sun = s.sun;
    if (points && !s.points.empty()) std::memcpy(points, s.points.data(), s.points.size() * sizeof(lgi_point_light));
    if (mats)
        for (size_t i = 0; i < s.materials.size(); ++i) {
            std::memcpy(mats[i].albedo, s.materials[i].albedo, 12);
            std::memcpy(mats[i].emissive, s.materials[i].emissive, 12);
        }
    if (tris && !s.tris.empty()) {
        if (ctx->bvh_ready) {  // BuildBVH leaves scene.triangles in Morton order
            std::vector<int32_t> perm(s.tris.size());
            LGI_HIP(ctx, hipMemcpy(perm.data(), ctx->dev.perm, perm.size() * 4, hipMemcpyDeviceToHost));
            for (size_t i = 0; i < perm.size(); ++i) tris[i] = s.tris[perm[i]];
        } else {
            std::memcpy(tris, s.tris.data(), s.tris.size() * sizeof(lgi_triangle));
        }
    }
    return LGI_OK;
}

lgi_status lgi_scene_set_camera(lgi_ctx* ctx, const lgi_camera* cam) {
    if (!ctx || !cam) return LGI_ERR_INVALID_ARG;
    ctx->scene.camera = *cam;
    return LGI_OK;
}

lgi_status lgi_camera_look_at(lgi_camera* cam, const float pos[3], const float target[3]) {
    if (!cam || !pos || !target) return LGI_ERR_INVALID_ARG;
    camera_look_at(*cam, pos, target);
    return LGI_OK;
}

lgi_status lgi_bvh_build(lgi_ctx* ctx) {
    if (!ctx) return LGI_ERR_INVALID_ARG;
    if (ctx->scene.tris.empty()) {
        ctx->bvh_ready = false;
        return lgi_set_err(ctx, LGI_ERR_NO_SCENE, "scene has no triangles");
    }
    lgi_status st = upload_scene(ctx);
    if (st != LGI_OK) return st;
    LGI_HIP(ctx, hipSetDevice(ctx->device));
    st = lgi_gpu_build_bvh(ctx);
    if (st != LGI_OK) return st;
    LGI_HIP(ctx, hipMemcpy(&ctx->root, ctx->dev.nodes, sizeof(lgi_bvh_node), hipMemcpyDeviceToHost));
    return LGI_OK;
}

lgi_status lgi_bvh_node_count(const lgi_ctx* ctx, int64_t* n_nodes) {
    if (!ctx || !n_nodes) return LGI_ERR_INVALID_ARG;
    *n_nodes = ctx->bvh_ready ? ctx->dev.n_nodes : 0;
    return LGI_OK;
}

lgi_status lgi_bvh_get(lgi_ctx* ctx, lgi_bvh_node* nodes, int32_t* tri_perm, uint32_t* keys) {
    if (!ctx) return LGI_ERR_INVALID_ARG;
    if (!ctx->bvh_ready) return lgi_set_err(ctx, LGI_ERR_NO_BVH, "no BVH");
    const DevScene& d = ctx->dev;
    LGI_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (nodes) LGI_HIP(ctx, hipMemcpy(nodes, d.nodes, d.n_nodes * sizeof(lgi_bvh_node), hipMemcpyDeviceToHost));
    if (tri_perm) LGI_HIP(ctx, hipMemcpy(tri_perm, d.perm, d.n_tris * 4, hipMemcpyDeviceToHost));
    if (keys) LGI_HIP(ctx, hipMemcpy(keys, d.keys, d.n_tris * 4, hipMemcpyDeviceToHost));
    return LGI_OK;
}

lgi_status lgi_bvh_set_treelet_passes(lgi_ctx* ctx, int32_t passes) {
    if (!ctx) return LGI_ERR_INVALID_ARG;
    if (passes < 0 || passes > 8) return lgi_set_err(ctx, LGI_ERR_INVALID_ARG, "treelet passes must be 0..8");
    ctx->bvh_treelet_passes = passes;
    return LGI_OK;
}

lgi_status lgi_bvh_get_traversal(lgi_ctx* ctx, lgi_bvh_node* nodes) {
    if (!ctx || !nodes) return LGI_ERR_INVALID_ARG;
    if (!ctx->bvh_ready) return lgi_set_err(ctx, LGI_ERR_NO_BVH, "no BVH");
    const DevScene& d = ctx->dev;
    LGI_HIP(ctx, hipStreamSynchronize(ctx->stream));
    LGI_HIP(ctx, hipMemcpy(nodes, d.trav_nodes, d.n_nodes * sizeof(lgi_bvh_node), hipMemcpyDeviceToHost));
    return LGI_OK;
}

lgi_status lgi_bvh_last_treelet_ms(const lgi_ctx* ctx, float* ms) {
    if (!ctx || !ms) return LGI_ERR_INVALID_ARG;
    *ms = ctx->last_treelet_ms;
    return LGI_OK;
}

lgi_status lgi_bvh_last_build_ms(const lgi_ctx* ctx, float* ms) {
    if (!ctx || !ms) return LGI_ERR_INVALID_ARG;
    *ms = ctx->last_build_ms;
    return LGI_OK;
}

lgi_status lgi_render_accumulate_noalloc(lgi_ctx* ctx, const lgi_camera* cam_in, int32_t spp, int32_t bounces,
                                         uint32_t flags, int32_t sample0, float* d_sum, float* d_comp,
                                         float* d_sumsq) {
    if (!ctx) return LGI_ERR_INVALID_ARG;
    const lgi_camera& cam = cam_in ? *cam_in : ctx->scene.camera;
    if (flags & LGI_RENDER_SORTED)
        return lgi_set_err(ctx, LGI_ERR_UNSUPPORTED,
                           "allocation-free render: the sorted renderer's queues are allocated per pass");
    if (ctx->tex_dirty)  // check_render_ready would reload (and allocate) the textures
        return lgi_set_err(ctx, LGI_ERR_UNSUPPORTED,
                           "allocation-free render: textures changed since the last render; render once first");
    lgi_status st = check_render_ready(ctx, cam);
    if (st != LGI_OK) return st;
    if (spp < 0 || bounces < 0 || bounces > 64 || sample0 < 0 || !d_sum || !d_comp)
        return lgi_set_err(ctx, LGI_ERR_INVALID_ARG, "bad render arguments");
    if (spp == 0) return LGI_OK;
    int* scratch_dev = ctx->noalloc_scratch;  // resolved when the context set up its device
    if (!scratch_dev) return lgi_set_err(ctx, LGI_ERR_HIP, "allocation-free render: no scratch address");
    TraceParams P = make_params(ctx, cam, spp, bounces, flags, sample0);
    P.sum = d_sum;
    P.comp = d_comp;
    P.sumsq = d_sumsq;
    ctx->kev_used = 0;
    lgi_launch_render_all_tiles_noalloc(P, scratch_dev, ctx->stream);
    LGI_HIP(ctx, hipGetLastError());
    return LGI_OK;
}

lgi_status lgi_render_accumulate(lgi_ctx* ctx, const lgi_camera* cam_in, int32_t spp, int32_t bounces, uint32_t flags,
                                 int32_t sample0, float* d_sum, float* d_comp, float* d_sumsq) {
    if (!ctx) return LGI_ERR_INVALID_ARG;
    const lgi_camera& cam = cam_in ? *cam_in : ctx->scene.camera;
    lgi_status st = check_render_ready(ctx, cam);
    if (st != LGI_OK) return st;
    if (spp < 0 || bounces < 0 || bounces > 64 || sample0 < 0 || !d_sum || !d_comp)
        return lgi_set_err(ctx, LGI_ERR_INVALID_ARG, "bad render arguments");
    if ((flags & LGI_RENDER_SORTED) && (flags & LGI_RNG_XORSHIFT_REF))
        return lgi_set_err(ctx, LGI_ERR_UNSUPPORTED,
                           "the sorted renderer needs per-sample RNG streams (PCG32); xorshift-ref chains "
                           "samples within an 8-sample chunk");
    if (spp == 0) return LGI_OK;
    TraceParams P = make_params(ctx, cam, spp, bounces, flags, sample0);
    P.sum = d_sum;
    P.comp = d_comp;
    P.sumsq = d_sumsq;
    if (flags & LGI_RENDER_SORTED) {
        const lgi_bvh_node& r = ctx->root;
        return lgi_render_wavefront(ctx, P, make_float3(r.bb_min[0], r.bb_min[1], r.bb_min[2]),
                                    make_float3(r.bb_max[0], r.bb_max[1], r.bb_max[2]));
    }
    if ((st = attach_spill(ctx, P, lgi_grid_threads(cam.width, cam.height))) != LGI_OK) return st;
    ctx->kev_used = 0;
    // timing events only outside stream capture, so that the call can be captured in a graph
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    LGI_HIP(ctx, hipStreamIsCapturing(ctx->stream, &cap));
    const bool timed = cap == hipStreamCaptureStatusNone;
    if (timed) LGI_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    lgi_launch_render(P, ctx->stream);
    LGI_HIP(ctx, hipGetLastError());
    if (timed) LGI_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    return LGI_OK;
}

lgi_status lgi_last_trace_ms(const lgi_ctx* ctx_c, float* ms, int32_t* launches) {
    lgi_ctx* ctx = const_cast<lgi_ctx*>(ctx_c);
    if (!ctx || !ms || !launches) return LGI_ERR_INVALID_ARG;
    if (!ctx->gpu_ready) return lgi_set_err(ctx, LGI_ERR_NO_BVH, "nothing rendered yet");
    LGI_HIP(ctx, hipEventSynchronize(ctx->ev1));
    if (ctx->kev_used == 0) {  // inline renderer: the megakernel is the traversal kernel
        LGI_HIP(ctx, hipEventElapsedTime(ms, ctx->ev0, ctx->ev1));
        *launches = 1;
        return LGI_OK;
    }
    if (ctx->kev_overflow) return lgi_set_err(ctx, LGI_ERR_UNSUPPORTED, "more traversal launches than event pairs");
    float total = 0.f;
    for (int i = 0; i < ctx->kev_used; ++i) {
        float t = 0.f;
        LGI_HIP(ctx, hipEventElapsedTime(&t, ctx->kev[2 * i], ctx->kev[2 * i + 1]));
        total += t;
    }
    *ms = total;
    *launches = ctx->kev_used;
    return LGI_OK;
}

lgi_status lgi_last_render_ms(const lgi_ctx* ctx_c, float* ms) {
    lgi_ctx* ctx = const_cast<lgi_ctx*>(ctx_c);
    if (!ctx || !ms) return LGI_ERR_INVALID_ARG;
    if (!ctx->gpu_ready) return lgi_set_err(ctx, LGI_ERR_NO_BVH, "nothing rendered yet");
    LGI_HIP(ctx, hipEventSynchronize(ctx->ev1));
    LGI_HIP(ctx, hipEventElapsedTime(ms, ctx->ev0, ctx->ev1));
    ctx->last_render_ms = *ms;
    return LGI_OK;
}

lgi_status lgi_tonemap(lgi_ctx* ctx, const float* d_sum, int32_t n_px, float inv_spp, uint8_t* d_rgb) {
    if (!ctx || !d_sum || !d_rgb || n_px < 0) return LGI_ERR_INVALID_ARG;
    if (n_px == 0) return LGI_OK;
    lgi_status st = ensure_gpu(ctx);
    if (st != LGI_OK) return st;
    lgi_launch_tonemap(d_sum, (int64_t)n_px * 3, inv_spp, d_rgb, ctx->stream);
    LGI_HIP(ctx, hipGetLastError());
    return LGI_OK;
}

lgi_status lgi_primary_hits(lgi_ctx* ctx, const lgi_camera* cam_in, int32_t* d_tri, float* d_t) {
    if (!ctx || !d_tri || !d_t) return LGI_ERR_INVALID_ARG;
    const lgi_camera& cam = cam_in ? *cam_in : ctx->scene.camera;
    lgi_status st = check_render_ready(ctx, cam);
    if (st != LGI_OK) return st;
    TraceParams P = make_params(ctx, cam, 1, 0, 0, 0);
    if ((st = attach_spill(ctx, P, lgi_grid_threads(cam.width, cam.height))) != LGI_OK) return st;
    lgi_launch_primary(P, d_tri, d_t, ctx->stream);
    LGI_HIP(ctx, hipGetLastError());
    return LGI_OK;
}

lgi_status lgi_gbuffer(lgi_ctx* ctx, const lgi_camera* cam_in, uint8_t* d_alb, uint8_t* d_nrm, uint8_t* d_wp) {
    if (!ctx || !d_alb || !d_nrm || !d_wp) return LGI_ERR_INVALID_ARG;
    const lgi_camera& cam = cam_in ? *cam_in : ctx->scene.camera;
    lgi_status st = check_render_ready(ctx, cam);
    if (st != LGI_OK) return st;
    // root extents (src/pathtracer.cu:728-734)
    lgi_bvh_node root;
    LGI_HIP(ctx, hipMemcpyAsync(&root, ctx->dev.nodes, sizeof(root), hipMemcpyDeviceToHost, ctx->stream));
    LGI_HIP(ctx, hipStreamSynchronize(ctx->stream));
    float3 mn = make_float3(root.bb_min[0], root.bb_min[1], root.bb_min[2]);
    float ex = root.bb_max[0] - root.bb_min[0], ey = root.bb_max[1] - root.bb_min[1],
          ez = root.bb_max[2] - root.bb_min[2];
    float3 inv = make_float3(1.0f / std::fmax(ex, 1e-6f), 1.0f / std::fmax(ey, 1e-6f), 1.0f / std::fmax(ez, 1e-6f));
    TraceParams P = make_params(ctx, cam, 1, 0, 0, 0);
    if ((st = attach_spill(ctx, P, lgi_grid_threads(cam.width, cam.height))) != LGI_OK) return st;
    lgi_launch_gbuffer(P, mn, inv, d_alb, d_nrm, d_wp, ctx->stream);
    LGI_HIP(ctx, hipGetLastError());
    return LGI_OK;
}

lgi_status lgi_render(lgi_ctx* ctx, const lgi_camera* cam_in, int32_t spp, int32_t bounces, uint32_t flags,
                      uint8_t* rgb, float* hdr, uint8_t* alb, uint8_t* nrm, uint8_t* wp) {
    if (!ctx || !rgb) return LGI_ERR_INVALID_ARG;
    const lgi_camera cam = cam_in ? *cam_in : ctx->scene.camera;
    lgi_status st = check_render_ready(ctx, cam);
    if (st != LGI_OK) return st;
    if (spp <= 0) return lgi_set_err(ctx, LGI_ERR_INVALID_ARG, "spp must be positive");
    std::printf("textures: %d loaded, %d missing (of %zu materials)\n", ctx->dev.tex_loaded, ctx->dev.tex_missing,
                ctx->scene.materials.size());
    const size_t npx = (size_t)cam.width * cam.height;
    float* sum = (float*)ctx->scratch_get(40, npx * 12);
    float* comp = (float*)ctx->scratch_get(41, npx * 12);
    uint8_t* img = (uint8_t*)ctx->scratch_get(42, npx * 3 * 4);
    if (!sum || !comp || !img) return lgi_set_err(ctx, LGI_ERR_HIP, "render buffer allocation failed");
    LGI_HIP(ctx, hipMemsetAsync(sum, 0, npx * 12, ctx->stream));
    LGI_HIP(ctx, hipMemsetAsync(comp, 0, npx * 12, ctx->stream));
    st = lgi_render_accumulate(ctx, &cam, spp, bounces, flags, 0, sum, comp, nullptr);
    if (st != LGI_OK) return st;
    lgi_launch_tonemap(sum, (int64_t)npx * 3, 1.0f / (float)spp, img, ctx->stream);
    LGI_HIP(ctx, hipGetLastError());
    LGI_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    float ms = 0;
    LGI_HIP(ctx, hipEventSynchronize(ctx->ev1));
    LGI_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    ctx->last_render_ms = ms;
    LGI_HIP(ctx, hipMemcpy(rgb, img, npx * 3, hipMemcpyDeviceToHost));
    if (hdr) {
        std::vector<float> h(npx * 3);
        LGI_HIP(ctx, hipMemcpy(h.data(), sum, npx * 12, hipMemcpyDeviceToHost));
        const float inv = 1.0f / (float)spp;
        for (size_t i = 0; i < npx * 3; ++i) hdr[i] = h[i] * inv;
    }
    if (alb || nrm || wp) {
        if (!(alb && nrm && wp)) return LGI_OK;  // reference: all three or none (:726)
        uint8_t* g = (uint8_t*)ctx->scratch_get(43, npx * 9);
        if (!g) return lgi_set_err(ctx, LGI_ERR_HIP, "gbuffer allocation failed");
        st = lgi_gbuffer(ctx, &cam, g, g + npx * 3, g + npx * 6);
        if (st != LGI_OK) return st;
        LGI_HIP(ctx, hipStreamSynchronize(ctx->stream));
        LGI_HIP(ctx, hipMemcpy(alb, g, npx * 3, hipMemcpyDeviceToHost));
        LGI_HIP(ctx, hipMemcpy(nrm, g + npx * 3, npx * 3, hipMemcpyDeviceToHost));
        LGI_HIP(ctx, hipMemcpy(wp, g + npx * 6, npx * 3, hipMemcpyDeviceToHost));
    }
    return LGI_OK;
}

lgi_status lgi_render_batch(lgi_ctx* ctx, const lgi_pose_job* jobs, int32_t n, int32_t bounces, uint32_t flags,
                            float* d_sum, float* d_comp, float* d_sumsq, int32_t use_graph) {
    if (!ctx || !jobs || n <= 0 || !d_sum || !d_comp || bounces < 0 || bounces > 64) return LGI_ERR_INVALID_ARG;
    const lgi_camera& cam0 = jobs[0].cam;
    lgi_status st = check_render_ready(ctx, cam0);
    if (st != LGI_OK) return st;
    if (flags & (LGI_RENDER_SORTED | LGI_RNG_XORSHIFT_REF))
        return lgi_set_err(ctx, LGI_ERR_UNSUPPORTED, "the batch queue renders the PCG32 inline path");
    if (flags & LGI_RENDER_SET_SUMS)
        return lgi_set_err(ctx, LGI_ERR_UNSUPPORTED, "the batch queue adds its items' samples to the sums");
    const int tiles = ((cam0.width + 7) / 8) * ((cam0.height + 7) / 8);
    std::vector<DevCamera> cams(n);
    std::vector<lgi_work_item> items;
    items.reserve((size_t)tiles * n);
    for (int p = 0; p < n; ++p) {
        if (jobs[p].cam.width != cam0.width || jobs[p].cam.height != cam0.height || jobs[p].spp < 0 ||
            jobs[p].sample0 < 0)
            return lgi_set_err(ctx, LGI_ERR_INVALID_ARG, "batch poses must share the image size");
        cams[p] = lgi_make_dev_camera(jobs[p].cam);
        for (int t = 0; t < tiles && jobs[p].spp > 0; ++t) items.push_back({p, t, jobs[p].spp, jobs[p].sample0});
    }
    // longest items first: the queue then ends on short items (less idle tail)
    std::stable_sort(items.begin(), items.end(),
                     [](const lgi_work_item& a, const lgi_work_item& b) { return a.spp > b.spp; });
    if (items.empty()) return LGI_OK;
    return run_queue(ctx, items, cams, cam0, bounces, flags, d_sum, d_comp, d_sumsq, use_graph != 0);
}

lgi_status lgi_render_tiles(lgi_ctx* ctx, const lgi_camera* cam_in, const int32_t* tiles, int32_t n_tiles, int32_t spp,
                            int32_t sample0, int32_t bounces, uint32_t flags, float* d_sum, float* d_comp,
                            float* d_sumsq) {
    if (!ctx || (n_tiles > 0 && !tiles) || n_tiles < 0 || spp < 0 || sample0 < 0 || !d_sum || !d_comp)
        return LGI_ERR_INVALID_ARG;
    const lgi_camera& cam = cam_in ? *cam_in : ctx->scene.camera;
    lgi_status st = check_render_ready(ctx, cam);
    if (st != LGI_OK) return st;
    if (flags & LGI_RENDER_SET_SUMS)
        return lgi_set_err(ctx, LGI_ERR_UNSUPPORTED, "tile renders add samples to the sums");
    if (flags & LGI_RNG_XORSHIFT_REF)
        return lgi_set_err(ctx, LGI_ERR_UNSUPPORTED, "tile renders use PCG32 per-sample streams");
    const int ntiles = ((cam.width + 7) / 8) * ((cam.height + 7) / 8);
    std::vector<uint8_t> seen((size_t)ntiles, 0);
    std::vector<lgi_work_item> items;
    for (int i = 0; i < n_tiles; ++i) {
        if (tiles[i] < 0 || tiles[i] >= ntiles) return lgi_set_err(ctx, LGI_ERR_INVALID_ARG, "tile index out of range");
        if (seen[tiles[i]]++) return lgi_set_err(ctx, LGI_ERR_INVALID_ARG, "tile listed twice");
        if (spp > 0) items.push_back({0, tiles[i], spp, sample0});
    }
    if (items.empty()) return LGI_OK;
    if (flags & LGI_RENDER_SORTED) {
        // all samples of all listed tiles in flight at once in the wavefront renderer
        int32_t* d_tiles = (int32_t*)ctx->scratch_get(48, (size_t)n_tiles * 4);
        if (!d_tiles) return lgi_set_err(ctx, LGI_ERR_HIP, "tile list allocation failed");
        LGI_HIP(ctx, hipMemcpyAsync(d_tiles, tiles, (size_t)n_tiles * 4, hipMemcpyHostToDevice, ctx->stream));
        TraceParams P = make_params(ctx, cam, spp, bounces, flags, sample0);
        P.sum = d_sum;
        P.comp = d_comp;
        P.sumsq = d_sumsq;
        const lgi_bvh_node& r = ctx->root;
        return lgi_render_wavefront(ctx, P, make_float3(r.bb_min[0], r.bb_min[1], r.bb_min[2]),
                                    make_float3(r.bb_max[0], r.bb_max[1], r.bb_max[2]), d_tiles, n_tiles);
    }
    std::vector<DevCamera> cams{lgi_make_dev_camera(cam)};
    return run_queue(ctx, items, cams, cam, bounces, flags, d_sum, d_comp, d_sumsq, false);
}

lgi_status lgi_batch_parse(const char* path, lgi_batch_job* jobs, int32_t cap, int32_t* n_jobs) {
    if (!path || !n_jobs || cap < 0 || (cap > 0 && !jobs)) return LGI_ERR_INVALID_ARG;
    std::vector<lgi_batch_job> v;
    std::string err;
    lgi_status st = parse_batch_file(path, v, err);
    if (st != LGI_OK) return st;
    *n_jobs = (int32_t)v.size();
    for (int32_t i = 0; i < cap && i < (int32_t)v.size(); ++i) jobs[i] = v[i];
    return LGI_OK;
}

lgi_status lgi_write_png(const char* path, const uint8_t* rgb, int32_t w, int32_t h) {
    return write_png_rgb(path, rgb, w, h);
}

lgi_status lgi_read_png(const char* path, int32_t* w, int32_t* h, uint8_t* rgba) {
    if (!path || !w || !h) return LGI_ERR_INVALID_ARG;
    int iw = 0, ih = 0;
    std::vector<uint8_t> px;
    std::string err;
    lgi_status st = read_png_rgba(path, iw, ih, px, err);
    if (st != LGI_OK) return st;
    *w = iw;
    *h = ih;
    if (rgba) std::memcpy(rgba, px.data(), px.size());
    return LGI_OK;
}

}  // extern "C"
