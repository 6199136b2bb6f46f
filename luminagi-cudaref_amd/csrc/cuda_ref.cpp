// cuda_ref.cpp — command-line renderer with the reference's argv surface, exit codes and
// log lines (src/main.cpp:93-229), built on the liblgi C ABI.
//
//   cuda_ref <scene.json | scene.lgis> [out.png] [-o out.png] [--spp N] [--bounces N] [--asset-root DIR]
//            [--sort] [--no-ray-sort] [--gbuffer] [--cam-pos x y z --cam-target x y z] [--batch file]
//            [--hdr out.f32] [--rng xorshift]
// --no-ray-sort (with --sort): the wavefront renderer without its key sort between bounces
// (same image bits).
// Exit codes: 1 usage, 2 scene load failure, 3 PNG write failure, 4 GPU failure.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <functional>
#include <string>
#include <vector>

#include "lgi.h"

namespace {

double ms_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

int fail_gpu(lgi_ctx* ctx, const char* what) {
    std::fprintf(stderr, "%s failed: %s\n", what, lgi_last_error(ctx));
    return 4;
}

bool write_gbuffer(const std::string& png, int W, int H, const std::vector<uint8_t>& alb,
                   const std::vector<uint8_t>& nrm, const std::vector<uint8_t>& wp) {
    std::filesystem::path p(png);
    std::string stem = (p.parent_path() / p.stem()).string();
    bool ok = lgi_write_png((stem + ".albedo.png").c_str(), alb.data(), W, H) == LGI_OK;
    ok = lgi_write_png((stem + ".normal.png").c_str(), nrm.data(), W, H) == LGI_OK && ok;
    ok = lgi_write_png((stem + ".worldpos.png").c_str(), wp.data(), W, H) == LGI_OK && ok;
    return ok;
}

// Command line. The flag names, their values, the positional rule (first non-flag word = scene,
// any later one = output PNG) and the defaults are the reference CLI's interface; a flag whose
// values are missing at the end of argv is ignored, as an unknown flag is.
struct Options {
    std::string scene;
    std::string png = "output/reference.png";
    std::string asset_root, batch, hdr;
    int spp = 64, bounces = 2;
    uint32_t rng = LGI_RNG_PCG32;
    bool sort = false, no_ray_sort = false, gbuffer = false;
    bool has_pos = false, has_tgt = false;
    float pos[3] = {0, 0, 0}, tgt[3] = {0, 0, 0};

    uint32_t render_flags() const {
        return rng | (sort ? LGI_RENDER_SORTED : 0u) | (sort && no_ray_sort ? LGI_RENDER_NO_RAY_SORT : 0u);
    }
};

bool parse_options(int argc, char** argv, Options& o) {
    using Values = const char* const*;
    auto vec3 = [](Values v, float out[3]) {
        for (int k = 0; k < 3; ++k) out[k] = (float)std::atof(v[k]);
    };
    struct Flag {
        const char* name;
        int values;
        std::function<void(Values)> apply;
    };
    const Flag flags[] = {
        {"--spp", 1, [&](Values v) { o.spp = std::atoi(v[0]); }},
        {"--bounces", 1, [&](Values v) { o.bounces = std::atoi(v[0]); }},
        {"--asset-root", 1, [&](Values v) { o.asset_root = v[0]; }},
        {"-o", 1, [&](Values v) { o.png = v[0]; }},
        {"--batch", 1, [&](Values v) { o.batch = v[0]; }},
        {"--hdr", 1, [&](Values v) { o.hdr = v[0]; }},
        {"--rng", 1, [&](Values v) { o.rng = std::strcmp(v[0], "xorshift") == 0 ? LGI_RNG_XORSHIFT_REF : LGI_RNG_PCG32; }},
        {"--sort", 0, [&](Values) { o.sort = true; }},
        {"--no-ray-sort", 0, [&](Values) { o.no_ray_sort = true; }},
        {"--gbuffer", 0, [&](Values) { o.gbuffer = true; }},
        {"--cam-pos", 3, [&](Values v) { vec3(v, o.pos); o.has_pos = true; }},
        {"--cam-target", 3, [&](Values v) { vec3(v, o.tgt); o.has_tgt = true; }},
    };
    int positional = 0;
    for (int i = 1; i < argc; ++i) {
        const Flag* f = nullptr;
        for (const Flag& c : flags)
            if (std::strcmp(argv[i], c.name) == 0) f = &c;
        if (f) {
            if (i + f->values < argc) {
                f->apply(argv + i + 1);
                i += f->values;
            }
            continue;
        }
        if (argv[i][0] == '-') continue;
        (positional++ == 0 ? o.scene : o.png) = argv[i];
    }
    return !o.scene.empty();
}

// One render per batch line (read_batch_file format), the scene and BVH loaded once; batch jobs
// use the inline renderer, as the reference's batch loop does (src/main.cpp:170).
int run_batch(lgi_ctx* ctx, const Options& o, const lgi_camera& cam, double bvh_ms, long long ntris) {
    int32_t n = 0;
    lgi_batch_parse(o.batch.c_str(), nullptr, 0, &n);
    std::vector<lgi_batch_job> jobs(n);
    if (n) lgi_batch_parse(o.batch.c_str(), jobs.data(), n, &n);
    std::printf("batch: %zu jobs  (scene+bvh loaded once in %.1f ms)\n", jobs.size(), bvh_ms);
    for (size_t k = 0; k < jobs.size(); ++k) {
        const lgi_batch_job& job = jobs[k];
        lgi_camera view = cam;
        lgi_camera_look_at(&view, job.cam_pos, job.cam_tgt);
        const size_t bytes = (size_t)view.width * view.height * 3;
        std::vector<uint8_t> rgb(bytes), alb(o.gbuffer ? bytes : 0), nrm(alb.size()), wp(alb.size());
        const auto start = std::chrono::steady_clock::now();
        if (lgi_render(ctx, &view, job.spp, job.bounces, o.rng, rgb.data(), nullptr, o.gbuffer ? alb.data() : nullptr,
                       o.gbuffer ? nrm.data() : nullptr, o.gbuffer ? wp.data() : nullptr) != LGI_OK)
            return fail_gpu(ctx, "render");
        float kernel_ms = 0;
        lgi_last_render_ms(ctx, &kernel_ms);
        std::printf("kernel: %.1f ms  (%d x %d, %d spp, %d bounces, %lld tris)\n", kernel_ms, view.width, view.height,
                    job.spp, job.bounces, ntris);
        const double wall_ms = ms_since(start);
        const std::filesystem::path out(job.out_png);
        if (out.has_parent_path()) std::filesystem::create_directories(out.parent_path());
        lgi_write_png(job.out_png, rgb.data(), view.width, view.height);
        if (o.gbuffer) write_gbuffer(job.out_png, view.width, view.height, alb, nrm, wp);
        std::printf("[%zu/%zu] %s  spp=%d  render=%.1f ms\n", k + 1, jobs.size(), job.out_png, job.spp, wall_ms);
    }
    return 0;
}

// The scene camera's image to o.png (+ G-buffer PNGs, + raw FP32 mean radiance with --hdr).
int run_single(lgi_ctx* ctx, const Options& o, const lgi_camera& cam, long long ntris) {
    const int W = cam.width, H = cam.height;
    const size_t bytes = (size_t)W * H * 3;
    const bool gbuffer = o.gbuffer && !o.sort;  // the sorted path has no G-buffer (src/main.cpp:19-27)
    std::vector<uint8_t> rgb(bytes), alb(gbuffer ? bytes : 0), nrm(alb.size()), wp(alb.size());
    std::vector<float> hdr(o.hdr.empty() ? 0 : bytes);
    const auto start = std::chrono::steady_clock::now();
    if (lgi_render(ctx, nullptr, o.spp, o.bounces, o.render_flags(), rgb.data(), hdr.empty() ? nullptr : hdr.data(),
                   gbuffer ? alb.data() : nullptr, gbuffer ? nrm.data() : nullptr, gbuffer ? wp.data() : nullptr) !=
        LGI_OK)
        return fail_gpu(ctx, "render");
    float kernel_ms = 0;
    lgi_last_render_ms(ctx, &kernel_ms);
    std::printf("kernel%s: %.1f ms  (%d x %d, %d spp, %d bounces, %lld tris)\n", o.sort ? " (sorted)" : "", kernel_ms, W,
                H, o.spp, o.bounces, ntris);
    std::printf("render: %.1f ms (%d x %d)\n", ms_since(start), W, H);
    const std::filesystem::path out(o.png);
    if (out.has_parent_path()) std::filesystem::create_directories(out.parent_path());
    if (lgi_write_png(o.png.c_str(), rgb.data(), W, H) != LGI_OK) {
        std::fprintf(stderr, "failed to write %s\n", o.png.c_str());
        return 3;
    }
    std::printf("wrote %s\n", o.png.c_str());
    if (gbuffer) write_gbuffer(o.png, W, H, alb, nrm, wp);
    if (!hdr.empty()) {
        if (FILE* f = std::fopen(o.hdr.c_str(), "wb")) {
            std::fwrite(hdr.data(), 4, hdr.size(), f);
            std::fclose(f);
        }
    }
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    Options o;
    if (!parse_options(argc, argv, o)) {
        std::fprintf(stderr, "usage: %s <scene.json> [output.png] [--spp N] [--bounces N] [--asset-root DIR]\n",
                     argv[0]);
        return 1;
    }
    lgi_ctx* ctx = nullptr;
    if (lgi_create(&ctx, 0, nullptr) != LGI_OK) {
        std::fprintf(stderr, "no HIP device\n");
        return 4;
    }
    // a .lgis path is the binary scene file (lgi_scene_save_bin), anything else scene.json
    const bool binary = std::filesystem::path(o.scene).extension() == ".lgis";
    if ((binary ? lgi_scene_load_bin(ctx, o.scene.c_str(), 1) : lgi_scene_load_json(ctx, o.scene.c_str(), 1)) !=
        LGI_OK) {
        lgi_destroy(ctx);
        return 2;
    }
    if (!o.asset_root.empty()) lgi_set_asset_root(ctx, o.asset_root.c_str());
    lgi_camera cam;
    lgi_scene_get(ctx, &cam, nullptr, nullptr, nullptr, nullptr);
    if (o.has_pos && o.has_tgt) {
        lgi_camera_look_at(&cam, o.pos, o.tgt);
        lgi_scene_set_camera(ctx, &cam);
        std::printf("camera override: pos (%.2f,%.2f,%.2f) -> target (%.2f,%.2f,%.2f)\n", o.pos[0], o.pos[1], o.pos[2],
                    o.tgt[0], o.tgt[1], o.tgt[2]);
    }
    const auto start = std::chrono::steady_clock::now();
    if (lgi_bvh_build(ctx) != LGI_OK) return fail_gpu(ctx, "bvh build");
    const double bvh_ms = ms_since(start);
    int64_t nodes = 0, ntris = 0;
    lgi_bvh_node_count(ctx, &nodes);
    lgi_scene_counts(ctx, &ntris, nullptr, nullptr, nullptr);
    int rc;
    if (!o.batch.empty()) {
        rc = run_batch(ctx, o, cam, bvh_ms, (long long)ntris);
    } else {
        std::printf("bvh: %.1f ms (%lld nodes)\n", bvh_ms, (long long)nodes);
        rc = run_single(ctx, o, cam, (long long)ntris);
    }
    lgi_destroy(ctx);
    return rc;
}
