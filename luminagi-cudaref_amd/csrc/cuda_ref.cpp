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
#include <filesystem>
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

}  // namespace

int main(int argc, char** argv) {
    const char* jsonPath = nullptr;
    const char* pngPath = "output/reference.png";
    std::string assetRoot, batchFile, hdrPath;
    int spp = 64, bounces = 2;
    bool useSort = false, noRaySort = false, gbuffer = false, haveCamPos = false, haveCamTgt = false;
    uint32_t rngFlag = LGI_RNG_PCG32;
    float camPos[3] = {0, 0, 0}, camTgt[3] = {0, 0, 0};
    auto read_vec3 = [&](int& i, float out[3]) -> bool {
        if (i + 3 >= argc) return false;
        out[0] = (float)std::atof(argv[++i]);
        out[1] = (float)std::atof(argv[++i]);
        out[2] = (float)std::atof(argv[++i]);
        return true;
    };
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        if (a == "--spp" && i + 1 < argc) spp = std::atoi(argv[++i]);
        else if (a == "--bounces" && i + 1 < argc) bounces = std::atoi(argv[++i]);
        else if (a == "--asset-root" && i + 1 < argc) assetRoot = argv[++i];
        else if (a == "-o" && i + 1 < argc) pngPath = argv[++i];
        else if (a == "--sort") useSort = true;
        else if (a == "--no-ray-sort") noRaySort = true;
        else if (a == "--gbuffer") gbuffer = true;
        else if (a == "--cam-pos") haveCamPos = read_vec3(i, camPos);
        else if (a == "--cam-target") haveCamTgt = read_vec3(i, camTgt);
        else if (a == "--batch" && i + 1 < argc) batchFile = argv[++i];
        else if (a == "--hdr" && i + 1 < argc) hdrPath = argv[++i];
        else if (a == "--rng" && i + 1 < argc) rngFlag = std::string(argv[++i]) == "xorshift" ? LGI_RNG_XORSHIFT_REF : LGI_RNG_PCG32;
        else if (a[0] != '-' && !jsonPath) jsonPath = argv[i];
        else if (a[0] != '-') pngPath = argv[i];
    }
    if (!jsonPath) {
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
    const std::string sp(jsonPath);
    const bool bin = sp.size() > 5 && sp.compare(sp.size() - 5, 5, ".lgis") == 0;
    if ((bin ? lgi_scene_load_bin(ctx, jsonPath, 1) : lgi_scene_load_json(ctx, jsonPath, 1)) != LGI_OK) {
        lgi_destroy(ctx);
        return 2;
    }
    if (!assetRoot.empty()) lgi_set_asset_root(ctx, assetRoot.c_str());
    lgi_camera cam;
    lgi_scene_get(ctx, &cam, nullptr, nullptr, nullptr, nullptr);
    if (haveCamPos && haveCamTgt) {
        lgi_camera_look_at(&cam, camPos, camTgt);
        lgi_scene_set_camera(ctx, &cam);
        std::printf("camera override: pos (%.2f,%.2f,%.2f) -> target (%.2f,%.2f,%.2f)\n", camPos[0], camPos[1],
                    camPos[2], camTgt[0], camTgt[1], camTgt[2]);
    }
    const uint32_t flags =
        rngFlag | (useSort ? LGI_RENDER_SORTED : 0u) | (useSort && noRaySort ? LGI_RENDER_NO_RAY_SORT : 0u);
    auto t0 = std::chrono::steady_clock::now();
    if (lgi_bvh_build(ctx) != LGI_OK) return fail_gpu(ctx, "bvh build");
    int64_t nn = 0;
    lgi_bvh_node_count(ctx, &nn);
    const double bvh_ms = ms_since(t0);
    int64_t ntris = 0;
    lgi_scene_counts(ctx, &ntris, nullptr, nullptr, nullptr);

    if (!batchFile.empty()) {
        int32_t n = 0;
        lgi_batch_parse(batchFile.c_str(), nullptr, 0, &n);
        std::vector<lgi_batch_job> jobs(n);
        if (n) lgi_batch_parse(batchFile.c_str(), jobs.data(), n, &n);
        std::printf("batch: %zu jobs  (scene+bvh loaded once in %.1f ms)\n", jobs.size(), bvh_ms);
        for (size_t k = 0; k < jobs.size(); ++k) {
            const lgi_batch_job& j = jobs[k];
            lgi_camera jc = cam;
            lgi_camera_look_at(&jc, j.cam_pos, j.cam_tgt);
            const int W = jc.width, H = jc.height;
            std::vector<uint8_t> rgb((size_t)W * H * 3), alb, nrm, wp;
            if (gbuffer) { alb.resize(rgb.size()); nrm.resize(rgb.size()); wp.resize(rgb.size()); }
            auto tj = std::chrono::steady_clock::now();
            // batch mode renders with the inline path, as the reference (src/main.cpp:170)
            if (lgi_render(ctx, &jc, j.spp, j.bounces, rngFlag, rgb.data(), nullptr, gbuffer ? alb.data() : nullptr,
                           gbuffer ? nrm.data() : nullptr, gbuffer ? wp.data() : nullptr) != LGI_OK)
                return fail_gpu(ctx, "render");
            float kms = 0;
            lgi_last_render_ms(ctx, &kms);
            std::printf("kernel: %.1f ms  (%d x %d, %d spp, %d bounces, %lld tris)\n", kms, W, H, j.spp, j.bounces,
                        (long long)ntris);
            const double rms = ms_since(tj);
            std::filesystem::path op(j.out_png);
            if (op.has_parent_path()) std::filesystem::create_directories(op.parent_path());
            lgi_write_png(j.out_png, rgb.data(), W, H);
            if (gbuffer) write_gbuffer(j.out_png, W, H, alb, nrm, wp);
            std::printf("[%zu/%zu] %s  spp=%d  render=%.1f ms\n", k + 1, jobs.size(), j.out_png, j.spp, rms);
        }
        lgi_destroy(ctx);
        return 0;
    }
    std::printf("bvh: %.1f ms (%lld nodes)\n", bvh_ms, (long long)nn);
    const int W = cam.width, H = cam.height;
    std::vector<uint8_t> rgb((size_t)W * H * 3), alb, nrm, wp;
    std::vector<float> hdr;
    if (!hdrPath.empty()) hdr.resize(rgb.size());
    const bool gb = gbuffer && !useSort;
    if (gb) { alb.resize(rgb.size()); nrm.resize(rgb.size()); wp.resize(rgb.size()); }
    auto t2 = std::chrono::steady_clock::now();
    if (lgi_render(ctx, nullptr, spp, bounces, flags, rgb.data(), hdr.empty() ? nullptr : hdr.data(),
                   gb ? alb.data() : nullptr, gb ? nrm.data() : nullptr, gb ? wp.data() : nullptr) != LGI_OK)
        return fail_gpu(ctx, "render");
    float kms = 0;
    lgi_last_render_ms(ctx, &kms);
    std::printf("kernel%s: %.1f ms  (%d x %d, %d spp, %d bounces, %lld tris)\n", useSort ? " (sorted)" : "", kms, W, H,
                spp, bounces, (long long)ntris);
    std::printf("render: %.1f ms (%d x %d)\n", ms_since(t2), W, H);
    std::filesystem::path pp(pngPath);
    if (pp.has_parent_path()) std::filesystem::create_directories(pp.parent_path());
    if (lgi_write_png(pngPath, rgb.data(), W, H) != LGI_OK) {
        std::fprintf(stderr, "failed to write %s\n", pngPath);
        lgi_destroy(ctx);
        return 3;
    }
    std::printf("wrote %s\n", pngPath);
    if (gb) write_gbuffer(pngPath, W, H, alb, nrm, wp);
    if (!hdr.empty()) {
        FILE* f = std::fopen(hdrPath.c_str(), "wb");
        if (f) {
            std::fwrite(hdr.data(), 4, hdr.size(), f);
            std::fclose(f);
        }
    }
    lgi_destroy(ctx);
    return 0;
}
