// scene_host.h — host-side scene model, scene.json I/O, procedural scenes, PNG and
// batch files. Host C++ only (built with g++ -ffp-contract=off).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "lgi.h"

struct HostMaterial {
    float albedo[3] = {0.8f, 0.8f, 0.8f};
    float emissive[3] = {0.f, 0.f, 0.f};
    std::string diffuse_tex;   // empty = untextured
};

// Mirrors the reference Scene (include/scene.h:47-56).
struct HostScene {
    lgi_camera camera{};
    lgi_dir_light sun{};
    std::vector<lgi_dir_light> dirs;
    std::vector<lgi_point_light> points;
    std::vector<HostMaterial> materials;
    std::vector<lgi_triangle> tris;
    HostScene();
};

// LoadSceneJSON (src/scene_loader.cpp:26-138): schema v1. Returns LGI_OK, LGI_ERR_IO,
// LGI_ERR_PARSE (syntax or missing key) or LGI_ERR_VERSION. verbose prints the
// reference's summary lines.
lgi_status scene_load_json(const char* path, HostScene& out, std::string& err, bool verbose);
lgi_status scene_save_json(const char* path, const HostScene& s, std::string& err);
// Binary scene file (the same model in bulk little-endian records): LGI_ERR_IO, LGI_ERR_PARSE
// (bad magic, truncated or inconsistent sizes) or LGI_ERR_VERSION.
lgi_status scene_save_bin(const char* path, const HostScene& s, std::string& err);
lgi_status scene_load_bin(const char* path, HostScene& out, std::string& err, bool verbose);

// Procedural scenes. kind 1: ground quad + 16 meshes (icosphere L4 / 8x8 box / 64x32
// torus), ~65k triangles. kind 2: ground + 62 meshes (geodesic spheres, tori, value-noise
// height fields) with log-uniform 2k..80k triangles each, ~1.05M in total.
lgi_status scene_generate(int kind, uint64_t seed, int width, int height, HostScene& out,
                          std::string& err);

// overwrite_camera (src/main.cpp:29-59).
void camera_look_at(lgi_camera& cam, const float pos[3], const float tgt[3]);

lgi_status write_png_rgb(const char* path, const uint8_t* rgb, int w, int h);
// PNG -> RGBA8 (non-interlaced; depth 8 or 16; grey, grey+alpha, RGB, RGBA, palette).
lgi_status read_png_rgba(const char* path, int& w, int& h, std::vector<uint8_t>& rgba, std::string& err);
lgi_status parse_batch_file(const char* path, std::vector<lgi_batch_job>& jobs, std::string& err);
