// ref_driver.cpp — drives the reference BuildBVH (compiled from the reference's own src/bvh.cpp,
// with src/scene_loader.cpp for scene.json input, by oracle/Makefile) and dumps the result for the
// golden-vector tests and the bench's reference CPU baseline. TEST INFRASTRUCTURE ONLY.
//
//   ref_bvh <scene.json | scene.lgis> <out.bin>
//
// A .lgis input is the project's binary scene file (64-byte header, camera, sun, lights,
// materials, then the 100-byte triangles); only its triangles are read, into the reference's
// Triangle array, which has the same layout. stdout: "BuildBVH <ms> ms" (steady clock around
// the BuildBVH call alone, one thread).
// Output (little endian): int64 n_tris, int64 n_nodes, n_nodes x BvhNode (40 B),
// n_tris x int32 original index of the triangle at each sorted position.
// The original index is carried through the reference's in-place permutation in
// Triangle::material, which BuildBVH never reads (src/bvh.cpp:164-211).
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "bvh.h"
#include "scene.h"

static_assert(sizeof(Triangle) == 100, "the binary scene file stores 100-byte triangles");

namespace {

bool read_exact(FILE* f, void* p, size_t n) { return n == 0 || std::fread(p, 1, n, f) == n; }

// Triangles of a binary scene file (layout of the project's scene_save_bin).
bool load_lgis_triangles(const char* path, Scene& scene) {
    FILE* f = std::fopen(path, "rb");
    if (!f) return false;
    char head[64];
    bool ok = read_exact(f, head, 64) && std::memcmp(head, "LGISCN1", 8) == 0;
    uint64_t n_tris = 0;
    uint32_t n_mats = 0, n_points = 0, n_dirs = 0;
    if (ok) {
        std::memcpy(&n_tris, head + 16, 8);
        std::memcpy(&n_mats, head + 24, 4);
        std::memcpy(&n_points, head + 28, 4);
        std::memcpy(&n_dirs, head + 32, 4);
    }
    // camera (100 B), sun (28 B), directional lights (28 B), point lights (32 B)
    const long skip = 100 + 28 + 28L * n_dirs + 32L * n_points;
    ok = ok && std::fseek(f, skip, SEEK_CUR) == 0;
    for (uint32_t m = 0; ok && m < n_mats; ++m) {  // albedo, emissive, texture path
        uint32_t len = 0;
        ok = std::fseek(f, 24, SEEK_CUR) == 0 && read_exact(f, &len, 4) && std::fseek(f, len, SEEK_CUR) == 0;
    }
    if (ok) {
        scene.triangles.resize(n_tris);
        ok = read_exact(f, scene.triangles.data(), n_tris * sizeof(Triangle));
    }
    std::fclose(f);
    return ok;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: ref_bvh <scene.json | scene.lgis> <out.bin>\n");
        return 1;
    }
    Scene scene;
    const std::string in = argv[1];
    const bool bin = in.size() > 5 && in.compare(in.size() - 5, 5, ".lgis") == 0;
    if (bin ? !load_lgis_triangles(argv[1], scene) : !LoadSceneJSON(argv[1], scene)) return 2;
    for (size_t i = 0; i < scene.triangles.size(); ++i) scene.triangles[i].material = (int)i;
    Bvh bvh;
    const auto t0 = std::chrono::steady_clock::now();
    BuildBVH(scene, bvh);
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    std::printf("BuildBVH %.3f ms\n", ms);
    FILE* f = std::fopen(argv[2], "wb");
    if (!f) return 3;
    int64_t nt = (int64_t)scene.triangles.size(), nn = (int64_t)bvh.nodes.size();
    std::fwrite(&nt, 8, 1, f);
    std::fwrite(&nn, 8, 1, f);
    if (nn) std::fwrite(bvh.nodes.data(), sizeof(BvhNode), (size_t)nn, f);
    for (const Triangle& t : scene.triangles) {
        int32_t idx = t.material;
        std::fwrite(&idx, 4, 1, f);
    }
    std::fclose(f);
    return 0;
}
