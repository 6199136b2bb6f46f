// ref_driver.cpp — drives the reference LoadSceneJSON + BuildBVH (compiled from the
// reference's own src/scene_loader.cpp and src/bvh.cpp by oracle/Makefile) and dumps the
// result for the golden-vector tests. TEST INFRASTRUCTURE ONLY.
//
// Output (little endian): int64 n_tris, int64 n_nodes, n_nodes x BvhNode (40 B),
// n_tris x int32 original index of the triangle at each sorted position.
// The original index is carried through the reference's in-place permutation in
// Triangle::material, which BuildBVH never reads (src/bvh.cpp:164-211).
#include <cstdint>
#include <cstdio>
#include <vector>

#include "bvh.h"
#include "scene.h"

int main(int argc, char** argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: ref_bvh <scene.json> <out.bin>\n");
        return 1;
    }
    Scene scene;
    if (!LoadSceneJSON(argv[1], scene)) return 2;
    for (size_t i = 0; i < scene.triangles.size(); ++i) scene.triangles[i].material = (int)i;
    Bvh bvh;
    BuildBVH(scene, bvh);
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
