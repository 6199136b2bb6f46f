// host_harness.cpp — drives the host-side parsers of liblgi (scene_host.cpp: scene.json, binary
// scene files, PNG, batch files) and the CPU oracle (oracle.cpp) under AddressSanitizer and
// UndefinedBehaviorSanitizer. Built and run by tests/test_sanitizers.py; test infrastructure.
//
//   host_harness <work dir> <scene.json> [more PNG files ...]
//
// Every well-formed input must load and round-trip; every truncated or byte-flipped variant
// must come back with a status (any status) and no sanitizer report. Exits 0 and prints
// "harness ok" when all of that holds.
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iterator>
#include <string>
#include <vector>

#include "scene_host.h"
#include "oracle.cpp"  // the oracle's C ABI and its structs, in this translation unit

namespace {

int failures = 0;
#define CHECK(cond, ...)                                                              \
    do {                                                                              \
        if (!(cond)) {                                                                \
            std::fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__);                 \
            std::fprintf(stderr, __VA_ARGS__);                                        \
            std::fprintf(stderr, "\n");                                               \
            ++failures;                                                               \
        }                                                                             \
    } while (0)

std::vector<uint8_t> slurp(const std::string& path) {
    std::ifstream in(path, std::ios::binary);
    return std::vector<uint8_t>((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
}

void spit(const std::string& path, const uint8_t* p, size_t n) {
    std::ofstream out(path, std::ios::binary | std::ios::trunc);
    out.write((const char*)p, (std::streamsize)n);
}

struct Mix {  // SplitMix64 (Steele, Lea, Flood 2014): deterministic mutation positions
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
};

// Every prefix up to max_prefix bytes (then every step-th one), and n_flips single-byte
// corruptions, each written to `tmp` and handed to `load`.
template <class Load>
int mutate(const std::vector<uint8_t>& good, const std::string& tmp, size_t max_prefix, size_t step, int n_flips,
           uint64_t seed, Load load) {
    int runs = 0;
    for (size_t n = 0; n < good.size(); n += (n < max_prefix ? 1 : step)) {
        spit(tmp, good.data(), n);
        load(tmp);
        ++runs;
    }
    Mix r{seed};
    std::vector<uint8_t> buf;
    for (int k = 0; k < n_flips && !good.empty(); ++k) {
        buf = good;
        const size_t at = r.next() % buf.size();
        const uint64_t v = r.next();
        buf[at] = (v & 1) ? (uint8_t)(buf[at] ^ (1u << ((v >> 1) & 7))) : (uint8_t)(v >> 8);
        spit(tmp, buf.data(), buf.size());
        load(tmp);
        ++runs;
    }
    return runs;
}

bool same_scene(const HostScene& a, const HostScene& b) {
    return a.tris.size() == b.tris.size() && a.materials.size() == b.materials.size() &&
           a.points.size() == b.points.size() &&
           std::memcmp(a.tris.data(), b.tris.data(), a.tris.size() * sizeof(lgi_triangle)) == 0 &&
           std::memcmp(&a.camera, &b.camera, sizeof(lgi_camera)) == 0;
}

void scenes(const std::string& dir, const std::string& mock) {
    std::string err;
    HostScene s, j, b;
    CHECK(scene_load_json(mock.c_str(), s, err, false) == LGI_OK, "mock scene: %s", err.c_str());
    CHECK(s.tris.size() == 2 && s.materials.size() == 2, "mock scene counts");
    HostScene g;
    CHECK(scene_generate(1, 7, 64, 32, g, err) == LGI_OK, "generate: %s", err.c_str());
    for (const HostScene* src : {&s, &g}) {
        const std::string jp = dir + "/rt.json", bp = dir + "/rt.lgis";
        CHECK(scene_save_json(jp.c_str(), *src, err) == LGI_OK, "save json: %s", err.c_str());
        CHECK(scene_load_json(jp.c_str(), j, err, false) == LGI_OK && same_scene(*src, j), "json round trip");
        CHECK(scene_save_bin(bp.c_str(), *src, err) == LGI_OK, "save bin: %s", err.c_str());
        CHECK(scene_load_bin(bp.c_str(), b, err, false) == LGI_OK && same_scene(*src, b), "bin round trip");
    }
    const std::string tmp = dir + "/mut.json";
    int ok = 0;
    int runs = mutate(slurp(mock), tmp, 1u << 20, 1, 3000, 1, [&](const std::string& p) {
        HostScene t;
        std::string e;
        ok += scene_load_json(p.c_str(), t, e, false) == LGI_OK;
    });
    std::printf("scene.json: %d mutated inputs, %d still valid\n", runs, ok);
    std::vector<uint8_t> bin = slurp(dir + "/rt.lgis");  // the generated scene, ~6.5 MB
    bin.resize(std::min<size_t>(bin.size(), 64 + 4096));  // header and the first records
    spit(dir + "/head.lgis", bin.data(), bin.size());
    runs = mutate(bin, dir + "/mut.lgis", 512, 61, 2000, 2, [&](const std::string& p) {
        HostScene t;
        std::string e;
        (void)scene_load_bin(p.c_str(), t, e, false);
    });
    std::printf("binary scene: %d mutated inputs\n", runs);
}

void pngs(const std::string& dir, int n_extra, char** extra) {
    const int W = 37, H = 23;
    std::vector<uint8_t> rgb((size_t)W * H * 3);
    Mix r{3};
    for (auto& v : rgb) v = (uint8_t)r.next();
    const std::string p = dir + "/img.png";
    CHECK(write_png_rgb(p.c_str(), rgb.data(), W, H) == LGI_OK, "write png");
    int w = 0, h = 0;
    std::vector<uint8_t> rgba;
    std::string err;
    CHECK(read_png_rgba(p.c_str(), w, h, rgba, err) == LGI_OK && w == W && h == H, "read png: %s", err.c_str());
    bool same = rgba.size() == (size_t)W * H * 4;
    for (size_t i = 0; same && i < (size_t)W * H; ++i)
        same = !std::memcmp(&rgba[4 * i], &rgb[3 * i], 3) && rgba[4 * i + 3] == 255;
    CHECK(same, "png round trip");
    std::vector<std::string> files{p};
    for (int k = 0; k < n_extra; ++k) {
        files.emplace_back(extra[k]);
        CHECK(read_png_rgba(extra[k], w, h, rgba, err) == LGI_OK, "%s: %s", extra[k], err.c_str());
    }
    int runs = 0;
    for (size_t k = 0; k < files.size(); ++k)
        runs += mutate(slurp(files[k]), dir + "/mut.png", 256, 7, 600, 10 + k, [&](const std::string& q) {
            int ww, hh;
            std::vector<uint8_t> px;
            std::string e;
            (void)read_png_rgba(q.c_str(), ww, hh, px, e);
        });
    std::printf("png: %zu files, %d mutated inputs\n", files.size(), runs);
}

void batches(const std::string& dir) {
    const std::string p = dir + "/batch.txt";
    const std::string text =
        "# out spp bounces cx cy cz tx ty tz\n"
        "a/0000.png 8 2 5.5542 -1.7855 1.0116 6.5052 -2.0692 1.1351\n"
        "b/0000.png 256 2 5.5542 -1.7855 1.0116 6.5052 -2.0692 1.1351  # trailing comment\n"
        "short line 1 2\n\n"
        "c/0001.png 8 2 1 2 3 4 5 6\n";
    spit(p, (const uint8_t*)text.data(), text.size());
    std::vector<lgi_batch_job> jobs;
    std::string err;
    CHECK(parse_batch_file(p.c_str(), jobs, err) == LGI_OK && jobs.size() == 3, "batch parse");
    std::string longp = std::string(600, 'x') + " 8 2 1 2 3 4 5 6\n";
    spit(p, (const uint8_t*)longp.data(), longp.size());
    jobs.clear();
    CHECK(parse_batch_file(p.c_str(), jobs, err) == LGI_ERR_INVALID_ARG, "over-long output path refused");
    const std::vector<uint8_t> good(text.begin(), text.end());
    const int runs = mutate(good, dir + "/mut.txt", 1u << 20, 1, 1000, 4, [&](const std::string& q) {
        std::vector<lgi_batch_job> jj;
        std::string e;
        (void)parse_batch_file(q.c_str(), jj, e);
    });
    std::printf("batch: %d mutated inputs\n", runs);
}

void oracle_render() {
    HostScene g;
    std::string err;
    CHECK(scene_generate(1, 5, 48, 24, g, err) == LGI_OK, "generate");
    const int64_t n = (int64_t)g.tris.size();
    std::vector<ONode> nodes(2 * n);
    std::vector<int32_t> perm(n);
    std::vector<uint32_t> keys(n);
    const int64_t nn = orc_bvh_build((const OTri*)g.tris.data(), n, nodes.data(), perm.data(), keys.data());
    CHECK(nn > 0 && nn < 2 * n, "oracle bvh");
    std::vector<OTri> sorted(n);
    for (int64_t i = 0; i < n; ++i) std::memcpy(&sorted[i], &g.tris[perm[i]], sizeof(OTri));
    std::vector<float> albedo;
    for (auto& m : g.materials) albedo.insert(albedo.end(), m.albedo, m.albedo + 3);
    OrcRenderArgs a{};
    a.nodes = nodes.data();
    a.tris = sorted.data();
    a.mat_albedo = albedo.data();
    a.points = (const OPoint*)g.points.data();
    a.n_points = (int32_t)g.points.size();
    std::memcpy(&a.sun, &g.sun, sizeof(OSun));
    std::memcpy(&a.cam, &g.camera, sizeof(OCam));
    a.spp = 4;
    a.bounces = 2;
    a.row_step = 1;
    a.threads = 2;
    const size_t npx = (size_t)g.camera.width * g.camera.height;
    std::vector<float> sum(npx * 3), comp(npx * 3), sq(npx);
    std::vector<double> s64(npx * 3);
    uint64_t cnt[5] = {};
    orc_render(&a, sum.data(), comp.data(), s64.data(), sq.data(), cnt);
    CHECK(cnt[3] > 0, "oracle rendered no rays");
    std::vector<uint8_t> alb(npx * 3), nrm(npx * 3), wp(npx * 3), rgb(npx * 3);
    orc_gbuffer(&a, alb.data(), nrm.data(), wp.data());
    orc_tonemap(sum.data(), (int64_t)npx, 0.25f, rgb.data());
    std::printf("oracle: %lld nodes, %llu rays\n", (long long)nn, (unsigned long long)cnt[3]);
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s <work dir> <scene.json> [png ...]\n", argv[0]);
        return 2;
    }
    const std::string dir = argv[1];
    scenes(dir, argv[2]);
    pngs(dir, argc - 3, argv + 3);
    batches(dir);
    oracle_render();
    if (failures) {
        std::fprintf(stderr, "%d checks failed\n", failures);
        return 1;
    }
    std::printf("harness ok\n");
    return 0;
}
