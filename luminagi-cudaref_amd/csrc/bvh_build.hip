// bvh_build.hip — GPU LBVH for gfx950, bit-exact to the reference CPU builder
// (src/bvh.cpp:11-211): same 30-bit Morton codes, same split positions, same leaf rule
// (<= 4 triangles, or a run of identical codes), same BFS node numbering.
//
// Pipeline (all on the context stream):
//   1. scene AABB (block min/max reduction)                       src/bvh.cpp:170-178
//   2. centroid Morton codes                                       src/bvh.cpp:180-192
//   3. stable LSD radix sort (key, triangle) — 4 x 8-bit passes    src/bvh.cpp:194-195
//   4. Karras binary radix tree over (key, index)-augmented codes; its split positions
//      are the reference find_split positions (src/bvh.cpp:55-65, 84-90)
//   5. reference nodes = root + children of "splittable" Karras nodes (count > 4 and
//      first key != last key), i.e. the reference leaf rule (src/bvh.cpp:92)
//   6. BFS numbering (src/bvh.cpp:124-162): sort reference nodes by (depth, first index);
//      within a level, left-to-right order is range order
//   7. bottom-up AABB refit, one launch per level from the deepest up (no atomics); leaf
//      boxes from the triangle vertices in order, parent box = left box then strict-union
//      right box (src/bvh.cpp:96-103, 115-117)
//   8. treelet restructuring of a copy of the tree for traversal (below); the reference node
//      array itself stays as built
//   9. traversal layout: TNode (child boxes in the parent), TriV (v0, e1, e2), TriS.
#include <algorithm>
#include <cstring>

#include "lgi_internal.h"
#include "lgi_math.h"
#include "trace_kernels.h"

namespace {

constexpr int kBlock = 256;
// 16 keys per thread: 4096-key tiles halve the per-tile digit counts (and the scans over them)
// against 8. The 1.05M-triangle frame renders 0.8% faster, while its LBVH build takes 1.42
// instead of 1.35 ms (32 keys: slower in both).
constexpr int kItems = 16;
constexpr int kTile = kBlock * kItems;   // keys per radix-sort block
constexpr int kDepthShift = 26;          // BFS key = depth << 26 | first
constexpr int kMaxLevels = 32;           // reference depth <= 30 (one split per Morton bit) + leaf
// Deepest tree the traversal stacks (kStackMax entries, one push per internal level) accept;
// the reference tree never exceeds it (a deeper one is rejected at step 6).
constexpr int kMaxTraversalDepth = kMaxLevels - 1;
static_assert(kMaxTraversalDepth <= kStackMax, "one stack entry per internal level");

inline int ceil_div(int64_t a, int64_t b) { return (int)((a + b - 1) / b); }

// ------------------------------------------------------------------ scan (exclusive, u32)
__device__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* s_wave, uint32_t& total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t x = v;
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) s_wave[wave] = x;
    __syncthreads();
    uint32_t wpre = 0, tot = 0;
    for (int w = 0; w < kBlock / 64; ++w) {
        uint32_t t = s_wave[w];
        if (w < wave) wpre += t;
        tot += t;
    }
    __syncthreads();
    total = tot;
    return wpre + x - v;
}

// In-place exclusive scan of one kTile chunk per block; block totals to sums. A thread's kItems
// consecutive values move as 16-byte vectors (data is 16-byte aligned) unless they straddle n.
__global__ __launch_bounds__(kBlock) void k_scan_tiles(uint32_t* data, int n, uint32_t* sums) {
    __shared__ uint32_t s_wave[kBlock / 64];
    const int base = blockIdx.x * kTile + threadIdx.x * kItems;
    const bool whole = base + kItems <= n;
    uint32_t v[kItems];
    uint32_t acc = 0;
    if (whole) {
        const uint4* src = reinterpret_cast<const uint4*>(data + base);
        for (int q = 0; q < kItems / 4; ++q) {
            const uint4 x = src[q];
            v[4 * q] = x.x; v[4 * q + 1] = x.y; v[4 * q + 2] = x.z; v[4 * q + 3] = x.w;
        }
    } else {
        for (int k = 0; k < kItems; ++k) v[k] = (base + k < n) ? data[base + k] : 0u;
    }
    for (int k = 0; k < kItems; ++k) acc += v[k];
    uint32_t total;
    uint32_t pre = block_exclusive_scan(acc, s_wave, total);
    for (int k = 0; k < kItems; ++k) {
        const uint32_t x = v[k];
        v[k] = pre;
        pre += x;
    }
    if (whole) {
        uint4* dst = reinterpret_cast<uint4*>(data + base);
        for (int q = 0; q < kItems / 4; ++q) dst[q] = make_uint4(v[4 * q], v[4 * q + 1], v[4 * q + 2], v[4 * q + 3]);
    } else {
        for (int k = 0; k < kItems; ++k)
            if (base + k < n) data[base + k] = v[k];
    }
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// Single-block exclusive scan of an arbitrary-length array (carry across chunks).
__global__ __launch_bounds__(kBlock) void k_scan_single(uint32_t* data, int n, uint32_t* total_out) {
    __shared__ uint32_t s_wave[kBlock / 64];
    uint32_t carry = 0;
    for (int off = 0; off < n; off += kTile) {
        const int base = off + threadIdx.x * kItems;
        uint32_t v[kItems];
        uint32_t acc = 0;
        for (int k = 0; k < kItems; ++k) {
            int i = base + k;
            v[k] = (i < n) ? data[i] : 0u;
            acc += v[k];
        }
        uint32_t total;
        uint32_t pre = block_exclusive_scan(acc, s_wave, total) + carry;
        for (int k = 0; k < kItems; ++k) {
            int i = base + k;
            if (i < n) data[i] = pre;
            pre += v[k];
        }
        carry += total;
    }
    if (threadIdx.x == 0 && total_out) *total_out = carry;
}

__global__ __launch_bounds__(kBlock) void k_scan_add(uint32_t* data, int n, const uint32_t* sums) {
    const uint32_t add = sums[blockIdx.x];
    const int base = blockIdx.x * kTile + threadIdx.x * kItems;
    if (base + kItems <= n) {
        uint4* d = reinterpret_cast<uint4*>(data + base);
        for (int q = 0; q < kItems / 4; ++q) {
            const uint4 x = d[q];
            d[q] = make_uint4(x.x + add, x.y + add, x.z + add, x.w + add);
        }
        return;
    }
    for (int k = 0; k < kItems; ++k) {
        int i = base + k;
        if (i < n) data[i] += add;
    }
}

// ------------------------------------------------------------------ radix sort
__global__ __launch_bounds__(kBlock) void k_radix_hist(const uint32_t* keys, int n, const uint32_t* n_dev, int shift,
                                                       uint32_t* counts, int nb) {
    __shared__ uint32_t h[256];
    if (n_dev) n = min(n, (int)*n_dev);
    h[threadIdx.x] = 0;
    __syncthreads();
    const int base = blockIdx.x * kTile;
    for (int j = 0; j < kItems; ++j) {
        int i = base + j * kBlock + threadIdx.x;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    counts[(size_t)threadIdx.x * nb + blockIdx.x] = h[threadIdx.x];
}

// Stable scatter of one kTile-key tile. Wave w owns the tile's keys [w * 1024, (w + 1) * 1024),
// 64 at a time in index order, and ranks each key among the wave's earlier keys of its digit:
// a 64-lane match from 8 ballots plus a per-wave running count per digit, with no block barrier
// inside the loop. The tile is then staged in LDS in digit order (digit start in the tile + keys
// of that digit in earlier waves + rank) and written out by consecutive threads to consecutive
// positions of each digit's output run, so the global stores coalesce. The result is the stable
// order: equal digits keep their index order.
constexpr int kWaveKeys = kTile / (kBlock / 64);
__global__ __launch_bounds__(kBlock) void k_radix_scatter(const uint32_t* kin, const int32_t* vin,
                                                          uint32_t* kout, int32_t* vout, int n, const uint32_t* n_dev,
                                                          int shift, const uint32_t* offs, const uint32_t* offs_tiles,
                                                          int nb) {
    __shared__ uint32_t s_wc[kBlock / 64][256];  // per wave: keys of each digit so far, then its prefix
    __shared__ uint32_t s_start[256];            // digit start in the tile
    __shared__ uint32_t s_out[256];              // digit's first output position
    __shared__ uint32_t s_key[kTile];
    __shared__ int32_t s_val[kTile];
    if (n_dev) n = min(n, (int)*n_dev);
    const int base = blockIdx.x * kTile;
    if (base >= n) return;  // whole block past the device-side count
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int w = 0; w < kBlock / 64; ++w) s_wc[w][tid] = 0;
    {
        // offs holds each digit-major count's offset within its kTile chunk of counts, offs_tiles
        // the chunks' exclusive prefix: their sum is the global output position
        const size_t c = (size_t)tid * nb + blockIdx.x;
        s_out[tid] = offs[c] + offs_tiles[c / kTile];
    }
    __syncthreads();
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    uint32_t key[kItems], rank[kItems];
    int32_t val[kItems];
    for (int j = 0; j < kItems; ++j) {
        const int i = base + wave * kWaveKeys + j * 64 + lane;
        const bool valid = i < n;
        key[j] = valid ? kin[i] : 0u;
        val[j] = valid ? vin[i] : 0;
        const uint32_t d = (key[j] >> shift) & 255u;
        uint64_t peers = __ballot(valid);
        for (int b = 0; b < 8; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        // the wave's earlier keys of digit d, then this round's lower lanes of digit d
        const uint32_t before = s_wc[wave][d];
        rank[j] = before + (uint32_t)__popcll(peers & lt_mask);
        __builtin_amdgcn_wave_barrier();
        if (valid && (__ffsll((long long)peers) - 1) == lane) s_wc[wave][d] = before + (uint32_t)__popcll(peers);
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
    // per digit: total in the tile and each wave's prefix; digit starts by a block scan
    uint32_t tot = 0;
    for (int w = 0; w < kBlock / 64; ++w) {
        const uint32_t c = s_wc[w][tid];
        s_wc[w][tid] = tot;
        tot += c;
    }
    {
        uint32_t x = tot;
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t y = __shfl_up(x, o, 64);
            if (lane >= o) x += y;
        }
        s_start[tid] = x;  // inclusive within the wave for now
    }
    __syncthreads();
    {
        uint32_t pre = 0;
        for (int w = 0; w < wave; ++w) pre += s_start[w * 64 + 63];
        const uint32_t excl = pre + s_start[tid] - tot;
        __syncthreads();
        s_start[tid] = excl;
    }
    __syncthreads();
    for (int j = 0; j < kItems; ++j) {
        const int i = base + wave * kWaveKeys + j * 64 + lane;
        if (i < n) {
            const uint32_t d = (key[j] >> shift) & 255u;
            const uint32_t L = s_start[d] + s_wc[wave][d] + rank[j];
            s_key[L] = key[j];
            s_val[L] = val[j];
        }
    }
    __syncthreads();
    const int n_tile = min(kTile, n - base);
    for (int L = tid; L < n_tile; L += kBlock) {
        const uint32_t k = s_key[L];
        const uint32_t d = (k >> shift) & 255u;
        const uint32_t pos = s_out[d] + ((uint32_t)L - s_start[d]);
        kout[pos] = k;
        vout[pos] = s_val[L];
    }
}

// ------------------------------------------------------------------ scene bbox + Morton
__global__ __launch_bounds__(kBlock) void k_bbox_partial(const lgi_triangle* tris, int n, float* partial) {
    __shared__ float s[6][kBlock];
    float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        const lgi_triangle& t = tris[i];
        for (int c = 0; c < 3; ++c)
            for (int k = 0; k < 3; ++k) {
                float p = t.v[c][k];
                mn[k] = fminf(mn[k], p);
                mx[k] = fmaxf(mx[k], p);
            }
    }
    for (int k = 0; k < 3; ++k) { s[k][threadIdx.x] = mn[k]; s[3 + k][threadIdx.x] = mx[k]; }
    __syncthreads();
    for (int o = kBlock / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o)
            for (int k = 0; k < 3; ++k) {
                s[k][threadIdx.x] = fminf(s[k][threadIdx.x], s[k][threadIdx.x + o]);
                s[3 + k][threadIdx.x] = fmaxf(s[3 + k][threadIdx.x], s[3 + k][threadIdx.x + o]);
            }
        __syncthreads();
    }
    if (threadIdx.x < 6) partial[blockIdx.x * 6 + threadIdx.x] = s[threadIdx.x][0];
}

// out = sceneMin[3], ext[3] with ext <= 0 replaced by 1 (src/bvh.cpp:175-178)
__global__ __launch_bounds__(kBlock) void k_bbox_final(const float* partial, int nb, float* out) {
    __shared__ float s[6][kBlock];
    float v[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    for (int b = threadIdx.x; b < nb; b += kBlock)
        for (int k = 0; k < 3; ++k) {
            v[k] = fminf(v[k], partial[b * 6 + k]);
            v[3 + k] = fmaxf(v[3 + k], partial[b * 6 + 3 + k]);
        }
    for (int k = 0; k < 6; ++k) s[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int o = kBlock / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o)
            for (int k = 0; k < 3; ++k) {
                s[k][threadIdx.x] = fminf(s[k][threadIdx.x], s[k][threadIdx.x + o]);
                s[3 + k][threadIdx.x] = fmaxf(s[3 + k][threadIdx.x], s[3 + k][threadIdx.x + o]);
            }
        __syncthreads();
    }
    if (threadIdx.x < 3) {
        const int k = threadIdx.x;
        const float e = s[3 + k][0] - s[k][0];
        out[k] = s[k][0];
        out[3 + k] = (e <= 0.0f) ? 1.0f : e;
    }
}

__device__ __forceinline__ uint32_t spread3(uint32_t v) {
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}
__device__ __forceinline__ uint32_t quant10(float f) {
    f = f * 1024.0f;
    if (f < 0.0f) f = 0.0f;
    if (f > 1023.0f) f = 1023.0f;
    return (uint32_t)f;
}

__global__ __launch_bounds__(kBlock) void k_morton(const lgi_triangle* tris, int n, const float* bb,
                                                   uint32_t* keys, int32_t* vals) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const lgi_triangle& t = tris[i];
    const float third = 1.0f / 3.0f;
    float cx = (t.v[0][0] + t.v[1][0] + t.v[2][0]) * third;
    float cy = (t.v[0][1] + t.v[1][1] + t.v[2][1]) * third;
    float cz = (t.v[0][2] + t.v[1][2] + t.v[2][2]) * third;
    uint32_t x = quant10(lgi_div(cx - bb[0], bb[3]));
    uint32_t y = quant10(lgi_div(cy - bb[1], bb[4]));
    uint32_t z = quant10(lgi_div(cz - bb[2], bb[5]));
    keys[i] = (spread3(x) << 2) | (spread3(y) << 1) | spread3(z);
    vals[i] = i;
}

// Sorted vertex array (9 floats per triangle) in Morton order. Each thread gathers one
// triangle's vertices into LDS; the block then writes its 256 x 9 floats as contiguous rows.
__global__ __launch_bounds__(kBlock) void k_gather_verts(const lgi_triangle* tris, const int32_t* perm, int n,
                                                         float* sv) {
    __shared__ float s_v[kBlock * 9];
    const int i0 = blockIdx.x * kBlock, i = i0 + threadIdx.x;
    if (i < n) {
        const lgi_triangle& t = tris[perm[i]];
        for (int c = 0; c < 3; ++c)
            for (int k = 0; k < 3; ++k) s_v[threadIdx.x * 9 + c * 3 + k] = t.v[c][k];
    }
    __syncthreads();
    const int count = min(kBlock, n - i0) * 9;
    for (int j = threadIdx.x; j < count; j += kBlock) sv[(size_t)i0 * 9 + j] = s_v[j];
}

// ------------------------------------------------------------------ Karras tree
__device__ __forceinline__ int delta(const uint32_t* k, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    const uint32_t a = k[i], b = k[j];
    if (a != b) return __clz(a ^ b);
    return 32 + __clz((uint32_t)i ^ (uint32_t)j);
}

// Internal node i (0..n-2). Node ids: internal i, leaf k -> n-1+k.
__global__ __launch_bounds__(kBlock) void k_karras(const uint32_t* keys, int n, int32_t* child, int32_t* parent,
                                                   int32_t* rfirst, int32_t* rlast, uint8_t* splittable) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n - 1) return;
    const int d = (delta(keys, n, i, i + 1) - delta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
    const int dmin = delta(keys, n, i, i - d);
    int lmax = 2;
    while (delta(keys, n, i, i + lmax * d) > dmin) lmax <<= 1;
    int l = 0;
    for (int t = lmax >> 1; t >= 1; t >>= 1)
        if (delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
    const int j = i + l * d;
    const int dnode = delta(keys, n, i, j);
    int s = 0, t = l;
    do {
        t = (t + 1) >> 1;
        if (delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
    } while (t > 1);
    const int gamma = i + s * d + min(d, 0);
    const int first = min(i, j), last = max(i, j);
    const int left = (first == gamma) ? (n - 1 + gamma) : gamma;
    const int right = (last == gamma + 1) ? (n - 1 + gamma + 1) : gamma + 1;
    child[2 * i] = left;
    child[2 * i + 1] = right;
    parent[left] = i;
    parent[right] = i;
    rfirst[i] = first;
    rlast[i] = last;
    splittable[i] = (last - first + 1 > 4) && (keys[first] != keys[last]);
    if (i == 0) parent[0] = -1;
}

// Reference nodes: the root and every child of a splittable node.
__device__ inline bool is_ref_node(int x, int n, const int32_t* parent, const uint8_t* splittable) {
    if (x == 0) return true;
    if (x >= 2 * n - 1) return false;
    const int p = parent[x];
    return p >= 0 && splittable[p];
}

// Reference-node count of each block of k_ref_nodes (their exclusive scan gives the blocks' output
// offsets: an atomic slot counter shared by all waves serialises in L2).
__global__ __launch_bounds__(kBlock) void k_ref_count(int n, const int32_t* parent, const uint8_t* splittable,
                                                      uint32_t* block_counts) {
    __shared__ uint32_t s_wave[kBlock / 64];
    const int x = blockIdx.x * kBlock + threadIdx.x;
    const uint64_t mask = __ballot(is_ref_node(x, n, parent, splittable));
    if ((threadIdx.x & 63) == 0) s_wave[threadIdx.x >> 6] = (uint32_t)__popcll(mask);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int w = 0; w < kBlock / 64; ++w) t += s_wave[w];
        block_counts[blockIdx.x] = t;
    }
}

// Emits (BFS key, id) of every reference node at its block's offset plus its rank in the block
// (node order; the BFS sort orders the keys, which are unique: depth and first primitive).
__global__ __launch_bounds__(kBlock) void k_ref_nodes(int n, const int32_t* parent, const int32_t* rfirst,
                                                      const uint8_t* splittable, const uint32_t* block_base,
                                                      uint32_t* out_keys, int32_t* out_ids) {
    __shared__ uint32_t s_wave[kBlock / 64];
    const int x = blockIdx.x * kBlock + threadIdx.x;
    const bool ref = is_ref_node(x, n, parent, splittable);
    uint32_t total;
    const uint32_t rank = block_exclusive_scan(ref ? 1u : 0u, s_wave, total);
    if (!ref) return;
    const int first = (x < n - 1) ? rfirst[x] : x - (n - 1);
    int depth = 0;
    for (int y = x; y != 0; y = parent[y]) ++depth;
    const uint32_t slot = block_base[blockIdx.x] + rank;
    out_keys[slot] = ((uint32_t)depth << kDepthShift) | (uint32_t)first;
    out_ids[slot] = x;
}

__global__ void k_init_single(int32_t* parent, int32_t* ids, int32_t* new_index) {
    parent[0] = -1;
    ids[0] = 0;
    new_index[0] = 0;
}

__global__ __launch_bounds__(kBlock) void k_new_index(const int32_t* bfs_ids, int m, int32_t* new_index) {
    const int p = blockIdx.x * kBlock + threadIdx.x;
    if (p < m) new_index[bfs_ids[p]] = p;
}

__device__ __forceinline__ void fit3(float* mn, float* mx, const float* p) {
    for (int k = 0; k < 3; ++k) {
        if (p[k] < mn[k]) mn[k] = p[k];
        if (p[k] > mx[k]) mx[k] = p[k];
    }
}

// Leaf boxes from the triangle vertices in leaf order (src/bvh.cpp:96-103).
__global__ __launch_bounds__(kBlock) void k_leaf_boxes(int n, int m, const int32_t* bfs_ids,
                                                       const int32_t* rfirst, const int32_t* rlast,
                                                       const uint8_t* splittable, const float* sv,
                                                       lgi_bvh_node* nodes) {
    const int p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= m) return;
    const int x = bfs_ids[p];
    const bool leaf = (x >= n - 1) || !splittable[x];
    if (!leaf) return;
    const int first = (x < n - 1) ? rfirst[x] : x - (n - 1);
    const int last = (x < n - 1) ? rlast[x] : first;
    float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int i = first; i <= last; ++i)
        for (int c = 0; c < 3; ++c) fit3(mn, mx, sv + (size_t)i * 9 + c * 3);
    lgi_bvh_node nd;
    for (int k = 0; k < 3; ++k) { nd.bb_min[k] = mn[k]; nd.bb_max[k] = mx[k]; }
    nd.left = -1; nd.right = -1; nd.first_tri = first; nd.tri_count = last - first + 1;
    nodes[p] = nd;
}

// Internal boxes of one BFS level [p0, p1): left box, then strict union with the right box
// (src/bvh.cpp:115-117). Levels run deepest first, so both children are final. BFS order
// stores each level contiguously, which replaces per-node arrival atomics and the
// agent-scope release/acquire pairs they need across the eight XCD L2 caches.
__global__ __launch_bounds__(kBlock) void k_level_boxes(int n, int p0, int p1, const int32_t* bfs_ids,
                                                        const int32_t* new_index, const int32_t* child,
                                                        const uint8_t* splittable, lgi_bvh_node* nodes) {
    const int p = p0 + blockIdx.x * kBlock + threadIdx.x;
    if (p >= p1) return;
    const int x = bfs_ids[p];
    if (x >= n - 1 || !splittable[x]) return;
    const int il = new_index[child[2 * x]], ir = new_index[child[2 * x + 1]];
    const lgi_bvh_node L = nodes[il];
    const lgi_bvh_node R = nodes[ir];
    lgi_bvh_node o;
    for (int k = 0; k < 3; ++k) {
        o.bb_min[k] = L.bb_min[k];
        o.bb_max[k] = L.bb_max[k];
        if (R.bb_min[k] < o.bb_min[k]) o.bb_min[k] = R.bb_min[k];
        if (R.bb_max[k] > o.bb_max[k]) o.bb_max[k] = R.bb_max[k];
    }
    o.left = il; o.right = ir; o.first_tri = -1; o.tri_count = 0;
    nodes[p] = o;
}

// level_start[d] = first BFS position of depth d; level_start[max_depth + 1] = m.
__global__ __launch_bounds__(kBlock) void k_level_starts(const uint32_t* bfs_keys, int m, int32_t* level_start) {
    const int p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= m) return;
    const int d = (int)(bfs_keys[p] >> kDepthShift);
    if (p == 0 || (int)(bfs_keys[p - 1] >> kDepthShift) != d) level_start[d] = p;
    if (p == m - 1) level_start[d + 1] = m;
}

// ------------------------------------------------------------------ treelet restructuring
// The traversal tree is a copy of the reference tree whose topology above the leaves is
// rebuilt treelet by treelet to lower its surface-area cost, after Karras and Aila, "Fast
// Parallel Construction of High-Quality Bounding Volume Hierarchies" (HPG 2013): a treelet
// is a node plus descendants, grown by repeatedly opening the treelet leaf of largest surface
// area until it has kTreeletLeaves leaves; dynamic programming over the leaf subsets finds the
// binary tree over those leaves of least cost, C(S) = area(S) * kSahNode + min over splits of
// C(P) + C(S \ P), and the treelet's internal node slots are rewired to it when it is cheaper.
// Leaves (triangle ranges) never change, and every box stays the exact FP32 union of its
// children's, so a ray enters a leaf's ancestors whenever it enters the leaf, and the closest
// hit is the same as on the reference tree except where visit order decides it: exact
// distance ties, and a hit inside a box the slab test rounds to missed.
// Every internal node of the tree as it stands at the start of a pass is a treelet root once in
// that pass, after all of its descendants: a root's treelet lies inside its own subtree (whose
// node set restructuring below it never changes), so roots in disjoint subtrees commute and
// every children-first order gives the same tree. The passes here take the levels of the tree,
// deepest first, one launch per level and one wave per treelet. (A single bottom-up launch per
// pass, where the second child to arrive at a node continues upwards, gives the same tree but
// ran 4x slower: every step up needs an agent-scope release and acquire, which write back and
// invalidate the L2 of the XCD.)
constexpr float kSahNode = 1.2f, kSahTri = 1.0f;
// 5 leaves (31 subsets): on the 1.05M-triangle frame two passes traverse as fast as with 7-leaf
// treelets (127 subsets; 19.24-19.27 vs 19.27-19.32 ms) and restructure in 1.90 instead of 2.49 ms.
// 6 leaves: 19.18-19.24 ms, 2.21 ms; 4 leaves: 19.50 ms, 1.68 ms; one 5-leaf pass: 19.56 ms, 1.00 ms.
constexpr int kTreeletLeaves = 5;
constexpr int kSubsets = 1 << kTreeletLeaves;
// A wave works on 4 treelets at once, 16 lanes each: forming one takes up to 8 lanes (the
// prefetch's deepest level) and its dynamic program at most 10 (the 10 subsets of 2 or of 3
// leaves, one lane each; the full set of 5 takes 2). The wide levels (up to ~85 k treelets) are
// bound by one treelet's chain of dependent loads, so four chains per wave.
constexpr int kTreeletLanes = 16, kTreeletsPerWave = 64 / kTreeletLanes;

__device__ __forceinline__ float half_area(const float* mn, const float* mx) {
    const float dx = mx[0] - mn[0], dy = mx[1] - mn[1], dz = mx[2] - mn[2];
    return dx * dy + dy * dz + dz * dx;
}

// A treelet as its root's lane formed it: the internal slots (in[0] = the root) and the
// leaves, each leaf's box and subtree cost, and the root's refreshed cost. An odd number of
// words, so that the 64 lanes' records start in different LDS banks.
struct TreeletForm {
    int leaf[kTreeletLeaves], in[kTreeletLeaves - 1], n;
    float root_cost;
    float mn[kTreeletLeaves][3], mx[kTreeletLeaves][3], c[kTreeletLeaves];
    int pad;
};
static_assert(sizeof(TreeletForm) / 4 % 2 == 1, "odd word count");
// Working state while a treelet is formed: the root's subtree to depth 3, prefetched by the
// wave in heap order (slot 1 = the root, slots 2h and 2h+1 = the children of slot h; valid only
// below an internal node), and per treelet leaf its children, area, whether it is internal,
// its subtree cost and its heap slot (0: outside the prefetched levels).
constexpr int kHeapSlots = 16;
struct TreeletGrow {
    lgi_bvh_node node[kHeapSlots];
    float cost[kHeapSlots];
    int kid[kTreeletLeaves][2];
    float area[kTreeletLeaves];
    int open[kTreeletLeaves];
    int heap[kTreeletLeaves];
};
// Dynamic programming over the leaf subsets of the treelet being optimized.
struct TreeletDp {
    float mn[kSubsets][3], mx[kSubsets][3], a[kSubsets], c[kSubsets];
    uint8_t part[kSubsets];
};

// The subsets of the treelet leaves ordered by size: sub[start[s] + i] is the i-th of size s.
struct SubsetsBySize {
    uint8_t sub[kSubsets];
    uint8_t start[kTreeletLeaves + 2];
};
constexpr SubsetsBySize subsets_by_size() {
    SubsetsBySize t{};
    int k = 0;
    for (int s = 0; s <= kTreeletLeaves; ++s) {
        t.start[s] = (uint8_t)k;
        for (int v = 0; v < kSubsets; ++v) {
            int c = 0;
            for (int b = v; b; b &= b - 1) ++c;
            if (c == s) t.sub[k++] = (uint8_t)v;
        }
    }
    t.start[kTreeletLeaves + 1] = (uint8_t)k;
    return t;
}
__constant__ SubsetsBySize kBySize = subsets_by_size();

// Records node x (loaded as c, subtree cost cx, heap slot h or 0) as treelet leaf `at`.
__device__ __forceinline__ void treelet_leaf(TreeletForm& F, TreeletGrow& G, int at, int x, const lgi_bvh_node& c,
                                             float cx, int h) {
    F.leaf[at] = x;
    for (int k = 0; k < 3; ++k) {
        F.mn[at][k] = c.bb_min[k];
        F.mx[at][k] = c.bb_max[k];
    }
    F.c[at] = cx;
    G.open[at] = c.tri_count == 0;
    G.area[at] = half_area(c.bb_min, c.bb_max);
    G.kid[at][0] = c.left;
    G.kid[at][1] = c.right;
    G.heap[at] = h;
}

// Forms the treelet of internal node p (the subtrees below it are final for this pass); the
// kTreeletLanes lanes of its group call it (lane: 0..15 within the group). The group first
// loads p's subtree to depth 3 with the subtree costs, one level per round, so that most
// treelets form from LDS. Lane 0 then refreshes p's cost and opens the internal treelet leaf
// of largest area (the first on ties) until there are kTreeletLeaves leaves, loading children
// from memory only below the prefetched levels.
__device__ void treelet_form(const lgi_bvh_node* nodes, float* cost, int p, TreeletForm& F, TreeletGrow& G,
                             int lane) {
    static_assert(kTreeletLanes >= 8, "the prefetch's deepest level takes 8 lanes");
    if (lane == 0) G.node[1] = nodes[p];
    __syncthreads();
    for (int lv = 1; lv <= 3; ++lv) {
        const int h = (1 << lv) + lane, up = h >> 1;
        if (lane < (1 << lv)) {
            if (G.node[up].tri_count == 0) {
                const int x = (h & 1) ? G.node[up].right : G.node[up].left;
                G.node[h] = nodes[x];
                G.cost[h] = cost[x];
            } else {
                G.node[h].tri_count = 1;  // below a leaf: no node (read as a leaf)
            }
        }
        __syncthreads();
    }
    if (lane != 0) return;
    const lgi_bvh_node& nd = G.node[1];
    F.root_cost = half_area(nd.bb_min, nd.bb_max) * kSahNode + G.cost[2] + G.cost[3];
    cost[p] = F.root_cost;
    int nl = 2, ni = 1;
    treelet_leaf(F, G, 0, nd.left, G.node[2], G.cost[2], 2);
    treelet_leaf(F, G, 1, nd.right, G.node[3], G.cost[3], 3);
    F.in[0] = p;
    while (nl < kTreeletLeaves) {
        int best = -1;
        float ba = -1.0f;
        for (int i = 0; i < nl; ++i)
            if (G.open[i] && G.area[i] > ba) { ba = G.area[i]; best = i; }
        if (best < 0) break;
        const int x = F.leaf[best], l = G.kid[best][0], r = G.kid[best][1], h = G.heap[best];
        F.in[ni++] = x;
        if (h > 0 && 2 * h + 1 < kHeapSlots) {
            const lgi_bvh_node cl = G.node[2 * h], cr = G.node[2 * h + 1];
            const float kl = G.cost[2 * h], kr = G.cost[2 * h + 1];
            treelet_leaf(F, G, best, l, cl, kl, 2 * h);
            treelet_leaf(F, G, nl++, r, cr, kr, 2 * h + 1);
        } else {
            const lgi_bvh_node cl = nodes[l], cr = nodes[r];
            const float kl = cost[l], kr = cost[r];
            treelet_leaf(F, G, best, l, cl, kl, 0);
            treelet_leaf(F, G, nl++, r, cr, kr, 0);
        }
    }
    F.n = nl;
}

// The kTreeletLanes lanes of a group find the least-cost topology of the formed treelet F and,
// when it is cheaper than the current one, lane 0 rewires F's internal slots to it.
__device__ void treelet_optimize(lgi_bvh_node* nodes, float* cost, const TreeletForm& F, TreeletDp& S,
                                 const SubsetsBySize& by_size, int lane) {
    const int nl = F.n, ns = 1 << nl;
    if (lane < nl) S.c[1 << lane] = F.c[lane];
    {
        // boxes and areas of every subset of two or more leaves, from the leaf boxes held in
        // registers; the union runs from the lowest leaf up
        float lmn[kTreeletLeaves][3], lmx[kTreeletLeaves][3];
        for (int i = 0; i < kTreeletLeaves; ++i)
            for (int k = 0; k < 3; ++k) {
                lmn[i][k] = F.mn[i][k];
                lmx[i][k] = F.mx[i][k];
            }
        for (int sub = lane; sub < ns; sub += kTreeletLanes) {
            if (__popc(sub) < 2) continue;
            float mn[3] = {0.f, 0.f, 0.f}, mx[3] = {0.f, 0.f, 0.f};
            const int first = __ffs(sub) - 1;
#pragma unroll
            for (int i = 0; i < kTreeletLeaves; ++i) {
                if (!((sub >> i) & 1)) continue;
                for (int k = 0; k < 3; ++k) {
                    mn[k] = i == first ? lmn[i][k] : fminf(mn[k], lmn[i][k]);
                    mx[k] = i == first ? lmx[i][k] : fmaxf(mx[k], lmx[i][k]);
                }
            }
            for (int k = 0; k < 3; ++k) { S.mn[sub][k] = mn[k]; S.mx[sub][k] = mx[k]; }
            S.a[sub] = half_area(mn, mx);
        }
    }
    __syncthreads();
    // Least cost of every subset, by size. A subset's candidate splits put its lowest leaf on
    // the left; larger subsets spread their splits over a group of lanes (1, 2, 8 or 64) and
    // reduce. The winner is the least sum below 3.4e38, the largest left set on ties, else the
    // lowest leaf alone; which lane evaluated a split does not change it.
    for (int size = 2; size <= nl; ++size) {
        const int la = size <= 4 ? 0 : size == 5 ? 1 : size == 6 ? 3 : 6;  // log2 of the lanes per subset
        static_assert(kTreeletLeaves <= 5 && kTreeletLanes >= 10, "a group's lanes cover every subset of a size");
        const int g = 1 << la, which = lane >> la, part = lane & (g - 1);
        const int k0 = by_size.start[size], count = by_size.start[size + 1] - k0;
        const int sub = which < count ? by_size.sub[k0 + which] : 0;
        const bool mine = which < count && sub < ns;  // all of its leaves exist
        const int low = sub & -sub, rest = sub ^ low;
        // left = low + any subset of rest but rest itself: the la lowest leaves of rest are
        // fixed by the lane (the bits of part), the others are enumerated
        int lo = 0, base = 0;
        for (int t = 0, b = rest; t < la; ++t, b &= b - 1) {
            lo |= b & -b;
            if ((part >> t) & 1) base |= b & -b;
        }
        const int hi = rest ^ lo, trips = 1 << (size - 1 - la);
        float best = 3.4e38f;
        int bp = low;
        if (mine) {
            for (int it = 0, r = 0; it < trips; it += 2) {  // two splits per round
                const int r1 = (r - hi) & hi;
                const int q0 = low | base | r, q1 = low | base | r1;
                const float v0 = S.c[q0] + S.c[sub ^ q0];
                const float v1 = it + 1 < trips ? S.c[q1] + S.c[sub ^ q1] : 3.4e38f;
                if (q0 != sub && (v0 < best || (v0 == best && v0 < 3.4e38f && q0 > bp))) { best = v0; bp = q0; }
                if (it + 1 < trips && q1 != sub && (v1 < best || (v1 == best && v1 < 3.4e38f && q1 > bp))) {
                    best = v1;
                    bp = q1;
                }
                r = (r1 - hi) & hi;
            }
        }
        for (int off = 1; off < g; off <<= 1) {
            const float ov = __shfl_xor(best, off, 64);
            const int oq = __shfl_xor(bp, off, 64);
            if (ov < best || (ov == best && ov < 3.4e38f && oq > bp)) { best = ov; bp = oq; }
        }
        if (mine && part == 0) {
            S.c[sub] = S.a[sub] * kSahNode + best;
            S.part[sub] = (uint8_t)bp;
        }
        __syncthreads();
    }
    if (lane == 0 && S.c[ns - 1] < F.root_cost * (1.0f - 1e-6f)) {
        // rewire: internal slot in[0] (the root) takes the full set, the others the inner splits
        int st_sub[kTreeletLeaves], st_node[kTreeletLeaves], sp = 0, next = 1;
        st_sub[sp] = ns - 1;
        st_node[sp++] = F.in[0];
        while (sp > 0) {
            --sp;
            const int sub = st_sub[sp], idx = st_node[sp];
            const int l = S.part[sub], r = sub ^ l;
            int ch[2];
            const int parts[2] = {l, r};
            for (int k = 0; k < 2; ++k) {
                if (__popc(parts[k]) == 1) {
                    ch[k] = F.leaf[__ffs(parts[k]) - 1];
                } else {
                    ch[k] = F.in[next++];
                    st_sub[sp] = parts[k];
                    st_node[sp++] = ch[k];
                }
            }
            lgi_bvh_node& nd = nodes[idx];
            nd.left = ch[0];
            nd.right = ch[1];
            for (int k = 0; k < 3; ++k) {
                nd.bb_min[k] = S.mn[sub][k];
                nd.bb_max[k] = S.mx[sub][k];
            }
            cost[idx] = S.c[sub];
        }
    }
    __syncthreads();  // the next treelet reuses S
}

// Cost of every leaf (area x triangles); internal nodes get theirs when they are processed.
__global__ __launch_bounds__(kBlock) void k_leaf_cost(const lgi_bvh_node* nodes, int m, float* cost) {
    const int p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= m) return;
    const lgi_bvh_node nd = nodes[p];
    if (nd.tri_count > 0) cost[p] = half_area(nd.bb_min, nd.bb_max) * kSahTri * (float)nd.tri_count;
}

// Restructures the treelets rooted at order[i] (i when order is null) for i in [p0, p1), four
// per wave (one per group of kTreeletLanes lanes). The treelets of one level are disjoint.
__global__ __launch_bounds__(64) void k_treelet_level(lgi_bvh_node* nodes, const int32_t* order, int p0, int p1,
                                                      float* cost) {
    __shared__ TreeletForm F[kTreeletsPerWave];
    __shared__ union {
        TreeletGrow g;
        TreeletDp dp;
    } U[kTreeletsPerWave];
    __shared__ SubsetsBySize by_size;
    // the subset table goes to LDS while the treelets form: the DP indexes it per lane, which
    // from constant memory would be a vector memory round trip per subset size
    for (int k = threadIdx.x; k < (int)sizeof(SubsetsBySize); k += 64)
        reinterpret_cast<uint8_t*>(&by_size)[k] = reinterpret_cast<const uint8_t*>(&kBySize)[k];
    const int grp = threadIdx.x / kTreeletLanes, lane = threadIdx.x % kTreeletLanes;
    const int i = p0 + blockIdx.x * kTreeletsPerWave + grp;
    if (i >= p1) return;
    const int p = order ? order[i] : i;
    if (nodes[p].tri_count > 0) return;
    treelet_form(nodes, cost, p, F[grp], U[grp].g, lane);
    __syncthreads();
    if (F[grp].n >= 3) treelet_optimize(nodes, cost, F[grp], U[grp].dp, by_size, lane);  // two leaves: one tree
}

// Levels of the current tree, in parallel: every slot finds its depth by walking its parent
// chain, then one stable 8-bit radix pass orders the slots by depth (by slot index within a
// level); the level-start pass writes the deepest depth to info[0]. info[1] is an error flag: set when a child ref is
// out of range, a slot other than the root has no parent or two, or a walk does not reach the
// root within kMaxWalk steps (a cycle). Only a tree of m slots passes.
constexpr int kMaxWalk = 63;
__global__ __launch_bounds__(kBlock) void k_tree_init(int m, int32_t* parent, uint32_t* nparent, int32_t* info) {
    const int p = blockIdx.x * kBlock + threadIdx.x;
    if (p < m) {
        parent[p] = -1;
        nparent[p] = 0u;
    }
    if (p < 2) info[p] = 0;
}

__global__ __launch_bounds__(kBlock) void k_tree_parents(const lgi_bvh_node* nodes, int m, int32_t* parent,
                                                         uint32_t* nparent, int32_t* info) {
    const int p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= m || nodes[p].tri_count > 0) return;
    for (int c : {nodes[p].left, nodes[p].right}) {
        if (c <= 0 || c >= m) {
            info[1] = 1;
            continue;
        }
        parent[c] = p;
        atomicAdd(&nparent[c], 1u);
    }
}

__global__ __launch_bounds__(kBlock) void k_tree_depth(int m, const int32_t* parent, const uint32_t* nparent,
                                                       uint32_t* depth, int32_t* slot, int32_t* info) {
    const int p = blockIdx.x * kBlock + threadIdx.x;
    int d = 0;
    if (p > 0 && p < m) {
        if (nparent[p] != 1u) info[1] = 1;
        int x = p;
        while (x > 0 && d < kMaxWalk) {  // depth <= kMaxWalk: level starts fit [0, kMaxWalk + 1]
            x = parent[x];
            ++d;
        }
        if (x != 0) {
            info[1] = 1;
            d = 0;
        }
    }
    if (p < m) {
        depth[p] = (uint32_t)d;
        slot[p] = p;
    }
}

// Level starts of the depth-sorted slots; the last slot holds the maximum depth (to max_depth).
__global__ __launch_bounds__(kBlock) void k_depth_starts(const uint32_t* depth_sorted, int m, int32_t* level_start,
                                                         int32_t* max_depth) {
    const int p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= m) return;
    const int d = (int)depth_sorted[p];
    if (p == 0 || (int)depth_sorted[p - 1] != d) level_start[d] = p;
    if (p == m - 1) {
        level_start[d + 1] = m;
        *max_depth = d;
    }
}

// Traversal numbering of the restructured tree: its internal nodes level by level (order[] from
// the depth sort), as the reference tree's are in its own array.
__global__ __launch_bounds__(kBlock) void k_bfs_flags(const lgi_bvh_node* nodes, const int32_t* order, int m,
                                                      uint32_t* flags) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < m) flags[i] = nodes[order[i]].tri_count == 0 ? 1u : 0u;
}

__global__ __launch_bounds__(kBlock) void k_bfs_index(const lgi_bvh_node* nodes, const int32_t* order, int m,
                                                      const uint32_t* rank, uint32_t* idx) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < m && nodes[order[i]].tri_count == 0) idx[order[i]] = rank[i];
}

// ------------------------------------------------------------------ traversal layout
__global__ __launch_bounds__(kBlock) void k_internal_flags(const lgi_bvh_node* nodes, int m, uint32_t* flags,
                                                           uint8_t* leaf_end) {
    const int p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= m) return;
    const lgi_bvh_node nd = nodes[p];
    flags[p] = nd.tri_count == 0 ? 1u : 0u;
    if (nd.tri_count > 0) leaf_end[nd.first_tri + nd.tri_count - 1] = 1;
}

__device__ __forceinline__ int child_ref(const lgi_bvh_node& c, const uint32_t* tidx, int ci) {
    return c.tri_count == 0 ? (int)tidx[ci] : ~c.first_tri;
}

__global__ __launch_bounds__(kBlock) void k_tnodes(const lgi_bvh_node* nodes, int m, const uint32_t* tidx,
                                                   TNode* out) {
    const int p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= m) return;
    const lgi_bvh_node nd = nodes[p];
    if (nd.tri_count != 0) return;
    const lgi_bvh_node a = nodes[nd.left], b = nodes[nd.right];
    TNode t;
    t.b0xy = make_float4(a.bb_min[0], a.bb_max[0], a.bb_min[1], a.bb_max[1]);
    t.b1xy = make_float4(b.bb_min[0], b.bb_max[0], b.bb_min[1], b.bb_max[1]);
    t.bz = make_float4(a.bb_min[2], a.bb_max[2], b.bb_min[2], b.bb_max[2]);
    t.ch = make_int4(child_ref(a, tidx, nd.left), child_ref(b, tidx, nd.right), 0, 0);
    out[tidx[p]] = t;
}

// Root-is-leaf case: one TNode whose two children are both the root leaf.
__global__ void k_tnode_root_leaf(const lgi_bvh_node* nodes, TNode* out) {
    const lgi_bvh_node a = nodes[0];
    TNode t;
    t.b0xy = make_float4(a.bb_min[0], a.bb_max[0], a.bb_min[1], a.bb_max[1]);
    t.b1xy = t.b0xy;
    t.bz = make_float4(a.bb_min[2], a.bb_max[2], a.bb_min[2], a.bb_max[2]);
    t.ch = make_int4(~a.first_tri, ~a.first_tri, 0, 0);
    out[0] = t;
}

__global__ __launch_bounds__(kBlock) void k_tri_layout(const lgi_triangle* tris, const int32_t* perm, int n,
                                                       const float* sv, const uint8_t* leaf_end, TriV* triv,
                                                       TriS* tris_s) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const float* v = sv + (size_t)i * 9;
    TriV t;
    // e1/e2 exactly as ray_tri forms them (src/pathtracer.cu:86-88)
    t.v0 = make_float4(v[0], v[1], v[2], __int_as_float(leaf_end[i] ? 1 : 0));
    t.e1 = make_float4(v[3] - v[0], v[4] - v[1], v[5] - v[2], __int_as_float(perm[i]));
    t.e2 = make_float4(v[6] - v[0], v[7] - v[1], v[8] - v[2], 0.0f);
    triv[i] = t;
    const lgi_triangle& src = tris[perm[i]];
    TriS s;
    for (int c = 0; c < 3; ++c)
        for (int k = 0; k < 3; ++k) s.n[c * 3 + k] = src.n[c][k];
    s.mat = src.material;
    for (int c = 0; c < 3; ++c)
        for (int k = 0; k < 2; ++k) s.uv[c * 2 + k] = src.uv[c][k];
    tris_s[i] = s;
}

struct SortBufs {
    uint32_t* k[2];
    int32_t* v[2];
    uint32_t* counts;
    uint32_t* sums;
};

// Stable LSD radix sort of (keys, vals) in buffer 0 over `bits` low bits (a multiple of 8);
// returns the buffer holding the result (0 when bits is a multiple of 16).
// n_dev (nullable): the key count on the device, at most n; the grids are sized for n.
int radix_sort(hipStream_t st, SortBufs& b, int n, int bits, const uint32_t* n_dev = nullptr) {
    if (n <= 1) return 0;
    const int nb = ceil_div(n, kTile);
    const int ncount = 256 * nb;
    const int nsb = ceil_div(ncount, kTile);
    int cur = 0;
    for (int shift = 0; shift < bits; shift += 8) {
        hipLaunchKernelGGL(k_radix_hist, dim3(nb), dim3(kBlock), 0, st, b.k[cur], n, n_dev, shift, b.counts, nb);
        // digit offsets: scanned within kTile chunks of the counts plus the chunks' scanned totals,
        // added by the scatter itself
        hipLaunchKernelGGL(k_scan_tiles, dim3(nsb), dim3(kBlock), 0, st, b.counts, ncount, b.sums);
        hipLaunchKernelGGL(k_scan_single, dim3(1), dim3(kBlock), 0, st, b.sums, nsb, (uint32_t*)nullptr);
        hipLaunchKernelGGL(k_radix_scatter, dim3(nb), dim3(kBlock), 0, st, b.k[cur], b.v[cur], b.k[cur ^ 1],
                           b.v[cur ^ 1], n, n_dev, shift, b.counts, b.sums, nb);
        cur ^= 1;
    }
    return cur;
}

template <class T>
lgi_status dev_alloc(lgi_ctx* ctx, T** p, size_t count) {
    if (*p) { (void)hipFree(*p); *p = nullptr; }
    if (count == 0) count = 1;
    LGI_HIP(ctx, hipMalloc((void**)p, count * sizeof(T)));
    return LGI_OK;
}

}  // namespace

// Scratch of lgi_sort_pairs: the second key and value buffers, the digit counts and their chunk
// totals, each starting on a 256-byte boundary (the scans move the counts as 16-byte vectors).
static size_t sort_part(size_t bytes) { return (bytes + 255) & ~(size_t)255; }

size_t lgi_sort_scratch_bytes(int64_t n) {
    const int nb = ceil_div(n, kTile);
    const size_t ncount = (size_t)256 * nb;
    return 2 * sort_part(4 * (size_t)n) + sort_part(4 * ncount) + sort_part(4 * ((size_t)ceil_div(ncount, kTile) + 64));
}

const int32_t* lgi_sort_pairs(hipStream_t st, uint32_t* keys, int32_t* vals, int n, int bits, void* scratch,
                              const uint32_t* n_dev) {
    SortBufs b;
    char* p = (char*)scratch;
    const int nb = ceil_div(n, kTile);
    const size_t ncount = (size_t)256 * nb;
    b.k[0] = keys;
    b.v[0] = vals;
    b.k[1] = (uint32_t*)p; p += sort_part(4 * (size_t)n);
    b.v[1] = (int32_t*)p; p += sort_part(4 * (size_t)n);
    b.counts = (uint32_t*)p; p += sort_part(4 * ncount);
    b.sums = (uint32_t*)p;
    return b.v[radix_sort(st, b, n, bits, n_dev)];
}

size_t lgi_scan_scratch_bytes(int64_t n) { return 4 * ((size_t)ceil_div(n, kTile) + 64); }

void lgi_exclusive_scan_chunks(hipStream_t st, uint32_t* data, int n, uint32_t* total, uint32_t* chunk_sums) {
    static_assert(kTile == kScanChunk, "chunk of the two-part scan");
    const int nb = ceil_div(n, kTile);
    hipLaunchKernelGGL(k_scan_tiles, dim3(std::max(1, nb)), dim3(kBlock), 0, st, data, n, chunk_sums);
    hipLaunchKernelGGL(k_scan_single, dim3(1), dim3(kBlock), 0, st, chunk_sums, nb, total);
}

void lgi_exclusive_scan(hipStream_t st, uint32_t* data, int n, uint32_t* total, void* scratch) {
    const int nb = ceil_div(n, kTile);
    uint32_t* sums = (uint32_t*)scratch;
    hipLaunchKernelGGL(k_scan_tiles, dim3(std::max(1, nb)), dim3(kBlock), 0, st, data, n, sums);
    hipLaunchKernelGGL(k_scan_single, dim3(1), dim3(kBlock), 0, st, sums, nb, total);
    hipLaunchKernelGGL(k_scan_add, dim3(std::max(1, nb)), dim3(kBlock), 0, st, data, n, sums);
}

void lgi_free_dev_scene(DevScene& d) {
    void* ptrs[] = {d.tri_in, d.nodes, d.tnodes, d.triv, d.tris, d.mat_albedo, d.points, d.perm, d.keys,
                    d.tex_texels, d.tex_desc, d.srgb_lut};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    d = DevScene();
}

lgi_status lgi_gpu_build_bvh(lgi_ctx* ctx) {
    const int64_t n64 = ctx->dev.n_tris;
    if (n64 <= 0) return lgi_set_err(ctx, LGI_ERR_NO_SCENE, "scene has no triangles");
    if (n64 >= (int64_t(1) << kDepthShift))
        return lgi_set_err(ctx, LGI_ERR_INVALID_ARG, "too many triangles for the GPU LBVH (max 2^26 - 1)");
    const int n = (int)n64;
    hipStream_t st = ctx->stream;
    DevScene& d = ctx->dev;
    lgi_status s;
    if ((s = dev_alloc(ctx, &d.perm, n)) != LGI_OK) return s;
    if ((s = dev_alloc(ctx, &d.keys, n)) != LGI_OK) return s;
    if ((s = dev_alloc(ctx, &d.triv, n)) != LGI_OK) return s;
    if ((s = dev_alloc(ctx, &d.tris, n)) != LGI_OK) return s;
    if ((s = dev_alloc(ctx, &d.nodes, 2 * (size_t)n)) != LGI_OK) return s;

    // scratch
    const int nb_sort = ceil_div(n, kTile);
    const int npart = std::min(1024, ceil_div(n, kBlock));
    auto S = [&](int slot, size_t bytes) { return ctx->scratch_get(slot, bytes); };
    float* partial = (float*)S(1, (size_t)npart * 6 * 4);
    float* bb = (float*)S(2, 6 * 4 + 64);
    uint32_t* counter = (uint32_t*)((char*)bb + 32);
    SortBufs sb;
    const size_t m_cap = 2 * (size_t)n;  // reference nodes <= 2n-1
    sb.k[0] = (uint32_t*)S(3, m_cap * 4);
    sb.k[1] = (uint32_t*)S(4, m_cap * 4);
    sb.v[0] = (int32_t*)S(5, m_cap * 4);
    sb.v[1] = (int32_t*)S(6, m_cap * 4);
    sb.counts = (uint32_t*)S(7, (size_t)256 * ceil_div(m_cap, kTile) * 4 + 64);
    sb.sums = (uint32_t*)S(8, (size_t)(ceil_div(256 * (size_t)ceil_div(m_cap, kTile), kTile) + 16) * 4);
    float* sv = (float*)S(9, (size_t)n * 9 * 4);
    int32_t* child = (int32_t*)S(10, (size_t)2 * n * 4);
    int32_t* parent = (int32_t*)S(11, (size_t)2 * n * 4);
    int32_t* rfirst = (int32_t*)S(12, (size_t)n * 4);
    int32_t* rlast = (int32_t*)S(13, (size_t)n * 4);
    uint8_t* splittable = (uint8_t*)S(14, (size_t)n + 64);
    int32_t* new_index = (int32_t*)S(15, (size_t)2 * n * 4);
    int32_t* level_start = (int32_t*)S(16, (kMaxLevels + 2) * 4);
    uint32_t* tflags = (uint32_t*)S(17, m_cap * 4 + 64);
    uint8_t* leaf_end = (uint8_t*)S(18, (size_t)n + 64);
    uint32_t* tsums = (uint32_t*)S(19, (size_t)(ceil_div(m_cap, kTile) + 16) * 4);
    for (void* p : {(void*)partial, (void*)sb.k[0], (void*)sv, (void*)child, (void*)new_index, (void*)level_start,
                    (void*)tflags, (void*)leaf_end, (void*)tsums})
        if (!p) return lgi_set_err(ctx, LGI_ERR_HIP, "scratch allocation failed");
    (void)nb_sort;

    LGI_HIP(ctx, hipEventRecord(ctx->ev0, st));
    LGI_HIP(ctx, hipMemsetAsync(level_start, 0xFF, (kMaxLevels + 2) * 4, st));
    hipLaunchKernelGGL(k_bbox_partial, dim3(npart), dim3(kBlock), 0, st, d.tri_in, n, partial);
    hipLaunchKernelGGL(k_bbox_final, dim3(1), dim3(kBlock), 0, st, partial, npart, bb);
    hipLaunchKernelGGL(k_morton, dim3(ceil_div(n, kBlock)), dim3(kBlock), 0, st, d.tri_in, n, bb, sb.k[0], sb.v[0]);
    radix_sort(st, sb, n, 32);
    LGI_HIP(ctx, hipMemcpyAsync(d.keys, sb.k[0], (size_t)n * 4, hipMemcpyDeviceToDevice, st));
    LGI_HIP(ctx, hipMemcpyAsync(d.perm, sb.v[0], (size_t)n * 4, hipMemcpyDeviceToDevice, st));
    hipLaunchKernelGGL(k_gather_verts, dim3(ceil_div(n, kBlock)), dim3(kBlock), 0, st, d.tri_in, d.perm, n, sv);

    int64_t m = 1;
    int32_t hls[kMaxLevels + 2];
    int max_depth = 0;
    if (n == 1) {
        // single triangle: one leaf
        hipLaunchKernelGGL(k_init_single, dim3(1), dim3(1), 0, st, parent, sb.v[0], new_index);
        hipLaunchKernelGGL(k_leaf_boxes, dim3(1), dim3(kBlock), 0, st, n, 1, sb.v[0], rfirst, rlast, splittable, sv,
                           d.nodes);
    } else {
        hipLaunchKernelGGL(k_karras, dim3(ceil_div(n - 1, kBlock)), dim3(kBlock), 0, st, d.keys, n, child, parent,
                           rfirst, rlast, splittable);
        // per-block counts and their offsets in tflags (free until the traversal layout), the
        // node count to counter
        const int nb_ref = ceil_div(2 * n - 1, kBlock);
        hipLaunchKernelGGL(k_ref_count, dim3(nb_ref), dim3(kBlock), 0, st, n, parent, splittable, tflags);
        lgi_exclusive_scan(st, tflags, nb_ref, counter, tsums);
        hipLaunchKernelGGL(k_ref_nodes, dim3(nb_ref), dim3(kBlock), 0, st, n, parent, rfirst, splittable, tflags,
                           sb.k[0], sb.v[0]);
        uint32_t hm = 0;
        LGI_HIP(ctx, hipMemcpyAsync(&hm, counter, 4, hipMemcpyDeviceToHost, st));
        LGI_HIP(ctx, hipStreamSynchronize(st));
        m = hm;
        radix_sort(st, sb, (int)m, 32);
        hipLaunchKernelGGL(k_new_index, dim3(ceil_div(m, kBlock)), dim3(kBlock), 0, st, sb.v[0], (int)m, new_index);
        hipLaunchKernelGGL(k_level_starts, dim3(ceil_div(m, kBlock)), dim3(kBlock), 0, st, sb.k[0], (int)m,
                           level_start);
        hipLaunchKernelGGL(k_leaf_boxes, dim3(ceil_div(m, kBlock)), dim3(kBlock), 0, st, n, (int)m, sb.v[0], rfirst,
                           rlast, splittable, sv, d.nodes);
        LGI_HIP(ctx, hipMemcpyAsync(hls, level_start, sizeof(hls), hipMemcpyDeviceToHost, st));
        LGI_HIP(ctx, hipStreamSynchronize(st));
        while (max_depth + 1 <= kMaxLevels && hls[max_depth + 1] != (int)m) ++max_depth;
        if (max_depth >= kMaxLevels) return lgi_set_err(ctx, LGI_ERR_INVALID_ARG, "BVH deeper than expected");
        for (int lv = max_depth - 1; lv >= 0; --lv) {
            const int p0 = hls[lv], p1 = hls[lv + 1];
            hipLaunchKernelGGL(k_level_boxes, dim3(ceil_div(p1 - p0, kBlock)), dim3(kBlock), 0, st, n, p0, p1,
                               sb.v[0], new_index, child, splittable, d.nodes);
        }
    }
    // traversal layout
    LGI_HIP(ctx, hipMemsetAsync(leaf_end, 0, (size_t)n, st));
    hipLaunchKernelGGL(k_internal_flags, dim3(ceil_div(m, kBlock)), dim3(kBlock), 0, st, d.nodes, (int)m, tflags,
                       leaf_end);
    const int nsb = ceil_div(m, kTile);
    hipLaunchKernelGGL(k_scan_tiles, dim3(nsb), dim3(kBlock), 0, st, tflags, (int)m, tsums);
    hipLaunchKernelGGL(k_scan_single, dim3(1), dim3(kBlock), 0, st, tsums, nsb, counter);
    hipLaunchKernelGGL(k_scan_add, dim3(nsb), dim3(kBlock), 0, st, tflags, (int)m, tsums);
    uint32_t n_int = 0;
    LGI_HIP(ctx, hipMemcpyAsync(&n_int, counter, 4, hipMemcpyDeviceToHost, st));
    LGI_HIP(ctx, hipStreamSynchronize(st));
    d.root_is_leaf = (n_int == 0);
    const int64_t n_t = std::max<int64_t>(1, n_int);
    if ((s = dev_alloc(ctx, &d.tnodes, n_t)) != LGI_OK) return s;
    // the tree the traversal layout is made from: the reference tree, or its restructured copy
    const lgi_bvh_node* tsrc = d.nodes;
    const int passes = ctx->bvh_treelet_passes;
    ctx->last_treelet_ms = 0.f;
    if (n_int > 1 && passes > 0) {
        LGI_HIP(ctx, hipEventRecord(ctx->ev_tree0, st));
        lgi_bvh_node* onodes = (lgi_bvh_node*)S(20, (size_t)m * sizeof(lgi_bvh_node));
        float* cost = (float*)S(21, (size_t)m * 4);
        uint32_t* rank = (uint32_t*)S(22, (size_t)m * 4);
        int32_t* parent = (int32_t*)S(24, (size_t)m * 4);
        constexpr int kBfsLevels = kMaxWalk + 1;
        int32_t* lstart = (int32_t*)S(25, (kBfsLevels + 4) * 4);  // level starts, then depth and error
        if (!onodes || !cost || !rank || !parent || !lstart)
            return lgi_set_err(ctx, LGI_ERR_HIP, "scratch allocation failed");
        const int32_t* order = nullptr;
        LGI_HIP(ctx, hipMemcpyAsync(onodes, d.nodes, (size_t)m * sizeof(lgi_bvh_node), hipMemcpyDeviceToDevice, st));
        hipLaunchKernelGGL(k_leaf_cost, dim3(ceil_div(m, kBlock)), dim3(kBlock), 0, st, onodes, (int)m, cost);
        // Pass 1 takes its treelet roots by the reference tree's BFS levels (the array order). A
        // restructured tree no longer follows them (a root of one level may now lie below
        // another), so each later pass first takes the levels of the current tree.
        bool tree_ok = true;
        int32_t lv_host[kBfsLevels + 4];
        const int32_t* lv_order = nullptr;
        int n_lev = max_depth + 1;
        std::copy(hls, hls + n_lev + 1, lv_host);
        // The levels are also taken after the last pass: the traversal uses the restructured tree
        // only when it is a tree of the m slots at most kMaxTraversalDepth levels deep (the
        // traversal stacks hold one entry per internal level). The depth sort reuses the
        // build's sort buffers.
        int32_t* info = lstart + kBfsLevels + 1;
        const int gm = ceil_div(m, kBlock);
        for (int pass = 0; pass <= passes; ++pass) {
            if (pass > 0) {
                hipLaunchKernelGGL(k_tree_init, dim3(gm), dim3(kBlock), 0, st, (int)m, parent, rank, info);
                hipLaunchKernelGGL(k_tree_parents, dim3(gm), dim3(kBlock), 0, st, onodes, (int)m, parent, rank, info);
                hipLaunchKernelGGL(k_tree_depth, dim3(gm), dim3(kBlock), 0, st, (int)m, parent, rank, sb.k[0], sb.v[0],
                                   info);
                const int cur = radix_sort(st, sb, (int)m, 8);
                hipLaunchKernelGGL(k_depth_starts, dim3(gm), dim3(kBlock), 0, st, sb.k[cur], (int)m, lstart,
                                   info);
                LGI_HIP(ctx, hipMemcpyAsync(lv_host, lstart, sizeof(lv_host), hipMemcpyDeviceToHost, st));
                LGI_HIP(ctx, hipStreamSynchronize(st));
                order = lv_order = sb.v[cur];
                if (lv_host[kBfsLevels + 2] != 0 || lv_host[kBfsLevels + 1] > kMaxWalk) { tree_ok = false; break; }
                n_lev = lv_host[kBfsLevels + 1] + 1;
            }
            if (pass == passes) break;
            for (int lv = n_lev - 2; lv >= 0; --lv)
                hipLaunchKernelGGL(k_treelet_level, dim3(ceil_div(lv_host[lv + 1] - lv_host[lv], kTreeletsPerWave)),
                                   dim3(64), 0, st, onodes, lv_order, lv_host[lv], lv_host[lv + 1], cost);
        }
        if (tree_ok && n_lev - 1 <= kMaxTraversalDepth) {
            hipLaunchKernelGGL(k_bfs_flags, dim3(ceil_div(m, kBlock)), dim3(kBlock), 0, st, onodes, order, (int)m,
                               rank);
            hipLaunchKernelGGL(k_scan_tiles, dim3(nsb), dim3(kBlock), 0, st, rank, (int)m, tsums);
            hipLaunchKernelGGL(k_scan_single, dim3(1), dim3(kBlock), 0, st, tsums, nsb, counter);
            hipLaunchKernelGGL(k_scan_add, dim3(nsb), dim3(kBlock), 0, st, rank, (int)m, tsums);
            hipLaunchKernelGGL(k_bfs_index, dim3(ceil_div(m, kBlock)), dim3(kBlock), 0, st, onodes, order, (int)m,
                               rank, tflags);
            tsrc = onodes;
        }
        LGI_HIP(ctx, hipEventRecord(ctx->ev_tree1, st));
        LGI_HIP(ctx, hipEventSynchronize(ctx->ev_tree1));
        LGI_HIP(ctx, hipEventElapsedTime(&ctx->last_treelet_ms, ctx->ev_tree0, ctx->ev_tree1));
    }
    d.trav_nodes = tsrc;
    if (n_int == 0)
        hipLaunchKernelGGL(k_tnode_root_leaf, dim3(1), dim3(1), 0, st, d.nodes, d.tnodes);
    else
        hipLaunchKernelGGL(k_tnodes, dim3(ceil_div(m, kBlock)), dim3(kBlock), 0, st, tsrc, (int)m, tflags, d.tnodes);
    hipLaunchKernelGGL(k_tri_layout, dim3(ceil_div(n, kBlock)), dim3(kBlock), 0, st, d.tri_in, d.perm, n, sv,
                       leaf_end, d.triv, d.tris);
    LGI_HIP(ctx, hipEventRecord(ctx->ev1, st));
    LGI_HIP(ctx, hipGetLastError());
    LGI_HIP(ctx, hipEventSynchronize(ctx->ev1));
    float ms = 0;
    LGI_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    ctx->last_build_ms = ms;
    d.n_nodes = m;
    d.n_tnodes = n_t;
    ctx->bvh_ready = true;
    return LGI_OK;
}
