// unet.hip — denoiser layers for gfx950: 3x3 convolution as implicit GEMM on MFMA
// (v_mfma_f32_16x16x32_bf16, FP32 accumulation), max-pool 2x2, bilinear 2x upsample, the
// 1x1 residual tail, and their backward passes.
//
// Layer semantics follow the reference UNet (ml/train.py:37-75: Conv2d 3x3 pad 1 + bias +
// activation, MaxPool2d(2), 2x upsample, channel concat, 1x1 tail, residual), with the
// denoiser configuration used here: ReLU, bilinear (align_corners=False) upsampling.
//
// Layout: NHWC, BF16 activations, weights [Cout][tap][Cin] (K = tap*Cin + ci contiguous per
// output channel), FP32 bias / master weights / weight gradients.
//
// Convolution tiling: a 256-thread block computes a (rows x 16 columns) pixel x TN-channel
// output tile. Per 32-channel input chunk it stages the input halo and the TN x 9 x 32 weight
// slice in LDS, then its four waves issue the MFMAs from LDS. Two tilings exist (k_conv: 8
// rows, fragments re-read per tap; k_conv_rb: 16 or 32 rows, register-blocked over kernel
// rows), and each layer shape is timed once to pick one. The channel concat of the decoder is
// fused into the halo gather (two source tensors), and the ReLU backward mask is fused into
// the gather of the data-gradient pass.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <cstdint>
#include <map>
#include <mutex>

#include "lgi.h"

namespace {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kBlock = 256;
constexpr int kTH = 8, kTW = 16;               // output tile (rows x cols)
constexpr int kHH = kTH + 2, kHW = kTW + 2;    // input halo tile
constexpr int kCK = 32;                        // input channels per K chunk
constexpr int kWRow = 9 * kCK + 8;             // LDS weight row (bf16), padded against bank conflicts

inline int cdiv(int64_t a, int64_t b) { return (int)((a + b - 1) / b); }

struct ConvArgs {
    const bf16* src0;      // [N,H,W,C0]
    const bf16* src1;      // [N,H,W,C1] or null (channel concat after src0)
    const bf16* mask0;     // optional ReLU-backward masks, same shapes as src0/src1
    const bf16* mask1;
    const bf16* wk;        // [Cout][9][Cin] (or [Cout][1][Cin] for 1x1)
    const float* bias;     // [Cout] or null
    bf16* out0;            // channels [0, Cs)
    bf16* out1;            // channels [Cs, Cout) or null
    int C0, C1, Cout, Cs;
    int N, H, W;
    int relu;
    bf16* pool_out;        // optional [N,H/2,W/2,Cout]: 2x2 max-pool of the output (k_conv32 only)
    // optional denoiser tail on the output (Cout 32, 32x32-MFMA tilings): tail_out[N,H,W,3] =
    // max(expm1(log1p(max(tail_x, 0)) + tail_w . y + tail_b), 0); the 32-channel y is not stored
    const float* tail_w;   // [3][32]
    const float* tail_b;   // [3]
    const float* tail_x;   // [N,H,W,3] radiance
    float* tail_out;
    // FP8 operands (k_conv_f8): wk holds E4M3 weights [Cout][9][Cin] (bytes), w_scale[Cout] their
    // per-output-channel scales, act_scale the per-tensor scale the BF16 inputs are quantized with
    const float* w_scale;
    float act_scale;
    int fp8;
};

__device__ __forceinline__ uint4 zero16() { return make_uint4(0u, 0u, 0u, 0u); }

// 8 bf16 channels (16 bytes) to and from fp32.
__device__ __forceinline__ void unpack8(const uint4& v, float f[8]) {
    const bf16x8 e = *reinterpret_cast<const bf16x8*>(&v);
    for (int k = 0; k < 8; ++k) f[k] = (float)e[k];
}
__device__ __forceinline__ uint4 pack8(const float f[8]) {
    bf16x8 e;
    for (int k = 0; k < 8; ++k) e[k] = (bf16)f[k];
    return *reinterpret_cast<const uint4*>(&e);
}

// Multiply 8 bf16 values by (mask > 0).
__device__ __forceinline__ uint4 relu_mask(uint4 v, uint4 m) {
    const bf16x8 mv = *reinterpret_cast<const bf16x8*>(&m);
    bf16x8 x = *reinterpret_cast<const bf16x8*>(&v);
    for (int i = 0; i < 8; ++i)
        if (!((float)mv[i] > 0.0f)) x[i] = (bf16)0.0f;
    return *reinterpret_cast<const uint4*>(&x);
}

template <int TN, int KS, bool MASK>
__global__ __launch_bounds__(kBlock) void k_conv(ConvArgs a) {
    constexpr int kTaps = KS * KS;
    constexpr int kPad = KS / 2;
    constexpr int kHalH = kTH + KS - 1, kHalW = kTW + KS - 1;
    constexpr int kHalo = kHalH * kHalW * 4;   // 16-byte loads of one chunk's halo
    constexpr int kWLd = TN * kTaps * 4;       // 16-byte loads of one chunk's weights
    constexpr int kHIt = (kHalo + kBlock - 1) / kBlock, kWIt = (kWLd + kBlock - 1) / kBlock;
    __shared__ __attribute__((aligned(16))) bf16 s_in[kHalH * kHalW * kCK];
    __shared__ __attribute__((aligned(16))) bf16 s_w[TN * kWRow];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int tiles_x = (a.W + kTW - 1) / kTW, tiles_y = (a.H + kTH - 1) / kTH;
    const int tile = blockIdx.x;
    const int n = tile / (tiles_x * tiles_y);
    const int ty = (tile / tiles_x) % tiles_y, tx = tile % tiles_x;
    const int y0 = ty * kTH, x0 = tx * kTW;
    const int co0 = blockIdx.y * TN;
    const int Cin = a.C0 + a.C1;
    f32x4 acc[2][TN / 16];
    for (int mi = 0; mi < 2; ++mi)
        for (int ni = 0; ni < TN / 16; ++ni) acc[mi][ni] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int g = lane >> 4, r16 = lane & 15;
    // Software pipeline: the global loads of chunk c + 1 are issued into registers before the
    // MFMAs of chunk c and written to LDS after them, so their latency hides behind the math.
    uint4 hv[kHIt], hm[MASK ? kHIt : 1], wv[kWIt];
    bool masked = false;
    auto load_chunk = [&](int ci0) {
        // source tensor of this chunk (chunks never straddle the concat boundary)
        const bool first = ci0 < a.C0;
        const bf16* src = first ? a.src0 : a.src1;
        const bf16* msk = first ? a.mask0 : a.mask1;
        const int Cs = first ? a.C0 : a.C1;
        const int cc0 = first ? ci0 : ci0 - a.C0;
        masked = MASK && msk != nullptr;
#pragma unroll
        for (int i = 0; i < kHIt; ++i) {
            const int t = tid + i * kBlock;
            hv[i] = zero16();
            if (MASK) hm[i] = zero16();
            if (t < kHalo) {
                const int px = t >> 2, part = t & 3;
                const int hy = px / kHalW, hx = px % kHalW;
                const int gy = y0 + hy - kPad, gx = x0 + hx - kPad;
                if (gy >= 0 && gy < a.H && gx >= 0 && gx < a.W) {
                    const int64_t off = (((int64_t)n * a.H + gy) * a.W + gx) * Cs + cc0 + part * 8;
                    hv[i] = *reinterpret_cast<const uint4*>(src + off);
                    if (MASK && msk) hm[i] = *reinterpret_cast<const uint4*>(msk + off);
                }
            }
        }
#pragma unroll
        for (int i = 0; i < kWIt; ++i) {
            const int t = tid + i * kBlock;
            wv[i] = zero16();
            if (t < kWLd) {
                const int co = t / (kTaps * 4), r = t % (kTaps * 4), tap = r >> 2, part = r & 3;
                wv[i] = *reinterpret_cast<const uint4*>(a.wk + ((int64_t)(co0 + co) * kTaps + tap) * Cin + ci0 +
                                                        part * 8);
            }
        }
    };
    load_chunk(0);
    for (int ci0 = 0; ci0 < Cin; ci0 += kCK) {
        // registers -> LDS: kHalH x kHalW pixels x 32 channels, TN x taps x 32 weights
#pragma unroll
        for (int i = 0; i < kHIt; ++i) {
            const int t = tid + i * kBlock;
            if (t < kHalo) {
                const int px = t >> 2, part = t & 3;
                uint4 v = hv[i];
                if (MASK && masked) v = relu_mask(v, hm[MASK ? i : 0]);
                *reinterpret_cast<uint4*>(&s_in[px * kCK + part * 8]) = v;
            }
        }
#pragma unroll
        for (int i = 0; i < kWIt; ++i) {
            const int t = tid + i * kBlock;
            if (t < kWLd) {
                const int co = t / (kTaps * 4), r = t % (kTaps * 4), tap = r >> 2, part = r & 3;
                *reinterpret_cast<uint4*>(&s_w[co * kWRow + tap * kCK + part * 8]) = wv[i];
            }
        }
        __syncthreads();
        if (ci0 + kCK < Cin) load_chunk(ci0 + kCK);
        for (int tap = 0; tap < kTaps; ++tap) {
            const int ky = tap / KS, kx = tap % KS;
            bf16x8 A[2], B[TN / 16];
            for (int mi = 0; mi < 2; ++mi) {
                const int row = wave * 2 + mi + ky;
                A[mi] = *reinterpret_cast<const bf16x8*>(&s_in[(row * kHalW + r16 + kx) * kCK + g * 8]);
            }
            for (int ni = 0; ni < TN / 16; ++ni)
                B[ni] = *reinterpret_cast<const bf16x8*>(&s_w[(ni * 16 + r16) * kWRow + tap * kCK + g * 8]);
            for (int mi = 0; mi < 2; ++mi)
                for (int ni = 0; ni < TN / 16; ++ni)
                    acc[mi][ni] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[mi], B[ni], acc[mi][ni], 0, 0, 0);
        }
        __syncthreads();
    }
    // epilogue: D[m = 4g + j][n = r16] -> pixel column 4g + j, channel ni*16 + r16
    for (int mi = 0; mi < 2; ++mi) {
        const int y = y0 + wave * 2 + mi;
        if (y >= a.H) continue;
        for (int ni = 0; ni < TN / 16; ++ni) {
            const int co = co0 + ni * 16 + r16;
            const float b = a.bias ? a.bias[co] : 0.0f;
            bf16* out = co < a.Cs ? a.out0 : a.out1;
            const int oc = co < a.Cs ? co : co - a.Cs;
            const int Co = co < a.Cs ? a.Cs : a.Cout - a.Cs;
            for (int j = 0; j < 4; ++j) {
                const int x = x0 + 4 * g + j;
                if (x >= a.W) continue;
                float v = acc[mi][ni][j] + b;
                if (a.relu) v = fmaxf(v, 0.0f);
                out[(((int64_t)n * a.H + y) * a.W + x) * Co + oc] = (bf16)v;
            }
        }
    }
}

// 3x3 convolution on v_mfma_f32_32x32x16_bf16. A block computes (4*RW rows x 32 columns) pixels
// x TN channels; wave w owns rows [w*RW, (w+1)*RW), all TN channels. One 32x32x16 MFMA does
// twice the multiply-adds of a 16x16x32 one for the same two 16-byte operand fragments, so per
// multiply-add the waves read half as many bytes from LDS, whose bandwidth bounds k_conv.
// Operand layouts (checked on the device): A[m][k] in lane m + 32*(k/8), element k%8, i.e. the
// pixel m of the row and 8 channels of the 16-channel half; B[k][n] in lane n + 32*(k/8); D[m][n]
// in lane n + 32*((m/4)%2), element 4*(m/8) + m%4.
constexpr int kTW32 = 32, kHW32 = kTW32 + 2;

// Denoiser input transform: radiance -> log1p(max(radiance, 0)), NaN kept (torch's clamp_min
// and log1p order), and its inverse on the output, max(expm1(v), 0).
__device__ __forceinline__ float to_log(float v) { return log1pf(v < 0.f ? 0.f : v); }
__device__ __forceinline__ float from_log(float v) {
    const float r = expm1f(v);
    return r < 0.f ? 0.f : r;
}

// Epilogue of the 32x32-MFMA convolutions: bias, ReLU, BF16 output tile of (4*RW rows x 32
// columns) x TN channels through LDS, optional 2x2 max-pool of it. Element j of lane l is pixel
// column 8*(j/4) + 4*(l/32) + j%4, channel l%32; the tile is written to s_out as [row][column]
// [channel] and then stored 8 channels (16 bytes) per thread. The caller's last barrier must
// have freed s_out (TH * 32 * TN BF16).
template <int TN, int RW>
__device__ __forceinline__ void conv32_epilogue(const ConvArgs& a, const f32x16 (&acc)[RW][TN / 32], bf16* s_out,
                                                int n, int y0, int x0, int co0) {
    constexpr int NG = TN / 32, TH = RW * (kBlock / 64);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int l32 = lane & 31, kg = lane >> 5;
    for (int mi = 0; mi < RW; ++mi)
        for (int ng = 0; ng < NG; ++ng) {
            const int c = ng * 32 + l32;
            const float b = a.bias ? a.bias[co0 + c] : 0.0f;
            for (int j = 0; j < 16; ++j) {
                float v = acc[mi][ng][j] + b;
                if (a.relu) v = fmaxf(v, 0.0f);
                s_out[((wave * RW + mi) * kTW32 + 8 * (j >> 2) + 4 * kg + (j & 3)) * TN + c] = (bf16)v;
            }
        }
    // the tail's 3 x 32 weights and 3 biases go to LDS first: read per pixel from memory, the
    // compiler reloads them after every output store (the pointers may alias)
    __shared__ float s_tail[TN == 32 ? 3 * 32 + 3 : 1];
    if (TN == 32 && a.tail_out && tid < 3 * 32 + 3) s_tail[tid] = tid < 3 * 32 ? a.tail_w[tid] : a.tail_b[tid - 3 * 32];
    __syncthreads();
    if (TN == 32 && a.tail_out) {
        // the tail on the BF16 features of each pixel, in k_tail_fwd's order of operations
        for (int px = tid; px < TH * kTW32; px += kBlock) {
            const int y = y0 + px / kTW32, x = x0 + px % kTW32;
            if (y >= a.H || x >= a.W) continue;
            float f[32];
            for (int q = 0; q < 4; ++q) unpack8(*reinterpret_cast<const uint4*>(&s_out[px * TN + q * 8]), f + q * 8);
            const int64_t p = ((int64_t)n * a.H + y) * a.W + x;
            for (int o = 0; o < 3; ++o) {
                float acc_t = s_tail[3 * 32 + o];
                for (int c = 0; c < 32; ++c) acc_t += s_tail[o * 32 + c] * f[c];
                a.tail_out[p * 3 + o] = from_log(to_log(a.tail_x[p * 3 + o]) + acc_t);
            }
        }
        return;
    }
    constexpr int kVecs = TH * kTW32 * TN / 8;
    for (int t = tid; t < kVecs; t += kBlock) {
        const int px = t / (TN / 8), cg = t % (TN / 8);
        const int y = y0 + px / kTW32, x = x0 + px % kTW32;
        if (y >= a.H || x >= a.W) continue;
        const int co = co0 + cg * 8;  // 8 channels never straddle Cs (a multiple of 32)
        bf16* out = co < a.Cs ? a.out0 : a.out1;
        const int oc = co < a.Cs ? co : co - a.Cs;
        const int Co = co < a.Cs ? a.Cs : a.Cout - a.Cs;
        *reinterpret_cast<uint4*>(out + (((int64_t)n * a.H + y) * a.W + x) * Co + oc) =
            *reinterpret_cast<const uint4*>(&s_out[px * TN + cg * 8]);
    }
    if (a.pool_out) {
        // MaxPool2d(2) of the tile, as k_pool_fwd computes it (first maximum in row-major
        // order, NaN propagates); H and W are even and tiles start at even rows and columns,
        // so every window lies inside one tile
        const int Hp = a.H >> 1, Wp = a.W >> 1;
        constexpr int kPool = (TH / 2) * (kTW32 / 2) * (TN / 8);
        for (int t = tid; t < kPool; t += kBlock) {
            const int pp = t / (TN / 8), cg = t % (TN / 8);
            const int ry = pp / (kTW32 / 2), rx = pp % (kTW32 / 2);
            const int py = (y0 >> 1) + ry, pxo = (x0 >> 1) + rx;
            if (py >= Hp || pxo >= Wp) continue;
            float m[8], v[8];
            unpack8(*reinterpret_cast<const uint4*>(&s_out[((2 * ry) * kTW32 + 2 * rx) * TN + cg * 8]), m);
            const int offs[3] = {1, kTW32, kTW32 + 1};
            for (int k = 0; k < 3; ++k) {
                unpack8(*reinterpret_cast<const uint4*>(&s_out[((2 * ry) * kTW32 + 2 * rx + offs[k]) * TN + cg * 8]),
                        v);
                for (int e = 0; e < 8; ++e)
                    if (v[e] > m[e] || v[e] != v[e]) m[e] = v[e];
            }
            *reinterpret_cast<uint4*>(a.pool_out + (((int64_t)n * Hp + py) * Wp + pxo) * a.Cout + co0 + cg * 8) =
                pack8(m);
        }
    }
}

template <int TN, int RW, bool MASK>
__global__ __launch_bounds__(kBlock) void k_conv32(ConvArgs a) {
    constexpr int NG = TN / 32;
    constexpr int TH = RW * (kBlock / 64), HH = TH + 2;
    constexpr int kHalo = HH * kHW32 * 4;   // 16-byte loads of one chunk's halo
    constexpr int kWLd = TN * 9 * 4;        // 16-byte loads of one chunk's weights
    constexpr int kHIt = (kHalo + kBlock - 1) / kBlock, kWIt = (kWLd + kBlock - 1) / kBlock;
    __shared__ __attribute__((aligned(16))) bf16 s_in[HH * kHW32 * kCK];
    __shared__ __attribute__((aligned(16))) bf16 s_w[TN * kWRow];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int tiles_x = (a.W + kTW32 - 1) / kTW32, tiles_y = (a.H + TH - 1) / TH;
    const int tile = blockIdx.x;
    const int n = tile / (tiles_x * tiles_y);
    const int ty = (tile / tiles_x) % tiles_y, tx = tile % tiles_x;
    const int y0 = ty * TH, x0 = tx * kTW32;
    const int co0 = blockIdx.y * TN;
    const int Cin = a.C0 + a.C1;
    f32x16 acc[RW][NG];
    for (int mi = 0; mi < RW; ++mi)
        for (int ng = 0; ng < NG; ++ng)
            for (int j = 0; j < 16; ++j) acc[mi][ng][j] = 0.f;
    const int l32 = lane & 31, kg = lane >> 5;
    // same software pipeline as k_conv: chunk c + 1 is loaded into registers during chunk c
    uint4 hv[kHIt], hm[MASK ? kHIt : 1], wv[kWIt];
    bool masked = false;
    auto load_chunk = [&](int ci0) {
        const bool first = ci0 < a.C0;
        const bf16* src = first ? a.src0 : a.src1;
        const bf16* msk = first ? a.mask0 : a.mask1;
        const int Cs = first ? a.C0 : a.C1;
        const int cc0 = first ? ci0 : ci0 - a.C0;
        masked = MASK && msk != nullptr;
#pragma unroll
        for (int i = 0; i < kHIt; ++i) {
            const int t = tid + i * kBlock;
            hv[i] = zero16();
            if (MASK) hm[i] = zero16();
            if (t < kHalo) {
                const int px = t >> 2, part = t & 3;
                const int hy = px / kHW32, hx = px % kHW32;
                const int gy = y0 + hy - 1, gx = x0 + hx - 1;
                if (gy >= 0 && gy < a.H && gx >= 0 && gx < a.W) {
                    const int64_t off = (((int64_t)n * a.H + gy) * a.W + gx) * Cs + cc0 + part * 8;
                    hv[i] = *reinterpret_cast<const uint4*>(src + off);
                    if (MASK && msk) hm[i] = *reinterpret_cast<const uint4*>(msk + off);
                }
            }
        }
#pragma unroll
        for (int i = 0; i < kWIt; ++i) {
            const int t = tid + i * kBlock;
            wv[i] = zero16();
            if (t < kWLd) {
                const int co = t / 36, r = t % 36, tap = r >> 2, part = r & 3;
                wv[i] = *reinterpret_cast<const uint4*>(a.wk + ((int64_t)(co0 + co) * 9 + tap) * Cin + ci0 + part * 8);
            }
        }
    };
    load_chunk(0);
    for (int ci0 = 0; ci0 < Cin; ci0 += kCK) {
#pragma unroll
        for (int i = 0; i < kHIt; ++i) {
            const int t = tid + i * kBlock;
            if (t < kHalo) {
                uint4 v = hv[i];
                if (MASK && masked) v = relu_mask(v, hm[MASK ? i : 0]);
                *reinterpret_cast<uint4*>(&s_in[(t >> 2) * kCK + (t & 3) * 8]) = v;
            }
        }
#pragma unroll
        for (int i = 0; i < kWIt; ++i) {
            const int t = tid + i * kBlock;
            if (t < kWLd) {
                const int co = t / 36, r = t % 36, tap = r >> 2, part = r & 3;
                *reinterpret_cast<uint4*>(&s_w[co * kWRow + tap * kCK + part * 8]) = wv[i];
            }
        }
        __syncthreads();
        if (ci0 + kCK < Cin) load_chunk(ci0 + kCK);
        for (int tap = 0; tap < 9; ++tap) {
            const int ky = tap / 3, kx = tap % 3;
#pragma unroll
            for (int kh = 0; kh < 2; ++kh) {
                bf16x8 A[RW], B[NG];
                for (int mi = 0; mi < RW; ++mi) {
                    const int row = wave * RW + mi + ky;
                    A[mi] = *reinterpret_cast<const bf16x8*>(&s_in[(row * kHW32 + l32 + kx) * kCK + kh * 16 + kg * 8]);
                }
                for (int ng = 0; ng < NG; ++ng)
                    B[ng] = *reinterpret_cast<const bf16x8*>(&s_w[(ng * 32 + l32) * kWRow + tap * kCK + kh * 16 + kg * 8]);
                for (int mi = 0; mi < RW; ++mi)
                    for (int ng = 0; ng < NG; ++ng)
                        acc[mi][ng] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A[mi], B[ng], acc[mi][ng], 0, 0, 0);
            }
        }
        __syncthreads();
    }
    static_assert(TN != 64 || TH * kTW32 * TN <= TN * kWRow, "64-channel tile fits the weight slice");
    static_assert(TN != 32 || TH * kTW32 * TN <= HH * kHW32 * kCK, "32-channel tile fits the halo");
    conv32_epilogue<TN, RW>(a, acc, TN == 64 ? s_w : s_in, n, y0, x0, co0);
}

// 3x3 convolution on FP8 (E4M3) operands: v_mfma_scale_f32_32x32x64_f8f6f4 with unit block
// scales, FP32 accumulation. One MFMA multiplies 32 pixels x 64 input channels by 64 input
// channels x 32 output channels: twice the multiply-adds per operand byte of the BF16 32x32x16
// instruction and twice its rate. The BF16 input is quantized while it is staged into LDS,
// q = e4m3(clamp(x / act_scale, +-448)) rounded to nearest even (the converter turns values
// beyond 448 into NaN, hence the clamp); the weights arrive quantized per output channel,
// w ~= q_w * w_scale[co]; the epilogue multiplies the FP32 sums by act_scale * w_scale[co] before
// the bias, ReLU and BF16 store. Operand layout (checked on the device): A byte j of lane
// m + 32 h holds k = 32 h + j of pixel m; B the same for output channel n; D as in k_conv32.
// Tile and pipeline as k_conv32: (4*RW rows x 32 columns) x TN channels, 64 input channels per
// chunk, the next chunk's loads in registers during the current chunk's MFMAs.
constexpr int kCK8 = 64;                 // input channels per FP8 chunk
constexpr int kWRow8 = 9 * kCK8 + 16;    // LDS weight row (bytes), padded against bank conflicts
typedef int i32x8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ uint2 quant8_e4m3(const uint4& v, float inv_scale) {
    float f[8];
    unpack8(v, f);
    float c[8];
    for (int k = 0; k < 8; ++k) c[k] = fminf(fmaxf(f[k] * inv_scale, -448.f), 448.f);
    int w0 = __builtin_amdgcn_cvt_pk_fp8_f32(c[0], c[1], 0, false);   // bytes 0, 1
    w0 = __builtin_amdgcn_cvt_pk_fp8_f32(c[2], c[3], w0, true);        // bytes 2, 3
    int w1 = __builtin_amdgcn_cvt_pk_fp8_f32(c[4], c[5], 0, false);
    w1 = __builtin_amdgcn_cvt_pk_fp8_f32(c[6], c[7], w1, true);
    return make_uint2((unsigned)w0, (unsigned)w1);
}

template <int TN, int RW>
__global__ __launch_bounds__(kBlock) void k_conv_f8(ConvArgs a) {
    constexpr int NG = TN / 32;
    constexpr int TH = RW * (kBlock / 64), HH = TH + 2;
    constexpr int kHalo = HH * kHW32 * 8;   // 16-byte BF16 loads (8 channels) of one chunk's halo
    constexpr int kWLd = TN * 9 * 4;        // 16-byte FP8 loads (16 channels) of one chunk's weights
    constexpr int kHIt = (kHalo + kBlock - 1) / kBlock, kWIt = (kWLd + kBlock - 1) / kBlock;
    __shared__ __attribute__((aligned(16))) uint8_t s_in[HH * kHW32 * kCK8];
    __shared__ __attribute__((aligned(16))) uint8_t s_w[TN * kWRow8];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int tiles_x = (a.W + kTW32 - 1) / kTW32, tiles_y = (a.H + TH - 1) / TH;
    const int tile = blockIdx.x;
    const int n = tile / (tiles_x * tiles_y);
    const int ty = (tile / tiles_x) % tiles_y, tx = tile % tiles_x;
    const int y0 = ty * TH, x0 = tx * kTW32;
    const int co0 = blockIdx.y * TN;
    const int Cin = a.C0 + a.C1;
    const float inv_sa = 1.0f / a.act_scale;
    const uint8_t* wq = reinterpret_cast<const uint8_t*>(a.wk);
    f32x16 acc[RW][NG];
    for (int mi = 0; mi < RW; ++mi)
        for (int ng = 0; ng < NG; ++ng)
            for (int j = 0; j < 16; ++j) acc[mi][ng][j] = 0.f;
    const int l32 = lane & 31, kg = lane >> 5;
    uint4 hv[kHIt], wv[kWIt];
    auto load_chunk = [&](int ci0) {
        const bool first = ci0 < a.C0;
        const bf16* src = first ? a.src0 : a.src1;
        const int Cs = first ? a.C0 : a.C1;
        const int cc0 = first ? ci0 : ci0 - a.C0;
#pragma unroll
        for (int i = 0; i < kHIt; ++i) {
            const int t = tid + i * kBlock;
            hv[i] = zero16();
            if (t < kHalo) {
                const int px = t >> 3, part = t & 7;
                const int hy = px / kHW32, hx = px % kHW32;
                const int gy = y0 + hy - 1, gx = x0 + hx - 1;
                if (gy >= 0 && gy < a.H && gx >= 0 && gx < a.W)
                    hv[i] = *reinterpret_cast<const uint4*>(src + (((int64_t)n * a.H + gy) * a.W + gx) * Cs + cc0 +
                                                            part * 8);
            }
        }
#pragma unroll
        for (int i = 0; i < kWIt; ++i) {
            const int t = tid + i * kBlock;
            wv[i] = zero16();
            if (t < kWLd) {
                const int co = t / 36, r = t % 36, tap = r >> 2, part = r & 3;
                wv[i] = *reinterpret_cast<const uint4*>(wq + ((int64_t)(co0 + co) * 9 + tap) * Cin + ci0 + part * 16);
            }
        }
    };
    load_chunk(0);
    for (int ci0 = 0; ci0 < Cin; ci0 += kCK8) {
#pragma unroll
        for (int i = 0; i < kHIt; ++i) {
            const int t = tid + i * kBlock;
            if (t < kHalo) *reinterpret_cast<uint2*>(&s_in[(t >> 3) * kCK8 + (t & 7) * 8]) = quant8_e4m3(hv[i], inv_sa);
        }
#pragma unroll
        for (int i = 0; i < kWIt; ++i) {
            const int t = tid + i * kBlock;
            if (t < kWLd) {
                const int co = t / 36, r = t % 36, tap = r >> 2, part = r & 3;
                *reinterpret_cast<uint4*>(&s_w[co * kWRow8 + tap * kCK8 + part * 16]) = wv[i];
            }
        }
        __syncthreads();
        if (ci0 + kCK8 < Cin) load_chunk(ci0 + kCK8);
        for (int tap = 0; tap < 9; ++tap) {
            const int ky = tap / 3, kx = tap % 3;
            i32x8 A[RW], B[NG];
            for (int mi = 0; mi < RW; ++mi) {
                const int row = wave * RW + mi + ky;
                const uint4* q = reinterpret_cast<const uint4*>(&s_in[(row * kHW32 + l32 + kx) * kCK8 + kg * 32]);
                const uint4 lo = q[0], hi = q[1];
                A[mi] = i32x8{(int)lo.x, (int)lo.y, (int)lo.z, (int)lo.w, (int)hi.x, (int)hi.y, (int)hi.z, (int)hi.w};
            }
            for (int ng = 0; ng < NG; ++ng) {
                const uint4* q = reinterpret_cast<const uint4*>(&s_w[(ng * 32 + l32) * kWRow8 + tap * kCK8 + kg * 32]);
                const uint4 lo = q[0], hi = q[1];
                B[ng] = i32x8{(int)lo.x, (int)lo.y, (int)lo.z, (int)lo.w, (int)hi.x, (int)hi.y, (int)hi.z, (int)hi.w};
            }
            for (int mi = 0; mi < RW; ++mi)
                for (int ng = 0; ng < NG; ++ng)
                    acc[mi][ng] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(A[mi], B[ng], acc[mi][ng], 0, 0, 0,
                                                                                  127, 0, 127);
        }
        __syncthreads();
    }
    for (int ng = 0; ng < NG; ++ng) {
        const float sc = a.act_scale * a.w_scale[co0 + ng * 32 + l32];
        for (int mi = 0; mi < RW; ++mi)
            for (int j = 0; j < 16; ++j) acc[mi][ng][j] *= sc;
    }
    static_assert(TH * kTW32 * TN * 2 <= TN * kWRow8, "output tile fits the weight slice");
    conv32_epilogue<TN, RW>(a, acc, reinterpret_cast<bf16*>(s_w), n, y0, x0, co0);
}

// 3x3 convolution with the operands staged by direct global-to-LDS loads (buffer_load_dwordx4
// ... lds): no staging registers, two LDS buffers, so chunk c + 1 streams into one buffer while
// the MFMAs of chunk c read the other, and one barrier per 32-channel chunk. A load instruction
// writes lane i's 16 bytes at the wave's LDS base + 16 i (checked on the device), so an LDS
// "slot" (16 bytes) is filled by the lane whose global address the kernel picks for it; lanes
// outside the image or past the data get an offset beyond the buffer's range and write zeros
// (the padding of the convolution).
// Slot layouts, with the 4 16-byte parts of a pixel (32 channels) or of a weight row (co, tap)
// XOR-swizzled by bits 2..3 of the pixel / output channel, so that the 32-lane fragment reads
// (16 lanes per LDS cycle, consecutive pixels or channels) hit distinct banks:
//   halo   slot p*4 + (part ^ ((p >> 2) & 3)), pixel p = row * 34 + column
//   weight slot tap*TN*4 + co*4 + (part ^ ((co >> 2) & 3))
// Register blocking over kernel rows: for one kernel column kx and K half kh a wave reads the
// RW + 2 halo rows it needs once and 3 weight fragments per 32 channels, and issues 3*RW*NG
// MFMAs (0.5 fragment reads per MFMA at RW 4, TN 64, against 1 in k_conv32).
constexpr unsigned kDlOut = 0x80000000u;  // buffer offset past the range: the load returns 0
constexpr int kDlRsrcWord3 = 0x00020000;  // raw buffer, 32-bit data format

template <int TN, int RW>
struct DlShape {
    static constexpr int NG = TN / 32, TH = RW * (kBlock / 64), HH = TH + 2;
    static constexpr int kHalo = HH * kHW32 * 4;
    static constexpr int kHaloSlots = (kHalo + kBlock - 1) / kBlock * kBlock;
    static constexpr int kW = TN * 36;
    static constexpr int kWSlots = (kW + kBlock - 1) / kBlock * kBlock;
    static constexpr int kBuf = kHaloSlots + kWSlots;  // 16-byte slots per buffer
};

typedef __attribute__((address_space(3))) void lds_void_t;

template <int TN, int RW>
__global__ __launch_bounds__(kBlock, 1) void k_conv_dl(ConvArgs a) {
    using S = DlShape<TN, RW>;
    constexpr int NG = S::NG, TH = S::TH;
    static_assert(2 * S::kBuf * 16 <= 160 * 1024, "two buffers fit the 160 KB of LDS");
    static_assert(TH * kTW32 * TN * 2 <= S::kBuf * 16, "output tile fits one buffer");
    // Two buffers as two LDS objects, each access naming one of them: the compiler's wait
    // counting then sees that loads into one buffer cannot alias reads of the other and does not
    // wait for them before the MFMA loop's LDS reads.
    __shared__ __attribute__((aligned(16))) uint4 s_b0[S::kBuf];
    __shared__ __attribute__((aligned(16))) uint4 s_b1[S::kBuf];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int tiles_x = (a.W + kTW32 - 1) / kTW32, tiles_y = (a.H + TH - 1) / TH;
    const int tile = blockIdx.x;
    const int n = tile / (tiles_x * tiles_y);
    const int ty = (tile / tiles_x) % tiles_y, tx = tile % tiles_x;
    const int y0 = ty * TH, x0 = tx * kTW32;
    const int co0 = blockIdx.y * TN;
    const int Cin = a.C0 + a.C1;
    const int nchunks = Cin / kCK;
    // buffer resources: byte ranges of the two sources and of the weights (sizes < 2^31 checked
    // by the launcher)
    const __amdgpu_buffer_rsrc_t r0 = __builtin_amdgcn_make_buffer_rsrc(
        (void*)a.src0, 0, (int)((int64_t)a.N * a.H * a.W * a.C0 * 2), kDlRsrcWord3);
    const __amdgpu_buffer_rsrc_t r1 = __builtin_amdgcn_make_buffer_rsrc(
        (void*)(a.src1 ? a.src1 : a.src0), 0, (int)((int64_t)a.N * a.H * a.W * a.C1 * 2), kDlRsrcWord3);
    const __amdgpu_buffer_rsrc_t rw = __builtin_amdgcn_make_buffer_rsrc(
        (void*)a.wk, 0, (int)((int64_t)a.Cout * 9 * Cin * 2), kDlRsrcWord3);
    auto issue = [&](int c, uint4* base) {
        const int ci0 = c * kCK;
        const bool first = ci0 < a.C0;
        const unsigned Cs = first ? a.C0 : a.C1;
        const unsigned cc0 = first ? ci0 : ci0 - a.C0;
#pragma unroll
        for (int i = 0; i < S::kHaloSlots / kBlock; ++i) {
            const int t = i * kBlock + tid;
            unsigned off = kDlOut;
            if (t < S::kHalo) {
                const int p = t >> 2, part = (t & 3) ^ ((p >> 2) & 3);
                const int hy = p / kHW32, hx = p % kHW32;
                const int gy = y0 + hy - 1, gx = x0 + hx - 1;
                if (gy >= 0 && gy < a.H && gx >= 0 && gx < a.W)
                    off = ((((unsigned)n * a.H + gy) * a.W + gx) * Cs + cc0 + part * 8) * 2u;
            }
            lds_void_t* dst = (lds_void_t*)(base + i * kBlock + wave * 64);
            if (first) __builtin_amdgcn_raw_ptr_buffer_load_lds(r0, dst, 16, off, 0, 0, 0);
            else __builtin_amdgcn_raw_ptr_buffer_load_lds(r1, dst, 16, off, 0, 0, 0);
        }
#pragma unroll
        for (int i = 0; i < S::kWSlots / kBlock; ++i) {
            const int t = i * kBlock + tid;
            unsigned off = kDlOut;
            if (t < S::kW) {
                const int tap = t / (TN * 4), r = t % (TN * 4), co = r >> 2, part = (r & 3) ^ ((co >> 2) & 3);
                off = ((((unsigned)(co0 + co) * 9 + tap) * Cin) + ci0 + part * 8) * 2u;
            }
            lds_void_t* dst = (lds_void_t*)(base + S::kHaloSlots + i * kBlock + wave * 64);
            __builtin_amdgcn_raw_ptr_buffer_load_lds(rw, dst, 16, off, 0, 0, 0);
        }
    };
    f32x16 acc[RW][NG];
    for (int mi = 0; mi < RW; ++mi)
        for (int ng = 0; ng < NG; ++ng)
            for (int j = 0; j < 16; ++j) acc[mi][ng][j] = 0.f;
    const int l32 = lane & 31, kg = lane >> 5;
    auto compute = [&](const uint4* sh) {
        const uint4* sw = sh + S::kHaloSlots;
#pragma unroll
        for (int kx = 0; kx < 3; ++kx) {
#pragma unroll
            for (int kh = 0; kh < 2; ++kh) {
                const int k = kh * 2 + kg;
                bf16x8 A[RW + 2], B[3][NG];
#pragma unroll
                for (int r = 0; r < RW + 2; ++r) {
                    const int p = (wave * RW + r) * kHW32 + l32 + kx;
                    const uint4 v = sh[p * 4 + (k ^ ((p >> 2) & 3))];
                    A[r] = *reinterpret_cast<const bf16x8*>(&v);
                }
#pragma unroll
                for (int ky = 0; ky < 3; ++ky)
#pragma unroll
                    for (int ng = 0; ng < NG; ++ng) {
                        const int co = ng * 32 + l32;
                        const uint4 v = sw[(ky * 3 + kx) * TN * 4 + co * 4 + (k ^ ((co >> 2) & 3))];
                        B[ky][ng] = *reinterpret_cast<const bf16x8*>(&v);
                    }
#pragma unroll
                for (int ky = 0; ky < 3; ++ky)
#pragma unroll
                    for (int mi = 0; mi < RW; ++mi)
#pragma unroll
                        for (int ng = 0; ng < NG; ++ng)
                            acc[mi][ng] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A[mi + ky], B[ky][ng], acc[mi][ng],
                                                                                  0, 0, 0);
            }
        }
    };
    issue(0, s_b0);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    // chunk c reads buffer c % 2 while chunk c + 1 streams into the other; each step ends with
    // this thread's loads waited for and a barrier (every load landed, nobody reads the buffer
    // the next step refills)
    for (int c = 0; c < nchunks; c += 2) {
        if (c + 1 < nchunks) issue(c + 1, s_b1);
        compute(s_b0);
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        if (c + 1 >= nchunks) break;
        if (c + 2 < nchunks) issue(c + 2, s_b0);
        compute(s_b1);
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
    }
    conv32_epilogue<TN, RW>(a, acc, reinterpret_cast<bf16*>(s_b0), n, y0, x0, co0);
}

// Register-blocked 3x3 convolution. Each wave owns RW output rows x 16 columns x 16*NB output
// channels; waves are stacked along rows (4 per block). For one kernel column kx the wave
// keeps the 3 x NB weight fragments of that column in registers and walks the RW + 2 input
// rows of its halo once: every input-row fragment read from LDS feeds up to 3 kernel rows x
// NB output fragments (3*NB MFMAs), instead of the NB of the simple tiling above. The halo
// keeps 40 channels of pitch per pixel (80 B) and the weight rows 296, so the 16-byte reads
// of a 16-lane group fall on distinct banks.
constexpr int kCKP = kCK + 8;

template <int RW, int NB>
__global__ __launch_bounds__(kBlock, RW * NB >= 32 ? 1 : 2) void k_conv_rb(ConvArgs a) {
    constexpr int kWaves = kBlock / 64;
    constexpr int TH = RW * kWaves;     // output rows per block
    constexpr int TN = 16 * NB;         // output channels per block
    constexpr int HH = TH + 2, HWd = kTW + 2;
    __shared__ __attribute__((aligned(16))) bf16 s_in[HH * HWd * kCKP];
    __shared__ __attribute__((aligned(16))) bf16 s_w[TN * kWRow];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int tiles_x = (a.W + kTW - 1) / kTW, tiles_y = (a.H + TH - 1) / TH;
    const int tile = blockIdx.x;
    const int n = tile / (tiles_x * tiles_y);
    const int ty = (tile / tiles_x) % tiles_y, tx = tile % tiles_x;
    const int y0 = ty * TH, x0 = tx * kTW;
    const int co0 = blockIdx.y * TN;
    const int Cin = a.C0 + a.C1;
    const int g = lane >> 4, r16 = lane & 15;
    f32x4 acc[RW][NB];
    for (int r = 0; r < RW; ++r)
        for (int ni = 0; ni < NB; ++ni) acc[r][ni] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int ci0 = 0; ci0 < Cin; ci0 += kCK) {
        const bool first = ci0 < a.C0;
        const bf16* src = first ? a.src0 : a.src1;
        const bf16* msk = first ? a.mask0 : a.mask1;
        const int Cs = first ? a.C0 : a.C1;
        const int cc0 = first ? ci0 : ci0 - a.C0;
        for (int t = tid; t < HH * HWd * 4; t += kBlock) {
            const int px = t >> 2, part = t & 3;
            const int hy = px / HWd, hx = px % HWd;
            const int gy = y0 + hy - 1, gx = x0 + hx - 1;
            uint4 v = zero16();
            if (gy >= 0 && gy < a.H && gx >= 0 && gx < a.W) {
                const int64_t off = (((int64_t)n * a.H + gy) * a.W + gx) * Cs + cc0 + part * 8;
                v = *reinterpret_cast<const uint4*>(src + off);
                if (msk) v = relu_mask(v, *reinterpret_cast<const uint4*>(msk + off));
            }
            *reinterpret_cast<uint4*>(&s_in[px * kCKP + part * 8]) = v;
        }
        for (int t = tid; t < TN * 9 * 4; t += kBlock) {
            const int co = t / 36, r = t % 36, tap = r >> 2, part = r & 3;
            const uint4 v =
                *reinterpret_cast<const uint4*>(a.wk + ((int64_t)(co0 + co) * 9 + tap) * Cin + ci0 + part * 8);
            *reinterpret_cast<uint4*>(&s_w[co * kWRow + tap * kCK + part * 8]) = v;
        }
        __syncthreads();
#pragma unroll
        for (int kx = 0; kx < 3; ++kx) {
            bf16x8 B[3][NB];
#pragma unroll
            for (int ky = 0; ky < 3; ++ky)
#pragma unroll
                for (int ni = 0; ni < NB; ++ni)
                    B[ky][ni] = *reinterpret_cast<const bf16x8*>(
                        &s_w[(ni * 16 + r16) * kWRow + (ky * 3 + kx) * kCK + g * 8]);
#pragma unroll
            for (int ir = 0; ir < RW + 2; ++ir) {
                const bf16x8 A =
                    *reinterpret_cast<const bf16x8*>(&s_in[((wave * RW + ir) * HWd + r16 + kx) * kCKP + g * 8]);
#pragma unroll
                for (int ky = 0; ky < 3; ++ky) {
                    const int r = ir - ky;  // output row fed by this input row through kernel row ky
                    if (r < 0 || r >= RW) continue;
#pragma unroll
                    for (int ni = 0; ni < NB; ++ni)
                        acc[r][ni] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, B[ky][ni], acc[r][ni], 0, 0, 0);
                }
            }
        }
        __syncthreads();
    }
    for (int r = 0; r < RW; ++r) {
        const int y = y0 + wave * RW + r;
        if (y >= a.H) continue;
        for (int ni = 0; ni < NB; ++ni) {
            const int co = co0 + ni * 16 + r16;
            const float b = a.bias ? a.bias[co] : 0.0f;
            bf16* out = co < a.Cs ? a.out0 : a.out1;
            const int oc = co < a.Cs ? co : co - a.Cs;
            const int Co = co < a.Cs ? a.Cs : a.Cout - a.Cs;
            for (int j = 0; j < 4; ++j) {
                const int x = x0 + 4 * g + j;
                if (x >= a.W) continue;
                float v = acc[r][ni][j] + b;
                if (a.relu) v = fmaxf(v, 0.0f);
                out[(((int64_t)n * a.H + y) * a.W + x) * Co + oc] = (bf16)v;
            }
        }
    }
}

// First convolution straight from the FP32 3-channel input: the 3x3x3 patch of a pixel is
// one K = 27 (+5 zero) row, so a 16-pixel x 16-channel output fragment is a single MFMA
// instead of the 9 taps x 32 (mostly zero) channels of the packed-input path. Output tile
// 8 x 16 pixels x 32 channels; weights [32][9][cin_w] BF16 (channels 0..2 used).
// kLog: x holds radiance and the layer reads log1p(max(x, 0)).
// Persistent blocks (a few per CU) stage the weights once and walk the tiles; each thread's
// halo pixel of the next tile is loaded into registers before the current tile is computed,
// so the loads' latency overlaps the MFMAs and the stores.
__device__ __forceinline__ void first_halo_load(const float* __restrict__ x, int tile, int tiles_x, int tiles_y,
                                                int H, int W, float& v0, float& v1, float& v2) {
    const int n = tile / (tiles_x * tiles_y);
    const int ty = (tile / tiles_x) % tiles_y, tx = tile % tiles_x;
    const int t = threadIdx.x, hy = t / kHW, hx = t % kHW;
    const int gy = ty * kTH + hy - 1, gx = tx * kTW + hx - 1;
    v0 = 0.f; v1 = 0.f; v2 = 0.f;
    if (gy >= 0 && gy < H && gx >= 0 && gx < W) {
        const float* px = x + (((int64_t)n * H + gy) * W + gx) * 3;
        v0 = px[0]; v1 = px[1]; v2 = px[2];
    }
}

template <bool kLog>
__global__ __launch_bounds__(kBlock) void k_conv_first(const float* __restrict__ x, const bf16* __restrict__ wk,
                                                       int cin_w, const float* __restrict__ bias, int N, int H, int W,
                                                       int relu, bf16* __restrict__ out) {
    static_assert(kHH * kHW <= kBlock, "one halo pixel per thread");
    __shared__ bf16 s_in[kHH * kHW * 4];  // halo pixels x 4 (3 channels + pad)
    __shared__ __attribute__((aligned(16))) bf16 s_w[32 * 40];  // [co][k], k = tap * 3 + c, 40 = padded row
    __shared__ __attribute__((aligned(16))) bf16 s_out[kTH * kTW * 32];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int tiles_x = (W + kTW - 1) / kTW, tiles_y = (H + kTH - 1) / kTH, ntiles = tiles_x * tiles_y * N;
    const bool halo = tid < kHH * kHW;
    for (int t = tid; t < 32 * 32; t += kBlock) {
        const int co = t >> 5, k = t & 31;
        s_w[co * 40 + k] = k < 27 ? wk[((int64_t)co * 9 + k / 3) * cin_w + k % 3] : (bf16)0.f;
    }
    const int g = lane >> 4, r16 = lane & 15;
    float c0 = 0.f, c1 = 0.f, c2 = 0.f;
    if (halo && (int)blockIdx.x < ntiles) first_halo_load(x, blockIdx.x, tiles_x, tiles_y, H, W, c0, c1, c2);
    for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        __syncthreads();  // the previous tile's readers of s_in and s_out are done
        if (halo) {
            float v0 = c0, v1 = c1, v2 = c2;
            if (kLog) {
                v0 = to_log(v0); v1 = to_log(v1); v2 = to_log(v2);
            }
            s_in[tid * 4 + 0] = (bf16)v0; s_in[tid * 4 + 1] = (bf16)v1; s_in[tid * 4 + 2] = (bf16)v2;
            s_in[tid * 4 + 3] = (bf16)0.f;
            const int next = tile + (int)gridDim.x;
            if (next < ntiles) first_halo_load(x, next, tiles_x, tiles_y, H, W, c0, c1, c2);
        }
        __syncthreads();
        const int n = tile / (tiles_x * tiles_y);
        const int ty = (tile / tiles_x) % tiles_y, tx = tile % tiles_x;
        const int y0 = ty * kTH, x0 = tx * kTW;
        f32x4 acc[2][2];
        for (int mi = 0; mi < 2; ++mi) {
            bf16x8 A;
            for (int i = 0; i < 8; ++i) {
                const int k = 8 * g + i;
                bf16 v = (bf16)0.f;
                if (k < 27) {
                    const int tap = k / 3, c = k % 3, ky = tap / 3, kx = tap % 3;
                    v = s_in[((wave * 2 + mi + ky) * kHW + r16 + kx) * 4 + c];
                }
                A[i] = v;
            }
            for (int ni = 0; ni < 2; ++ni) {
                const bf16x8 B = *reinterpret_cast<const bf16x8*>(&s_w[(ni * 16 + r16) * 40 + g * 8]);
                acc[mi][ni] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, B, f32x4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
            }
        }
        // epilogue through LDS: the 8 x 16 x 32 tile as [row][column][channel], then 16-byte stores
        for (int mi = 0; mi < 2; ++mi)
            for (int ni = 0; ni < 2; ++ni) {
                const int co = ni * 16 + r16;
                const float b = bias ? bias[co] : 0.0f;
                for (int j = 0; j < 4; ++j) {
                    float v = acc[mi][ni][j] + b;
                    if (relu) v = fmaxf(v, 0.0f);
                    s_out[((wave * 2 + mi) * kTW + 4 * g + j) * 32 + co] = (bf16)v;
                }
            }
        __syncthreads();
        for (int t = tid; t < kTH * kTW * 4; t += kBlock) {
            const int px = t >> 2, cg = t & 3;
            const int y = y0 + px / kTW, xx = x0 + px % kTW;
            if (y >= H || xx >= W) continue;
            *reinterpret_cast<uint4*>(out + (((int64_t)n * H + y) * W + xx) * 32 + cg * 8) =
                *reinterpret_cast<const uint4*>(&s_out[px * 32 + cg * 8]);
        }
    }
}

// Wt[ci][tap][co] = W[co][taps-1-tap][ci]: weights of the data-gradient convolution.
__global__ __launch_bounds__(kBlock) void k_flip_weights(const bf16* w, int Cout, int Cin, int taps, bf16* wt) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int64_t total = (int64_t)Cout * taps * Cin;
    if (i >= total) return;
    const int ci = (int)(i % Cin);
    const int tap = (int)((i / Cin) % taps);
    const int co = (int)(i / ((int64_t)Cin * taps));
    wt[((int64_t)ci * taps + (taps - 1 - tap)) * Cout + co] = w[i];
}

// ------------------------------------------------------------------ weight gradient
// dW[co][tap][ci] = sum over pixels of dYm[p][co] * X[p + tap offset][ci]. One block: 32
// output channels x 32 input channels x all 9 taps (three waves, one kernel row each), over a
// contiguous range of 8x16 pixel tiles; writes its FP32 partial sum (reduced afterwards).
constexpr int kWgBlock = 192;
__global__ __launch_bounds__(kWgBlock) void k_conv_wgrad(const bf16* dy, const bf16* ymask, int Cout,
                                                         const bf16* src0, int C0, const bf16* src1, int C1, int N,
                                                         int H, int W, int tiles_per_split, float* partial) {
    // LDS: dY tile transposed [32 co][128 px]; X halo transposed and shifted by kx:
    // [3 kx][32 ci][10 rows][16 cols] so that every MFMA operand read is 16-byte aligned.
    __shared__ __attribute__((aligned(16))) bf16 s_dy[32 * 128];
    __shared__ __attribute__((aligned(16))) bf16 s_x[3 * 32 * kHH * kTW];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;  // wave = ky
    const int g = lane >> 4, r16 = lane & 15;
    const int co0 = blockIdx.x * 32, ci0 = blockIdx.y * 32, split = blockIdx.z;
    const int Cin = C0 + C1;
    const bool first = ci0 < C0;
    const bf16* src = first ? src0 : src1;
    const int Cs = first ? C0 : C1;
    const int cc0 = first ? ci0 : ci0 - C0;
    const int tiles_x = (W + kTW - 1) / kTW, tiles_y = (H + kTH - 1) / kTH;
    const int ntiles = tiles_x * tiles_y * N;
    f32x4 acc[3][2][2];  // [kx][co subtile][ci subtile]
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 2; ++b)
            for (int c = 0; c < 2; ++c) acc[a][b][c] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int t_begin = split * tiles_per_split;
    const int t_end = min(ntiles, t_begin + tiles_per_split);
    for (int tile = t_begin; tile < t_end; ++tile) {
        const int n = tile / (tiles_x * tiles_y);
        const int ty = (tile / tiles_x) % tiles_y, tx = tile % tiles_x;
        const int y0 = ty * kTH, x0 = tx * kTW;
        // dY (masked by ReLU) -> s_dy[co][px]
        for (int t = tid; t < 128 * 4; t += kWgBlock) {
            const int px = t >> 2, part = t & 3;
            const int y = y0 + (px >> 4), x = x0 + (px & 15);
            uint4 v = zero16();
            if (y < H && x < W) {
                const int64_t off = (((int64_t)n * H + y) * W + x) * Cout + co0 + part * 8;
                v = *reinterpret_cast<const uint4*>(dy + off);
                if (ymask) v = relu_mask(v, *reinterpret_cast<const uint4*>(ymask + off));
            }
            const bf16x8 e = *reinterpret_cast<const bf16x8*>(&v);
            for (int i = 0; i < 8; ++i) s_dy[(part * 8 + i) * 128 + px] = e[i];
        }
        // X halo -> s_x[kx][ci][row][col] = X[y0 + row - 1][x0 + col + kx - 1][ci]
        for (int t = tid; t < kHH * kHW * 4; t += kWgBlock) {
            const int px = t >> 2, part = t & 3;
            const int hy = px / kHW, hx = px % kHW;
            const int gy = y0 + hy - 1, gx = x0 + hx - 1;
            uint4 v = zero16();
            if (gy >= 0 && gy < H && gx >= 0 && gx < W)
                v = *reinterpret_cast<const uint4*>(src + (((int64_t)n * H + gy) * W + gx) * Cs + cc0 + part * 8);
            const bf16x8 e = *reinterpret_cast<const bf16x8*>(&v);
            for (int kx = 0; kx < 3; ++kx) {
                const int col = hx - kx;
                if (col < 0 || col >= kTW) continue;
                for (int i = 0; i < 8; ++i) s_x[((kx * 32 + part * 8 + i) * kHH + hy) * kTW + col] = e[i];
            }
        }
        __syncthreads();
        // K = 128 pixels in 4 steps of 32 (two 16-pixel rows per step)
        for (int ks = 0; ks < 4; ++ks) {
            const int prow = ks * 2 + (g >> 1), pcol = (g & 1) * 8;
            bf16x8 A[2];
            for (int mi = 0; mi < 2; ++mi)
                A[mi] = *reinterpret_cast<const bf16x8*>(&s_dy[(mi * 16 + r16) * 128 + prow * 16 + pcol]);
            for (int kx = 0; kx < 3; ++kx)
                for (int ni = 0; ni < 2; ++ni) {
                    const bf16x8 B = *reinterpret_cast<const bf16x8*>(
                        &s_x[((kx * 32 + ni * 16 + r16) * kHH + prow + wave) * kTW + pcol]);
                    for (int mi = 0; mi < 2; ++mi)
                        acc[kx][mi][ni] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[mi], B, acc[kx][mi][ni], 0, 0, 0);
                }
        }
        __syncthreads();
    }
    // partial[split][co][tap][ci]; D[m = 4g + j][n = r16]: co = m, ci = n
    for (int kx = 0; kx < 3; ++kx) {
        const int tap = wave * 3 + kx;
        for (int mi = 0; mi < 2; ++mi)
            for (int ni = 0; ni < 2; ++ni)
                for (int j = 0; j < 4; ++j) {
                    const int co = co0 + mi * 16 + 4 * g + j, ci = ci0 + ni * 16 + r16;
                    partial[(((int64_t)split * Cout + co) * 9 + tap) * Cin + ci] = acc[kx][mi][ni][j];
                }
    }
}

// Column sums of rows of `size` floats: block (x, c) adds rows [c * per_chunk, (c + 1) *
// per_chunk) of `rows` (row k at partial + k * row_stride * size) over 64 columns and writes
// out + c * out_stride. The four waves take every fourth row, each lane into 8 running sums
// (8 loads in flight) added in order at the end, and the wave sums are added in wave order:
// every sum has a fixed order, so results are deterministic.
__global__ __launch_bounds__(kBlock) void k_reduce_rows(const float* partial, int64_t size, int rows,
                                                        int64_t row_stride, int per_chunk, float* out,
                                                        int64_t out_stride, int accumulate) {
    __shared__ float s[kBlock];
    constexpr int kWaves = kBlock / 64, kIlp = 8;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t i = (int64_t)blockIdx.x * 64 + lane;
    const int r0 = blockIdx.y * per_chunk, r1 = min(rows, r0 + per_chunk);
    float sum = 0.f;
    if (i < size) {
        float acc[kIlp] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        for (int k0 = r0 + wave; k0 < r1; k0 += kWaves * kIlp)
#pragma unroll
            for (int u = 0; u < kIlp; ++u) {
                const int k = k0 + u * kWaves;
                if (k < r1) acc[u] += partial[(int64_t)k * row_stride * size + i];
            }
        for (int u = 0; u < kIlp; ++u) sum += acc[u];
    }
    s[threadIdx.x] = sum;
    __syncthreads();  // every row of the chunk is read before its first row is overwritten
    if (wave == 0 && i < size) {
        const float t = ((s[lane] + s[lane + 64]) + s[lane + 128]) + s[lane + 192];
        float* o = out + blockIdx.y * out_stride;
        o[i] = accumulate ? o[i] + t : t;
    }
}

// out[i] (+)= sum over s of partial[s][i], in a fixed order. Up to 64 splits in one pass;
// more are summed in chunks of 64 first (in place, into each chunk's first row) so that the
// reads of a small weight tensor with many splits spread over many blocks.
void reduce_splits(hipStream_t st, float* partial, int64_t size, int splits, float* out, int accumulate) {
    constexpr int kChunk = 64;
    const int gx = cdiv(size, 64);
    if (splits <= kChunk) {
        hipLaunchKernelGGL(k_reduce_rows, dim3(gx, 1), dim3(kBlock), 0, st, partial, size, splits, (int64_t)1, splits,
                           out, (int64_t)0, accumulate);
        return;
    }
    const int chunks = cdiv(splits, kChunk);
    hipLaunchKernelGGL(k_reduce_rows, dim3(gx, chunks), dim3(kBlock), 0, st, partial, size, splits, (int64_t)1, kChunk,
                       partial, (int64_t)kChunk * size, 0);
    hipLaunchKernelGGL(k_reduce_rows, dim3(gx, 1), dim3(kBlock), 0, st, partial, size, chunks, (int64_t)kChunk, chunks,
                       out, (int64_t)0, accumulate);
}

// partial[split][c] = sum over the split's pixels of dYm[p][c] (the bias gradient before the
// split sum). A thread reads 8 channels (16 bytes) of a pixel per step; the block's threads
// cover kBlock / (C/8) pixels side by side, and their sums are added per channel in LDS in a
// fixed order. C is a multiple of 8, at most 8 * kBlock.
__global__ __launch_bounds__(kBlock) void k_bias_grad(const bf16* dy, const bf16* ymask, int64_t npx, int C,
                                                      int64_t px_per_split, float* partial) {
    __shared__ float s[kBlock * 8];
    const int c8 = C / 8, rows = kBlock / c8, t = threadIdx.x, g = t % c8, r = t / c8;
    float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    if (r < rows) {
        const int64_t p0 = (int64_t)blockIdx.x * px_per_split, p1 = min(npx, p0 + px_per_split);
        for (int64_t p = p0 + r; p < p1; p += rows) {
            const uint4 v = *reinterpret_cast<const uint4*>(dy + p * C + g * 8);
            const bf16x8 e = *reinterpret_cast<const bf16x8*>(&v);
            if (ymask) {
                const uint4 mv = *reinterpret_cast<const uint4*>(ymask + p * C + g * 8);
                const bf16x8 m = *reinterpret_cast<const bf16x8*>(&mv);
                for (int j = 0; j < 8; ++j) acc[j] += (float)m[j] > 0.0f ? (float)e[j] : 0.0f;
            } else {
                for (int j = 0; j < 8; ++j) acc[j] += (float)e[j];
            }
        }
        for (int j = 0; j < 8; ++j) s[r * C + g * 8 + j] = acc[j];
    }
    __syncthreads();
    for (int c = t; c < C; c += kBlock) {
        float sum = 0.f;
        for (int rr = 0; rr < rows; ++rr) sum += s[rr * C + c];
        partial[(int64_t)blockIdx.x * C + c] = sum;
    }
}

// ------------------------------------------------------------------ pool / upsample
// Elementwise layers move 8 channels (16 bytes) per thread; C is a multiple of 8.

// Thread mapping of the per-pixel NHWC kernels below: blockIdx.y = n * rows + row, and the
// threads along x cover the (column, 8-channel group) pairs of that row. Index math stays in
// 32 bits, with one division per thread (64-bit division is a long instruction sequence).
__device__ __forceinline__ bool row_thread(int rows, int cols, int C8, int& n, int& y, int& x, int& c) {
    const int j = blockIdx.x * kBlock + threadIdx.x;
    if (j >= cols * C8) return false;
    n = (int)blockIdx.y / rows;
    y = (int)blockIdx.y - n * rows;
    x = j / C8;
    c = (j - x * C8) * 8;
    return true;
}

inline dim3 row_grid(int n, int rows, int cols, int C8) { return dim3(cdiv((int64_t)cols * C8, kBlock), n * rows); }
// The row-mapped kernels need n * rows <= 65535 (grid y) and cols * C/8 < 2^31.
static bool row_grid_fits(int64_t n, int64_t rows, int64_t cols, int64_t c8) {
    return n > 0 && rows > 0 && cols > 0 && n * rows <= 65535 && cols * c8 < (int64_t(1) << 31);
}

// MaxPool2d(2): out[n,y,x,c] = max over the 2x2 window (first maximum in row-major order).
__global__ __launch_bounds__(kBlock) void k_pool_fwd(const bf16* in, int N, int H, int W, int C, bf16* out) {
    const int Ho = H / 2, Wo = W / 2, C8 = C / 8;
    int n, y, x, c;
    if (!row_thread(Ho, Wo, C8, n, y, x, c)) return;
    const int64_t p = ((int64_t)n * Ho + y) * Wo + x;
    const int64_t base = (((int64_t)n * H + 2 * y) * W + 2 * x) * C + c;
    const int64_t o[4] = {0, C, (int64_t)W * C, (int64_t)W * C + C};
    float m[8], v[8];
    unpack8(*reinterpret_cast<const uint4*>(in + base), m);
    for (int k = 1; k < 4; ++k) {
        unpack8(*reinterpret_cast<const uint4*>(in + base + o[k]), v);
        for (int e = 0; e < 8; ++e)
            if (v[e] > m[e] || v[e] != v[e]) m[e] = v[e];
    }
    *reinterpret_cast<uint4*>(out + p * C + c) = pack8(m);
}

// Gradient to the first maximum of each window (the pooled input is re-read, not stored).
__global__ __launch_bounds__(kBlock) void k_pool_bwd(const bf16* in, const bf16* dout, int N, int H, int W, int C,
                                                     bf16* din) {
    const int Ho = H / 2, Wo = W / 2, C8 = C / 8;
    int n, y, x, c;
    if (!row_thread(Ho, Wo, C8, n, y, x, c)) return;
    const int64_t p = ((int64_t)n * Ho + y) * Wo + x;
    const int64_t o[4] = {0, C, (int64_t)W * C, (int64_t)W * C + C};
    const int64_t base = (((int64_t)n * H + 2 * y) * W + 2 * x) * C + c;
    float m[8], v[8];
    int arg[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    unpack8(*reinterpret_cast<const uint4*>(in + base), m);
    for (int k = 1; k < 4; ++k) {
        unpack8(*reinterpret_cast<const uint4*>(in + base + o[k]), v);
        for (int e = 0; e < 8; ++e)
            if (v[e] > m[e] || v[e] != v[e]) { m[e] = v[e]; arg[e] = k; }
    }
    float g[8], z[8];
    unpack8(*reinterpret_cast<const uint4*>(dout + p * C + c), g);
    for (int k = 0; k < 4; ++k) {
        for (int e = 0; e < 8; ++e) z[e] = arg[e] == k ? g[e] : 0.0f;
        *reinterpret_cast<uint4*>(din + base + o[k]) = pack8(z);
    }
}

// Bilinear 2x (align_corners = False): source coordinate (d + 0.5) / 2 - 0.5 clamped at 0.
__device__ __forceinline__ void lin_src(int d, int in, int& i0, int& i1, float& l1) {
    float s = (d + 0.5f) * 0.5f - 0.5f;
    if (s < 0.f) s = 0.f;
    i0 = (int)s;
    i1 = min(i0 + 1, in - 1);
    l1 = s - (float)i0;
}

// Output (y, x) = the bilinear blend of the input pixels lin_src picks, in the expression order
// below; e8 = 8 channels at column offset c.
__device__ __forceinline__ void up_blend(const float a[8], const float b[8], const float cc[8], const float d[8],
                                         float ly, float lx, float v[8]) {
    for (int e = 0; e < 8; ++e)
        v[e] = (1.f - ly) * ((1.f - lx) * a[e] + lx * b[e]) + ly * ((1.f - lx) * cc[e] + lx * d[e]);
}

// One thread per input pixel (i, j) and 8 channels writes the 2 x 2 output pixels (2i + a,
// 2j + b). Away from the border their sources are the 3 x 3 input neighbourhood of (i, j):
// output row 2i reads rows i-1 and i, row 2i + 1 rows i and i + 1 (lin_src), so the 9 pixels
// are loaded once for the 4 outputs (instead of 4 loads per output). Border threads evaluate
// each output from lin_src's indices. Same arithmetic either way.
__global__ __launch_bounds__(kBlock) void k_up_fwd(const bf16* in, int N, int H, int W, int C, bf16* out) {
    const int Ho = 2 * H, Wo = 2 * W, C8 = C / 8;
    int n, i, j, c;
    if (!row_thread(H, W, C8, n, i, j, c)) return;
    auto px = [&](int yy, int xx) {
        return *reinterpret_cast<const uint4*>(in + (((int64_t)n * H + yy) * W + xx) * C + c);
    };
    if (i >= 1 && i + 1 < H && j >= 1 && j + 1 < W) {
        uint4 q[3][3];
#pragma unroll
        for (int ry = 0; ry < 3; ++ry)
#pragma unroll
            for (int rx = 0; rx < 3; ++rx) q[ry][rx] = px(i - 1 + ry, j - 1 + rx);
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            int y0, y1, x0, x1;
            float ly, lx, v[8];
            lin_src(2 * i + a, H, y0, y1, ly);  // y0 = i - 1 + a, y1 = i + a
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                lin_src(2 * j + b, W, x0, x1, lx);  // x0 = j - 1 + b, x1 = j + b
                float fa[8], fb[8], fc[8], fd[8];
                unpack8(q[a][b], fa); unpack8(q[a][b + 1], fb); unpack8(q[a + 1][b], fc); unpack8(q[a + 1][b + 1], fd);
                up_blend(fa, fb, fc, fd, ly, lx, v);
                *reinterpret_cast<uint4*>(out + (((int64_t)n * Ho + 2 * i + a) * Wo + 2 * j + b) * C + c) = pack8(v);
            }
        }
        return;
    }
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) {
            const int y = 2 * i + a, x = 2 * j + b;
            int y0, y1, x0, x1;
            float ly, lx, fa[8], fb[8], fc[8], fd[8], v[8];
            lin_src(y, H, y0, y1, ly);
            lin_src(x, W, x0, x1, lx);
            unpack8(px(y0, x0), fa); unpack8(px(y0, x1), fb); unpack8(px(y1, x0), fc); unpack8(px(y1, x1), fd);
            up_blend(fa, fb, fc, fd, ly, lx, v);
            *reinterpret_cast<uint4*>(out + (((int64_t)n * Ho + y) * Wo + x) * C + c) = pack8(v);
        }
}

// Adjoint of k_up_fwd as a gather: each input pixel sums the weighted gradients of the
// (at most 4 x 4) output pixels that read it.
__global__ __launch_bounds__(kBlock) void k_up_bwd(const bf16* dout, int N, int H, int W, int C, bf16* din) {
    const int C8 = C / 8;
    int n, y, x, c;
    if (!row_thread(H, W, C8, n, y, x, c)) return;
    const int64_t p = ((int64_t)n * H + y) * W + x;
    const int Ho = 2 * H, Wo = 2 * W;
    float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f}, g[8];
    for (int oy = max(0, 2 * y - 2); oy <= min(Ho - 1, 2 * y + 2); ++oy) {
        int y0, y1;
        float ly;
        lin_src(oy, H, y0, y1, ly);
        float wy = (y0 == y ? 1.f - ly : 0.f) + (y1 == y ? ly : 0.f);
        if (wy == 0.f) continue;
        for (int ox = max(0, 2 * x - 2); ox <= min(Wo - 1, 2 * x + 2); ++ox) {
            int x0, x1;
            float lx;
            lin_src(ox, W, x0, x1, lx);
            float wx = (x0 == x ? 1.f - lx : 0.f) + (x1 == x ? lx : 0.f);
            if (wx == 0.f) continue;
            unpack8(*reinterpret_cast<const uint4*>(dout + (((int64_t)n * Ho + oy) * Wo + ox) * C + c), g);
            const float w = wy * wx;
            for (int e = 0; e < 8; ++e) acc[e] += w * g[e];
        }
    }
    *reinterpret_cast<uint4*>(din + p * C + c) = pack8(acc);
}

// ------------------------------------------------------------------ input / tail
// log1p radiance (FP32, NHWC with 3 channels) -> BF16 NHWC padded to 32 channels; one pixel
// (64 bytes out) per thread.
__global__ __launch_bounds__(kBlock) void k_pack_input(const float* x, int64_t npx, bf16* out) {
    const int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (p >= npx) return;
    float f[8] = {x[p * 3], x[p * 3 + 1], x[p * 3 + 2], 0.f, 0.f, 0.f, 0.f, 0.f};
    uint4* o = reinterpret_cast<uint4*>(out + p * 32);
    o[0] = pack8(f);
    o[1] = o[2] = o[3] = zero16();
}

// out = x + W_t * feat + b_t (1x1 conv, 32 -> 3, FP32 math).
// kRadiance: x holds radiance; the residual adds to log1p(max(x, 0)) and out is max(expm1(.), 0).
template <bool kRadiance>
__global__ __launch_bounds__(kBlock) void k_tail_fwd(const bf16* feat, const float* x, const float* w, const float* b,
                                                     int64_t npx, float* out) {
    const int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (p >= npx) return;
    float f[32];
    const uint4* fp = reinterpret_cast<const uint4*>(feat + p * 32);
    for (int q = 0; q < 4; ++q) {
        const uint4 v = fp[q];
        const bf16x8 e = *reinterpret_cast<const bf16x8*>(&v);
        for (int i = 0; i < 8; ++i) f[q * 8 + i] = (float)e[i];
    }
    for (int o = 0; o < 3; ++o) {
        float s = b[o];
        for (int c = 0; c < 32; ++c) s += w[o * 32 + c] * f[c];
        if (kRadiance) out[p * 3 + o] = from_log(to_log(x[p * 3 + o]) + s);
        else out[p * 3 + o] = x[p * 3 + o] + s;
    }
}

// dfeat = W_t^T dout (BF16), and per-block partial sums of dW_t (3 x 32) and db_t (3) as 99
// values [o][33] (32 weights, then the bias). Threads stride over the pixels with the grid,
// keep their 99 sums in registers, and reduce them once per block (waves, then LDS in wave
// order): the order of every sum is fixed by the grid.
constexpr int kTailBwdBlocks = 1024;
__global__ __launch_bounds__(kBlock) void k_tail_bwd(const bf16* feat, const float* dout, const float* w, int64_t npx,
                                                     bf16* dfeat, float* partial /* [gridDim.x][99] */) {
    __shared__ float s[kBlock / 64][99];
    float acc[99];
    for (int k = 0; k < 99; ++k) acc[k] = 0.f;
    for (int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x; p < npx; p += (int64_t)gridDim.x * kBlock) {
        float f[32], d[3];
        const uint4* fp = reinterpret_cast<const uint4*>(feat + p * 32);
        for (int q = 0; q < 4; ++q) {
            const uint4 v = fp[q];
            const bf16x8 e = *reinterpret_cast<const bf16x8*>(&v);
            for (int i = 0; i < 8; ++i) f[q * 8 + i] = (float)e[i];
        }
        for (int o = 0; o < 3; ++o) d[o] = dout[p * 3 + o];
        bf16x8 g8[4];
        for (int c = 0; c < 32; ++c) g8[c / 8][c % 8] = (bf16)(w[c] * d[0] + w[32 + c] * d[1] + w[64 + c] * d[2]);
        uint4* dp = reinterpret_cast<uint4*>(dfeat + p * 32);
        for (int q = 0; q < 4; ++q) dp[q] = *reinterpret_cast<const uint4*>(&g8[q]);
        for (int o = 0; o < 3; ++o) {
            for (int c = 0; c < 32; ++c) acc[o * 33 + c] += d[o] * f[c];
            acc[o * 33 + 32] += d[o];
        }
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int k = 0; k < 99; ++k) {
        float v = acc[k];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        if (lane == 0) s[wave][k] = v;
    }
    __syncthreads();
    if (threadIdx.x < 99) {
        float t = s[0][threadIdx.x];
        for (int wv = 1; wv < kBlock / 64; ++wv) t += s[wv][threadIdx.x];
        partial[(int64_t)blockIdx.x * 99 + threadIdx.x] = t;
    }
}

// dW_t and db_t from the blocks' partials: column k of [blocks][99] summed as in
// k_reduce_rows (64 columns per block, four waves over every fourth row, wave order).
__global__ __launch_bounds__(kBlock) void k_tail_reduce(const float* partial, int blocks, float* dw, float* db) {
    __shared__ float s[kBlock];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int k = blockIdx.x * 64 + lane;
    float sum = 0.f;
    if (k < 99)
        for (int b = wave; b < blocks; b += kBlock / 64) sum += partial[(int64_t)b * 99 + k];
    s[threadIdx.x] = sum;
    __syncthreads();
    if (wave == 0 && k < 99) {
        const float t = ((s[lane] + s[lane + 64]) + s[lane + 128]) + s[lane + 192];
        const int o = k / 33, c = k % 33;
        if (c < 32) dw[o * 32 + c] = t;
        else db[o] = t;
    }
}

// ------------------------------------------------------------------ upsample + concat + conv
// Inference form of a decoder level's first convolution, relu(conv3x3(cat(up2(lo), skip)) + b),
// that never writes the upsampled tensor. Bilinear 2x upsampling (align_corners = False, the
// source index clamped into the image) gives, along each axis, up[2i] = .25 x[i-1] + .75 x[i]
// and up[2i+1] = .75 x[i] + .25 x[i+1], with x replicated past its edges. So at an output pixel
// of phase (py, px) = (Y & 1, X & 1) the 3x3 convolution of up equals a 3x3 convolution of x
// around (Y >> 1, X >> 1) whose weights mix the original taps:
//   w_p[dy][dx] = sum_{ky,kx} M_py[dy][ky] M_px[dx][kx] w[ky][kx],
//   M_0 = [[.75, .25, 0], [.25, .75, .75], [0, 0, .25]], M_1 = [[.25, 0, 0], [.75, .75, .25], [0, .25, .75]]
// (row: low-resolution tap -1, 0, +1; column: full-resolution tap -1, 0, +1). The identity holds
// wherever the 3x3 window stays inside the upsampled image. On the outermost ring of output
// pixels the window reaches the convolution's zero padding, which the phase form would read as
// replicated values, so the ring is computed from explicitly upsampled neighbourhoods.
//
// Tiles: 8 low-resolution rows x 32 columns = 16 x 64 output pixels x 32 output channels. Each
// of the kUpNW waves owns kUpRW low-resolution rows and the four phases of their output rows
// (4 kUpRW 32x32 accumulators; an MFMA covers 32 pixels of one phase along a row). Per 16-channel chunk of lo
// the block stages the 10 x 34 low-resolution halo (replicated at the edges) and the four phase
// weight slices; per chunk of skip the 18 x 66 full-resolution halo (zero padded), even and odd
// columns in separate planes so that the 32 same-phase pixels of an A fragment (every other
// column) are consecutive in LDS, and the plain weight slice. A wave reads every distinct
// fragment of a chunk once: lo, per kernel column, kUpRW + 2 halo rows and 12 phase-weight
// fragments; skip, the 9 weight fragments and 4 (2 kUpRW + 2) halo fragments (output row + kernel
// row and output column parity + kernel column coincide across phases) for 36 kUpRW MFMAs.
// The grid is persistent, one block per CU: a block walks its tiles as one stream of chunks in
// two LDS buffers, loaded by direct global-to-LDS loads one chunk ahead of the MFMAs, also
// across tile boundaries. A tile's epilogue stages its output through the buffer of its last
// chunk in two halves and stores 16 bytes per thread; pixels on the ring are not stored.
// Ring work: each wave computes a strip of 32 ring pixels (part of the top or bottom row, or of
// the left or right column without its corners) as one 32x32 MFMA tile over an explicitly
// upsampled 3 x 34 neighbourhood (k_up_fwd's arithmetic, BF16, zero outside the image) with the
// original weights. Groups of 4 strips go to the blocks with the fewest tiles, before their
// tiles.
// LDS slots are 16 bytes (8 channels); a pixel of a 16-channel chunk has two, XOR-swizzled by bit
// 3 of the pixel index, and a weight row (phase, output channel) 18, its two parts per tap
// swizzled by bit 3 of the row, so that the 32-lane fragment reads hit distinct banks.
constexpr int kUpNW = 8, kUpT = kUpNW * 64;           // waves per block, threads
constexpr int kUpLR = 8, kUpLC = 32;                  // low-resolution rows and columns per tile
constexpr int kUpRW = kUpLR / kUpNW;                  // low-resolution rows per wave
constexpr int kUpUH = kUpLR + 2, kUpUW = kUpLC + 2;   // low-resolution halo: 10 x 34
constexpr int kUpSH = 2 * kUpLR + 2, kUpSW = 2 * kUpLC + 2, kUpSP = kUpSW / 2;  // 18 x 66, planes of 33
// slots of a chunk: halo (padded to whole waves, so that each load instruction of a wave reads
// one array) + weights; lo 704 + 2304, skip 2432 + 576
constexpr int kUpUHalo = (kUpUH * kUpUW * 2 + 63) / 64 * 64, kUpUWt = 4 * 32 * 18;
constexpr int kUpSHalo = (kUpSH * kUpSW * 2 + 63) / 64 * 64, kUpSWt = 32 * 18;
constexpr int kUpUsed = kUpUHalo + kUpUWt > kUpSHalo + kUpSWt ? kUpUHalo + kUpUWt : kUpSHalo + kUpSWt;
constexpr int kUpBuf = (kUpUsed + kUpT - 1) / kUpT * kUpT;        // slots per buffer (3072)
constexpr int kUpLoads = kUpBuf / kUpT;              // buffer loads per thread per chunk
constexpr int kUpRing = 3 * kUpUW * 2;                // slots of one ring strip's neighbourhood
constexpr int kUpMaxCout = 1024;
static_assert(kUpNW * kUpRing + kUpSWt <= kUpBuf, "ring neighbourhoods and weights fit one buffer");
static_assert(8 * 64 * 32 * 2 <= kUpBuf * 16, "half of the output tile fits one buffer");

struct UpConvArgs {
    const bf16* lo;       // [N,h,w,Cu]: input of the upsample
    const bf16* skip;     // [N,2h,2w,Cs] or null (Cs 0)
    const bf16* wp;       // [4][Cout][9][Cu]: phase weights, phase = 2 py + px
    const bf16* wk;       // [Cout][9][Cu + Cs]: original weights (skip part, ring pixels)
    const float* bias;    // [Cout] or null
    bf16* out;            // [N,2h,2w,Cout]
    int Cu, Cs, Cout, N, h, w, relu;
    int tiles;            // tiles x output-channel groups (work items, tile-major)
    int ring_groups;      // groups of kUpNW ring strips x output-channel groups
};

__device__ __forceinline__ int up_px_slot(int p, int k) { return p * 2 + (k ^ ((p >> 3) & 1)); }
__device__ __forceinline__ int up_w_slot(int row, int tap, int k) { return row * 18 + tap * 2 + (k ^ ((row >> 3) & 1)); }

// 8 channels of up2(lo) at full-resolution (Y, X), rounded to BF16 as k_up_fwd stores them.
__device__ __forceinline__ uint4 up_value(const bf16* lo, int n, int h, int w, int C, int c, int Y, int X) {
    int y0, y1, x0, x1;
    float ly, lx;
    lin_src(Y, h, y0, y1, ly);
    lin_src(X, w, x0, x1, lx);
    const bf16* r0 = lo + ((int64_t)n * h + y0) * w * C + c;
    const bf16* r1 = lo + ((int64_t)n * h + y1) * w * C + c;
    const bf16x8 a = *reinterpret_cast<const bf16x8*>(r0 + (int64_t)x0 * C);
    const bf16x8 b = *reinterpret_cast<const bf16x8*>(r0 + (int64_t)x1 * C);
    const bf16x8 cc = *reinterpret_cast<const bf16x8*>(r1 + (int64_t)x0 * C);
    const bf16x8 d = *reinterpret_cast<const bf16x8*>(r1 + (int64_t)x1 * C);
    bf16x8 v;
#pragma unroll
    for (int e = 0; e < 8; ++e)
        v[e] = (bf16)((1.f - ly) * ((1.f - lx) * (float)a[e] + lx * (float)b[e]) +
                      ly * ((1.f - lx) * (float)cc[e] + lx * (float)d[e]));
    return *reinterpret_cast<const uint4*>(&v);
}

// Ring work item `item` (group of kUpNW strips x output-channel group) into LDS buffer s.
__device__ __forceinline__ void up_ring(const UpConvArgs a, int item, uint4* s) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int l32 = lane & 31, kg = lane >> 5;
    const int cog = a.Cout / 32, co0 = (item % cog) * 32, group = item / cog;
    const int H = 2 * a.h, W = 2 * a.w, Cin = a.Cu + a.Cs;
    const int rs = (W + 31) / 32, cs = (H - 2 + 31) / 32, per = 2 * rs + 2 * cs;
    // this wave's strip: n, kind (0 top, 1 bottom, 2 left, 3 right), start along the strip
    const int strip = group * kUpNW + wave;
    const bool live = strip < a.N * per;
    int n = 0, kind = 0, s0 = 0, len = 0;
    if (live) {
        n = strip / per;
        int r = strip - n * per;
        if (r < 2 * rs) { kind = r / rs; s0 = (r % rs) * 32; len = W; }
        else { r -= 2 * rs; kind = 2 + r / cs; s0 = 1 + (r % cs) * 32; len = H - 1; }
    }
    const bool row = kind < 2;
    const int fixed = kind == 0 ? 0 : kind == 1 ? H - 1 : kind == 2 ? 0 : W - 1;
    uint4* nb = s + wave * kUpRing;       // [3][34] pixels, across the strip x along it
    uint4* sw = s + kUpNW * kUpRing;      // 32 weight rows
    f32x16 acc;
    for (int j = 0; j < 16; ++j) acc[j] = 0.f;
    for (int ci0 = 0; ci0 < Cin; ci0 += 16) {
        const bool up = ci0 < a.Cu;
        __syncthreads();
        if (live)
            for (int t = lane; t < kUpRing; t += 64) {
                const int p = t >> 1, k = t & 1, across = p / kUpUW, along = p % kUpUW;
                const int Y = row ? fixed + across - 1 : s0 + along - 1;
                const int X = row ? s0 + along - 1 : fixed + across - 1;
                uint4 v = zero16();
                if (Y >= 0 && Y < H && X >= 0 && X < W) {
                    if (up) v = up_value(a.lo, n, a.h, a.w, a.Cu, ci0 + k * 8, Y, X);
                    else v = *reinterpret_cast<const uint4*>(a.skip + (((int64_t)n * H + Y) * W + X) * a.Cs +
                                                             (ci0 - a.Cu) + k * 8);
                }
                nb[up_px_slot(p, k)] = v;
            }
        for (int t = tid; t < kUpSWt; t += kUpT) {
            const int co = t / 18, r = t % 18, tap = r >> 1, k = r & 1;
            sw[up_w_slot(co, tap, k)] =
                *reinterpret_cast<const uint4*>(a.wk + ((int64_t)(co0 + co) * 9 + tap) * Cin + ci0 + k * 8);
        }
        __syncthreads();
        if (!live) continue;
        for (int tap = 0; tap < 9; ++tap) {
            const int ky = tap / 3, kx = tap % 3;
            const int p = (row ? ky : kx) * kUpUW + l32 + (row ? kx : ky);
            const uint4 av = nb[up_px_slot(p, kg)];
            const uint4 bv = sw[up_w_slot(l32, tap, kg)];
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(*reinterpret_cast<const bf16x8*>(&av),
                                                          *reinterpret_cast<const bf16x8*>(&bv), acc, 0, 0, 0);
        }
    }
    __syncthreads();  // the buffer is free for the caller's next use
    if (!live) return;
    const float b = a.bias ? a.bias[co0 + l32] : 0.f;
    for (int j = 0; j < 16; ++j) {
        const int m = 8 * (j >> 2) + 4 * kg + (j & 3), at = s0 + m;
        if (at >= len) continue;
        const int Y = row ? fixed : at, X = row ? at : fixed;
        float v = acc[j] + b;
        if (a.relu) v = fmaxf(v, 0.f);
        a.out[(((int64_t)n * H + Y) * W + X) * a.Cout + co0 + l32] = (bf16)v;
    }
}

// Per-block constants of k_conv_up's tile stream, for the force-inlined steps below (no lambdas:
// their by-reference closures kept such values in scratch memory, and so does a buffer resource
// inside a struct; the four resources travel as separate arguments).
struct UpK {
    bf16* out;
    int h, w, Cu, Cout, H, W, Cin, cog, tiles_x, tiles_y, nu, nch, P, b, tid, wave, l32, kg, relu;
};
struct UpTile { int n, i0, j0, co0; };

__device__ __forceinline__ UpTile up_tile(const UpK& k, int t) {
    const int item = k.b + t * k.P, q = item / k.cog;
    const int n = q / (k.tiles_x * k.tiles_y), ty = (q / k.tiles_x) % k.tiles_y, tx = q % k.tiles_x;
    return UpTile{n, ty * kUpLR, tx * kUpLC, (item % k.cog) * 32};
}

// Issues chunk s (tile s / nch, channel chunk c = s % nch: c < nu lo channels [16c, 16c + 16),
// else skip channels [16(c - nu), ...)) into buffer base. Slot t of the buffer is filled by
// thread t % kUpT in step t / kUpT with the 16 bytes the layout puts there (offset past the
// range: zeros).
__device__ __forceinline__ void up_issue(const UpK& k, __amdgpu_buffer_rsrc_t rl, __amdgpu_buffer_rsrc_t rs,
                                         __amdgpu_buffer_rsrc_t rp, __amdgpu_buffer_rsrc_t rw, int s, uint4* base) {
    const UpTile T = up_tile(k, s / k.nch);
    const int c = s % k.nch;
    const bool up = c < k.nu;
    const unsigned ch0 = up ? c * 16 : (c - k.nu) * 16;
    const int Y0 = 2 * T.i0, X0 = 2 * T.j0;
#pragma unroll
    for (int i = 0; i < kUpLoads; ++i) {
        int t = i * kUpT + k.tid;
        // opaque to the optimizer: the slot decode below is recomputed at each call instead of
        // being hoisted out of the chunk loop for all kUpLoads slots (registers)
        asm volatile("" : "+v"(t));
        lds_void_t* dst = (lds_void_t*)(base + i * kUpT + k.wave * 64);
        // halo or weights: the same for the whole wave (the halo regions are whole waves); taken
        // from lane 0 so that the branch is uniform and each load has one scalar buffer resource
        const int tw = __builtin_amdgcn_readfirstlane(t);
        const bool halo = tw < (up ? kUpUHalo : kUpSHalo);
        unsigned off = kDlOut;
        if (up && halo) {
            const int p = t >> 1, q = (t & 1) ^ ((p >> 3) & 1), r = p / kUpUW, cc = p % kUpUW;
            const int yy = min(max(T.i0 + r - 1, 0), k.h - 1), xx = min(max(T.j0 + cc - 1, 0), k.w - 1);
            const unsigned o = ((((unsigned)T.n * k.h + yy) * k.w + xx) * k.Cu + ch0 + q * 8) * 2u;
            off = r < kUpUH ? o : kDlOut;
        } else if (up) {
            const int u = t - kUpUHalo, row = u / 18, r = u % 18, tap = r >> 1;
            const int q = (r & 1) ^ ((row >> 3) & 1);
            const unsigned o =
                (((((unsigned)(row >> 5) * k.Cout + T.co0 + (row & 31)) * 9 + tap) * k.Cu) + ch0 + q * 8) * 2u;
            off = u < kUpUWt ? o : kDlOut;
        } else if (halo) {
            const int p = t >> 1, q = (t & 1) ^ ((p >> 3) & 1), r = p / kUpSW, v = p % kUpSW;
            const int cc = 2 * (v % kUpSP) + v / kUpSP;
            const int Y = Y0 + r - 1, X = X0 + cc - 1;
            const unsigned o = ((((unsigned)T.n * k.H + Y) * k.W + X) * (k.Cin - k.Cu) + ch0 + q * 8) * 2u;
            off = (r < kUpSH && Y >= 0 && Y < k.H && X >= 0 && X < k.W) ? o : kDlOut;
        } else {
            const int u = t - kUpSHalo, co = u / 18, r = u % 18, tap = r >> 1;
            const int q = (r & 1) ^ ((co >> 3) & 1);
            const unsigned o = ((((unsigned)(T.co0 + co) * 9 + tap) * k.Cin) + k.Cu + ch0 + q * 8) * 2u;
            off = u < kUpSWt ? o : kDlOut;
        }
        const __amdgpu_buffer_rsrc_t rsrc = up ? (halo ? rl : rp) : (halo ? rs : rw);
        __builtin_amdgcn_raw_ptr_buffer_load_lds(rsrc, dst, 16, off, 0, 0, 0);
    }
}

__device__ __forceinline__ void up_mfma(f32x16& acc, const uint4& av, const uint4& bv) {
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(*reinterpret_cast<const bf16x8*>(&av),
                                                  *reinterpret_cast<const bf16x8*>(&bv), acc, 0, 0, 0);
}

// The MFMAs of channel chunk c from buffer sb; acc [low-resolution row of the wave][phase].
__device__ __forceinline__ void up_compute(const UpK& k, int c, const uint4* sb, f32x16 (&acc)[kUpRW][4]) {
    int wave = k.wave, l32 = k.l32, kg = k.kg;
    // opaque lane indices: the fragment addresses are computed here, not hoisted out of the
    // chunk loop for every read of both chunk kinds (hundreds of registers, spills)
    asm volatile("" : "+v"(wave), "+v"(l32), "+v"(kg));
    if (c < k.nu) {
        const uint4* sw = sb + kUpUHalo;
#pragma unroll
        for (int dx = 0; dx < 3; ++dx) {
            // keeps the scheduler from hoisting the next column's LDS reads (registers)
            __builtin_amdgcn_sched_barrier(0);
            uint4 bv[3][4];
#pragma unroll
            for (int dy = 0; dy < 3; ++dy)
#pragma unroll
                for (int ph = 0; ph < 4; ++ph) bv[dy][ph] = sw[up_w_slot(ph * 32 + l32, dy * 3 + dx, kg)];
#pragma unroll
            for (int R = 0; R < kUpRW + 2; ++R) {   // halo row RW wave + R = row r of the wave + tap row dy
                const uint4 av = sb[up_px_slot((kUpRW * wave + R) * kUpUW + l32 + dx, kg)];
#pragma unroll
                for (int r = 0; r < kUpRW; ++r) {
                    const int dy = R - r;
                    if (dy < 0 || dy > 2) continue;
#pragma unroll
                    for (int ph = 0; ph < 4; ++ph) up_mfma(acc[r][ph], av, bv[dy][ph]);
                }
            }
        }
    } else {
        const uint4* sw = sb + kUpSHalo;
        uint4 bv[9];
#pragma unroll
        for (int tap = 0; tap < 9; ++tap) bv[tap] = sw[up_w_slot(l32, tap, kg)];
#pragma unroll
        for (int R = 0; R < 2 * kUpRW + 2; ++R) {   // halo row 2 RW wave + R = output row o + kernel row ky
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int C = 0; C < 4; ++C) {   // halo column 2 l32 + C = output parity px + kernel column kx
                const int cc = 2 * l32 + C;
                const uint4 av = sb[up_px_slot((2 * kUpRW * wave + R) * kUpSW + (C & 1) * kUpSP + (cc >> 1), kg)];
#pragma unroll
                for (int o = 0; o < 2 * kUpRW; ++o) {
                    const int ky = R - o;
                    if (ky < 0 || ky > 2) continue;
#pragma unroll
                    for (int px = 0; px < 2; ++px) {
                        const int kx = C - px;
                        if (kx < 0 || kx > 2) continue;
                        up_mfma(acc[o >> 1][(o & 1) * 2 + px], av, bv[ky * 3 + kx]);
                    }
                }
            }
        }
    }
}

// Epilogue of the block's tile t through buffer sb: bias, ReLU, output rows 0..7 (the first half
// of the waves) then 8..15 as [row][column][channel] BF16, 16-byte stores of the pixels off the
// ring.
__device__ __forceinline__ void up_epilogue(const UpK& k, int t, uint4* sb, const float* s_bias,
                                            const f32x16 (&acc)[kUpRW][4]) {
    typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
    typedef __attribute__((address_space(1))) u32x4 global_u32x4;
    const UpTile T = up_tile(k, t);
    const int Y0 = 2 * T.i0, X0 = 2 * T.j0;
    int wave = k.wave, l32 = k.l32, kg = k.kg;
    asm volatile("" : "+v"(wave), "+v"(l32), "+v"(kg));   // as in up_compute
    const float bias = s_bias[T.co0 + l32];
    bf16* so = reinterpret_cast<bf16*>(sb);
#pragma unroll
    for (int half = 0; half < 2; ++half) {
        __syncthreads();
        if (wave / (kUpNW / 2) == half) {
#pragma unroll
            for (int r = 0; r < kUpRW; ++r)
#pragma unroll
                for (int ph = 0; ph < 4; ++ph)
#pragma unroll
                    for (int j = 0; j < 16; ++j) {
                        const int m = 8 * (j >> 2) + 4 * kg + (j & 3);
                        float v = acc[r][ph][j] + bias;
                        if (k.relu) v = fmaxf(v, 0.f);
                        const int orow = 2 * kUpRW * (wave % (kUpNW / 2)) + 2 * r + (ph >> 1);
                        so[(orow * 64 + 2 * m + (ph & 1)) * 32 + l32] = (bf16)v;
                    }
        }
        __syncthreads();
        for (int q = k.tid; q < 8 * 64 * 4; q += kUpT) {
            const int px = q >> 2, part = q & 3, ry = px >> 6, rx = px & 63;
            const int Y = Y0 + 8 * half + ry, X = X0 + rx;
            if (Y <= 0 || Y >= k.H - 1 || X <= 0 || X >= k.W - 1) continue;
            // a global (not flat) store: it joins the in-order vector memory count the chunk
            // stream's waits rely on
            *(global_u32x4*)(k.out + (((int64_t)T.n * k.H + Y) * k.W + X) * k.Cout + T.co0 + part * 8) =
                *reinterpret_cast<const u32x4*>(so + (ry * 64 + rx) * 32 + part * 8);
        }
    }
}

// Step s of the chunk stream: chunk s is in buffer cur, and chunk s + 1 goes to buffer next,
// which chunk s - 1 used. The wait (all of this thread's loads: chunk s, and the previous tile's
// stores) and the barrier make chunk s visible to every thread and free buffer next.
__device__ __forceinline__ void up_step(const UpK& k, __amdgpu_buffer_rsrc_t rl, __amdgpu_buffer_rsrc_t rs,
                                        __amdgpu_buffer_rsrc_t rp, __amdgpu_buffer_rsrc_t rw, int s, int S,
                                        const uint4* cur, uint4* next, const float* s_bias,
                                        f32x16 (&acc)[kUpRW][4]) {
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    if (s + 1 < S) up_issue(k, rl, rs, rp, rw, s + 1, next);
    const int c = s % k.nch;
    if (c == 0) {
#pragma unroll
        for (int r = 0; r < kUpRW; ++r)
#pragma unroll
            for (int ph = 0; ph < 4; ++ph)
#pragma unroll
                for (int j = 0; j < 16; ++j) acc[r][ph][j] = 0.f;
    }
    up_compute(k, c, cur, acc);
    if (c == k.nch - 1) up_epilogue(k, s / k.nch, const_cast<uint4*>(cur), s_bias, acc);
}

__global__ __launch_bounds__(kUpT, 1) void k_conv_up(UpConvArgs a) {
    // two buffers as two LDS objects (see k_conv_dl)
    __shared__ __attribute__((aligned(16))) uint4 s_b0[kUpBuf];
    __shared__ __attribute__((aligned(16))) uint4 s_b1[kUpBuf];
    __shared__ float s_bias[kUpMaxCout];   // read from LDS so that no global load joins the vmcnt stream
    for (int t = threadIdx.x; t < a.Cout; t += kUpT) s_bias[t] = a.bias ? a.bias[t] : 0.f;
    __syncthreads();
    const int P = gridDim.x, b = blockIdx.x;
    // ring groups to the blocks with the fewest tiles (the last ones), before their tiles
    for (int r = P - 1 - b; r < a.ring_groups; r += P) up_ring(a, r, s_b0);
    UpK k;
    k.h = a.h; k.w = a.w; k.Cu = a.Cu; k.Cout = a.Cout; k.H = 2 * a.h; k.W = 2 * a.w; k.Cin = a.Cu + a.Cs;
    k.cog = a.Cout / 32; k.tiles_x = (a.w + kUpLC - 1) / kUpLC; k.tiles_y = (a.h + kUpLR - 1) / kUpLR;
    k.nu = a.Cu / 16; k.nch = k.nu + a.Cs / 16; k.P = P; k.b = b; k.relu = a.relu; k.out = a.out;
    k.tid = threadIdx.x; k.wave = k.tid >> 6; k.l32 = k.tid & 31; k.kg = (k.tid >> 5) & 1;
    // buffer resources (byte sizes < 2^31 checked by the launcher)
    const __amdgpu_buffer_rsrc_t rl =
        __builtin_amdgcn_make_buffer_rsrc((void*)a.lo, 0, (int)((int64_t)a.N * a.h * a.w * a.Cu * 2), kDlRsrcWord3);
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
        (void*)(a.skip ? a.skip : a.lo), 0, (int)((int64_t)a.N * k.H * k.W * a.Cs * 2), kDlRsrcWord3);
    const __amdgpu_buffer_rsrc_t rp =
        __builtin_amdgcn_make_buffer_rsrc((void*)a.wp, 0, (int)((int64_t)4 * a.Cout * 9 * a.Cu * 2), kDlRsrcWord3);
    const __amdgpu_buffer_rsrc_t rw =
        __builtin_amdgcn_make_buffer_rsrc((void*)a.wk, 0, (int)((int64_t)a.Cout * 9 * k.Cin * 2), kDlRsrcWord3);
    const int my_tiles = b < a.tiles ? (a.tiles - 1 - b) / P + 1 : 0;
    const int S = my_tiles * k.nch;   // chunks this block streams
    f32x16 acc[kUpRW][4];
    if (S > 0) up_issue(k, rl, rs, rp, rw, 0, s_b0);
    for (int s = 0; s < S; s += 2) {
        up_step(k, rl, rs, rp, rw, s, S, s_b0, s_b1, s_bias, acc);
        if (s + 1 >= S) break;
        up_step(k, rl, rs, rp, rw, s + 1, S, s_b1, s_b0, s_bias, acc);
    }
}

// Phase weights of k_conv_up from the FP32 weights [Cout][9][Cin] (first Cu input channels):
// wp[ph][co][dy * 3 + dx][ci] = sum_{ky,kx} M_py[dy][ky] M_px[dx][kx] w[co][ky * 3 + kx][ci],
// summed in FP32 in (ky, kx) order and rounded to BF16.
__global__ __launch_bounds__(kBlock) void k_up_phase_weights(const float* w, int Cout, int Cin, int Cu, bf16* wp) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int64_t total = (int64_t)4 * Cout * 9 * Cu;
    if (i >= total) return;
    const int ci = (int)(i % Cu);
    const int tap = (int)((i / Cu) % 9);
    const int co = (int)((i / ((int64_t)Cu * 9)) % Cout);
    const int ph = (int)(i / ((int64_t)Cu * 9 * Cout));
    const float M[2][3][3] = {{{.75f, .25f, 0.f}, {.25f, .75f, .75f}, {0.f, 0.f, .25f}},
                              {{.25f, 0.f, 0.f}, {.75f, .75f, .25f}, {0.f, .25f, .75f}}};
    const int py = ph >> 1, px = ph & 1, dy = tap / 3, dx = tap % 3;
    float s = 0.f;
    for (int ky = 0; ky < 3; ++ky)
        for (int kx = 0; kx < 3; ++kx)
            s += M[py][dy][ky] * M[px][dx][kx] * w[((int64_t)co * 9 + ky * 3 + kx) * Cin + ci];
    wp[i] = (bf16)s;
}

// ------------------------------------------------------------------ tile-shape selection
// Convolution variants: 0 = 8x16-pixel tile x 32/64 channels (k_conv), 1..4 = register-blocked
// k_conv_rb<RW, NB>, 5..7 = k_conv32 on 32x32x16 MFMAs. Which one is fastest depends on the
// layer (pixels, channels, how many blocks fill 256 CUs), so the first call of each shape
// times every variant that fits it on
// the caller's own buffers and stream, and later calls reuse the winner (the role cuDNN's
// benchmark mode plays for the reference's PyTorch UNet). Under stream capture, or with
// tuning off, the shape heuristic below is used.
enum ConvVariant {
    kVarSimple = 0, kVarRb4x4, kVarRb8x2, kVarRb4x2, kVarRb8x4,
    kVarM32r2n64, kVarM32r2n32, kVarM32r4n32,  // k_conv32<TN, RW> (RW 4 with TN 64 spills)
    kVarDl4n64, kVarDl2n64, kVarDl4n32,        // k_conv_dl<TN, RW> (direct-to-LDS staging)
    kVarF8r2n64, kVarF8r1n64, kVarF8r2n32,     // k_conv_f8<TN, RW> (FP8 operands only)
    kNumVariants
};
static_assert(kNumVariants == LGI_UNET_CONV_VARIANTS, "variant count in the C ABI");
constexpr int kVarRW[kNumVariants] = {0, 4, 8, 4, 8, 2, 2, 4, 4, 2, 4, 2, 1, 2};
constexpr int kVarNB[kNumVariants] = {0, 4, 2, 2, 4, 4, 2, 2, 4, 4, 2, 4, 4, 2};  // output channels / 16

struct TuneEntry {
    int variant;
    float us[kNumVariants];
};
std::mutex g_tune_mu;
std::map<std::array<int, 9>, TuneEntry> g_tune;
std::atomic<int> g_tune_mode{1};
std::atomic<int> g_force_variant{-1};

// With pool_out, the k_conv32 variants write the pooled copy from their output tile and the
// others are followed by a k_pool_fwd launch; the tuner times each pair as one candidate.
bool variant_fits(int v, const ConvArgs& a, int taps) {
    // FP8 operands run only on the FP8 kernel, which needs 64-channel chunks of one source
    const bool f8v = v >= kVarF8r2n64 && v < kNumVariants;
    if (a.fp8 != 0 || f8v)
        return a.fp8 != 0 && f8v && taps == 9 && !a.mask0 && !a.mask1 && a.C0 % kCK8 == 0 && a.C1 % kCK8 == 0 &&
               a.Cout % (16 * kVarNB[v]) == 0;
    // the fused tail lives in the epilogue of the 32-channel 32x32-MFMA tilings
    if (a.tail_out && !(v == kVarM32r2n32 || v == kVarM32r4n32 || v == kVarDl4n32)) return false;
    if (v == kVarSimple) return true;
    if (taps != 9 || v < 0 || v >= kNumVariants) return false;
    if (v >= kVarDl4n64) {
        // direct loads stage raw input (no ReLU-mask multiply) through 32-bit buffer offsets
        const int64_t px = (int64_t)a.N * a.H * a.W, lim = int64_t(1) << 31;
        if (a.mask0 || a.mask1 || px * a.C0 * 2 >= lim || px * a.C1 * 2 >= lim ||
            (int64_t)a.Cout * 9 * (a.C0 + a.C1) * 2 >= lim)
            return false;
    }
    return a.Cout % (16 * kVarNB[v]) == 0;
}

int heuristic_variant(const ConvArgs& a, int taps) {
    (void)taps;
    if (a.fp8) return a.Cout % 64 == 0 ? kVarF8r2n64 : kVarF8r2n32;
    return (a.pool_out || a.tail_out) ? kVarM32r2n32 : kVarSimple;
}

void launch_variant(int v, const ConvArgs& a, int taps, hipStream_t st) {
    if (a.pool_out && v < kVarM32r2n64) {
        ConvArgs c = a;
        c.pool_out = nullptr;
        launch_variant(v, c, taps, st);
        hipLaunchKernelGGL(k_pool_fwd, row_grid(a.N, a.H / 2, a.W / 2, a.Cout / 8), dim3(kBlock), 0, st, a.out0, a.N,
                           a.H, a.W, a.Cout, a.pool_out);
        return;
    }
    if (v == kVarSimple) {
        const int tiles = cdiv(a.W, kTW) * cdiv(a.H, kTH) * a.N;
        const bool tn64 = (a.Cout % 64 == 0) && (a.Cs % 64 == 0);
        const bool m = a.mask0 || a.mask1;
        const dim3 g64(tiles, a.Cout / 64), g32(tiles, a.Cout / 32);
        if (taps == 9) {
            if (tn64 && m) hipLaunchKernelGGL((k_conv<64, 3, true>), g64, dim3(kBlock), 0, st, a);
            else if (tn64) hipLaunchKernelGGL((k_conv<64, 3, false>), g64, dim3(kBlock), 0, st, a);
            else if (m) hipLaunchKernelGGL((k_conv<32, 3, true>), g32, dim3(kBlock), 0, st, a);
            else hipLaunchKernelGGL((k_conv<32, 3, false>), g32, dim3(kBlock), 0, st, a);
        } else {
            if (tn64) hipLaunchKernelGGL((k_conv<64, 1, true>), g64, dim3(kBlock), 0, st, a);
            else hipLaunchKernelGGL((k_conv<32, 1, true>), g32, dim3(kBlock), 0, st, a);
        }
        return;
    }
    const int th = kVarRW[v] * (kBlock / 64), tn = 16 * kVarNB[v];
    if (v >= kVarF8r2n64) {
        const dim3 grid(cdiv(a.W, kTW32) * cdiv(a.H, th) * a.N, a.Cout / tn);
        switch (v) {
            case kVarF8r2n64: hipLaunchKernelGGL((k_conv_f8<64, 2>), grid, dim3(kBlock), 0, st, a); break;
            case kVarF8r1n64: hipLaunchKernelGGL((k_conv_f8<64, 1>), grid, dim3(kBlock), 0, st, a); break;
            case kVarF8r2n32: hipLaunchKernelGGL((k_conv_f8<32, 2>), grid, dim3(kBlock), 0, st, a); break;
            default: break;
        }
        return;
    }
    if (v >= kVarDl4n64) {
        const dim3 grid(cdiv(a.W, kTW32) * cdiv(a.H, th) * a.N, a.Cout / tn);
        switch (v) {
            case kVarDl4n64: hipLaunchKernelGGL((k_conv_dl<64, 4>), grid, dim3(kBlock), 0, st, a); break;
            case kVarDl2n64: hipLaunchKernelGGL((k_conv_dl<64, 2>), grid, dim3(kBlock), 0, st, a); break;
            case kVarDl4n32: hipLaunchKernelGGL((k_conv_dl<32, 4>), grid, dim3(kBlock), 0, st, a); break;
            default: break;
        }
        return;
    }
    if (v >= kVarM32r2n64) {
        const dim3 grid(cdiv(a.W, kTW32) * cdiv(a.H, th) * a.N, a.Cout / tn);
        const bool m = a.mask0 || a.mask1;
        switch (v) {
            case kVarM32r2n64:
                if (m) hipLaunchKernelGGL((k_conv32<64, 2, true>), grid, dim3(kBlock), 0, st, a);
                else hipLaunchKernelGGL((k_conv32<64, 2, false>), grid, dim3(kBlock), 0, st, a);
                break;
            case kVarM32r2n32:
                if (m) hipLaunchKernelGGL((k_conv32<32, 2, true>), grid, dim3(kBlock), 0, st, a);
                else hipLaunchKernelGGL((k_conv32<32, 2, false>), grid, dim3(kBlock), 0, st, a);
                break;
            case kVarM32r4n32:
                if (m) hipLaunchKernelGGL((k_conv32<32, 4, true>), grid, dim3(kBlock), 0, st, a);
                else hipLaunchKernelGGL((k_conv32<32, 4, false>), grid, dim3(kBlock), 0, st, a);
                break;
            default: break;
        }
        return;
    }
    const dim3 grid(cdiv(a.W, kTW) * cdiv(a.H, th) * a.N, a.Cout / tn);
    switch (v) {
        case kVarRb4x4: hipLaunchKernelGGL((k_conv_rb<4, 4>), grid, dim3(kBlock), 0, st, a); break;
        case kVarRb8x2: hipLaunchKernelGGL((k_conv_rb<8, 2>), grid, dim3(kBlock), 0, st, a); break;
        case kVarRb4x2: hipLaunchKernelGGL((k_conv_rb<4, 2>), grid, dim3(kBlock), 0, st, a); break;
        case kVarRb8x4: hipLaunchKernelGGL((k_conv_rb<8, 4>), grid, dim3(kBlock), 0, st, a); break;
        default: break;
    }
}

// Times every fitting variant (one warm-up and three timed launches each) and returns the
// fastest; us[] receives the per-launch microseconds (0 for variants that do not fit).
int tune_variant(const ConvArgs& a, int taps, hipStream_t st, float* us) {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        if (e0) (void)hipEventDestroy(e0);
        return -1;
    }
    int best = -1;
    for (int v = 0; v < kNumVariants; ++v) {
        us[v] = 0.f;
        if (!variant_fits(v, a, taps)) continue;
        launch_variant(v, a, taps, st);
        (void)hipEventRecord(e0, st);
        for (int r = 0; r < 3; ++r) launch_variant(v, a, taps, st);
        (void)hipEventRecord(e1, st);
        float ms = 0.f;
        if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess) continue;
        us[v] = ms * 1000.f / 3.f;
        if (best < 0 || us[v] < us[best]) best = v;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return best;
}

lgi_status conv_launch(const ConvArgs& a, int taps, hipStream_t st) {
    const int Cin = a.C0 + a.C1;
    if (a.N <= 0 || a.H <= 0 || a.W <= 0 || Cin <= 0 || a.Cout <= 0) return LGI_ERR_INVALID_ARG;
    if (a.C0 % kCK || a.C1 % kCK || a.Cout % 32 || a.Cs % 32 || a.Cs > a.Cout || (a.C1 && !a.src1))
        return LGI_ERR_INVALID_ARG;
    if (a.Cs < a.Cout && !a.out1) return LGI_ERR_INVALID_ARG;
    if (taps != 9 && taps != 1) return LGI_ERR_INVALID_ARG;
    if (a.pool_out && (taps != 9 || a.Cs != a.Cout || a.mask0 || a.mask1 || a.H % 2 || a.W % 2 ||
                       !row_grid_fits(a.N, a.H / 2, a.W / 2, a.Cout / 8)))
        return LGI_ERR_INVALID_ARG;
    if (a.tail_out && (taps != 9 || a.Cout != 32 || a.Cs != 32 || a.mask0 || a.mask1 || a.pool_out || !a.tail_w ||
                       !a.tail_b || !a.tail_x))
        return LGI_ERR_INVALID_ARG;
    int v = g_force_variant.load();
    if (v >= 0) {
        if (!variant_fits(v, a, taps)) return LGI_ERR_UNSUPPORTED;
    } else {
        v = heuristic_variant(a, taps);
        if (taps == 9 && g_tune_mode.load() == 1) {
            // input kind: 0 plain, 1 ReLU-masked (data gradient), 2 plain with a pooled copy,
            // 3 with the denoiser tail in the epilogue; + 8 with FP8 operands
            const int kind = (a.mask0 != nullptr ? 1 : (a.pool_out ? 2 : (a.tail_out ? 3 : 0))) + (a.fp8 ? 8 : 0);
            const std::array<int, 9> key{a.N, a.H, a.W, a.C0, a.C1, a.Cout, a.Cs, kind, a.relu};
            std::lock_guard<std::mutex> lk(g_tune_mu);
            auto it = g_tune.find(key);
            if (it != g_tune.end()) {
                v = it->second.variant;
            } else {
                // timing needs a synchronize, which a capturing stream does not allow
                hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
                if (hipStreamIsCapturing(st, &cs) == hipSuccess && cs == hipStreamCaptureStatusNone) {
                    TuneEntry e{};
                    const int best = tune_variant(a, taps, st, e.us);
                    if (best >= 0) {
                        e.variant = v = best;
                        g_tune.emplace(key, e);
                    }
                }
            }
        }
    }
    launch_variant(v, a, taps, st);
    return hipGetLastError() == hipSuccess ? LGI_OK : LGI_ERR_HIP;
}

}  // namespace

extern "C" {

lgi_status lgi_unet_conv3x3_fwd(const void* src0, int32_t c0, const void* src1, int32_t c1, const void* w,
                                const float* bias, int32_t cout, int32_t n, int32_t h, int32_t wd, int32_t relu,
                                void* out, void* stream) {
    ConvArgs a{};
    a.src0 = (const bf16*)src0; a.src1 = (const bf16*)src1; a.C0 = c0; a.C1 = src1 ? c1 : 0;
    a.wk = (const bf16*)w; a.bias = bias; a.out0 = (bf16*)out; a.out1 = nullptr;
    a.Cout = cout; a.Cs = cout; a.N = n; a.H = h; a.W = wd; a.relu = relu;
    if (!src0 || !w || !out) return LGI_ERR_INVALID_ARG;
    return conv_launch(a, 9, (hipStream_t)stream);
}

lgi_status lgi_unet_conv3x3_pool_fwd(const void* src0, int32_t c0, const void* src1, int32_t c1, const void* w,
                                     const float* bias, int32_t cout, int32_t n, int32_t h, int32_t wd, int32_t relu,
                                     void* out, void* pool_out, void* stream) {
    ConvArgs a{};
    a.src0 = (const bf16*)src0; a.src1 = (const bf16*)src1; a.C0 = c0; a.C1 = src1 ? c1 : 0;
    a.wk = (const bf16*)w; a.bias = bias; a.out0 = (bf16*)out; a.out1 = nullptr;
    a.Cout = cout; a.Cs = cout; a.N = n; a.H = h; a.W = wd; a.relu = relu; a.pool_out = (bf16*)pool_out;
    if (!src0 || !w || !out || !pool_out || h % 2 || wd % 2) return LGI_ERR_INVALID_ARG;
    return conv_launch(a, 9, (hipStream_t)stream);
}

lgi_status lgi_unet_conv3x3_tail_radiance_fwd(const void* src0, int32_t c0, const void* w, const float* bias,
                                              int32_t n, int32_t h, int32_t wd, int32_t relu, const float* tail_w,
                                              const float* tail_b, const float* hdr, float* out, void* stream) {
    ConvArgs a{};
    a.src0 = (const bf16*)src0; a.src1 = nullptr; a.C0 = c0; a.C1 = 0;
    a.wk = (const bf16*)w; a.bias = bias; a.out0 = nullptr; a.out1 = nullptr;
    a.Cout = 32; a.Cs = 32; a.N = n; a.H = h; a.W = wd; a.relu = relu;
    a.tail_w = tail_w; a.tail_b = tail_b; a.tail_x = hdr; a.tail_out = out;
    if (!src0 || !w || !out || !tail_w || !tail_b || !hdr) return LGI_ERR_INVALID_ARG;
    return conv_launch(a, 9, (hipStream_t)stream);
}

lgi_status lgi_unet_conv3x3_fp8_fwd(const void* src0, int32_t c0, const void* src1, int32_t c1, const void* w_e4m3,
                                    const float* w_scale, float act_scale, const float* bias, int32_t cout, int32_t n,
                                    int32_t h, int32_t wd, int32_t relu, void* out, void* pool_out, void* stream) {
    ConvArgs a{};
    a.src0 = (const bf16*)src0; a.src1 = (const bf16*)src1; a.C0 = c0; a.C1 = src1 ? c1 : 0;
    a.wk = (const bf16*)w_e4m3; a.w_scale = w_scale; a.act_scale = act_scale; a.fp8 = 1;
    a.bias = bias; a.out0 = (bf16*)out; a.out1 = nullptr;
    a.Cout = cout; a.Cs = cout; a.N = n; a.H = h; a.W = wd; a.relu = relu; a.pool_out = (bf16*)pool_out;
    if (!src0 || !w_e4m3 || !w_scale || !out || !(act_scale > 0.0f) || (pool_out && (h % 2 || wd % 2)))
        return LGI_ERR_INVALID_ARG;
    return conv_launch(a, 9, (hipStream_t)stream);
}

lgi_status lgi_unet_up_phase_weights(const float* w, int32_t cout, int32_t cin, int32_t cu, void* w_phase,
                                     void* stream) {
    if (!w || !w_phase || cout <= 0 || cu <= 0 || cu > cin) return LGI_ERR_INVALID_ARG;
    const int64_t total = (int64_t)4 * cout * 9 * cu;
    hipLaunchKernelGGL(k_up_phase_weights, dim3(cdiv(total, kBlock)), dim3(kBlock), 0, (hipStream_t)stream, w, cout,
                       cin, cu, (bf16*)w_phase);
    return hipGetLastError() == hipSuccess ? LGI_OK : LGI_ERR_HIP;
}

lgi_status lgi_unet_conv3x3_up_fwd(const void* lo, int32_t cu, const void* skip, int32_t cs, const void* w_phase,
                                   const void* w, const float* bias, int32_t cout, int32_t n, int32_t h, int32_t wd,
                                   int32_t relu, void* out, void* stream) {
    if (!skip) cs = 0;
    if (!lo || !w_phase || !w || !out || n <= 0 || h <= 0 || wd <= 0 || cu <= 0 || cu % 32 || cs < 0 || cs % 32 ||
        cout <= 0 || cout % 32)
        return LGI_ERR_INVALID_ARG;
    if (cout > kUpMaxCout) return LGI_ERR_UNSUPPORTED;
    const int64_t H = 2 * (int64_t)h, W = 2 * (int64_t)wd, cog = cout / 32;
    const int64_t strips = (int64_t)n * (2 * ((W + 31) / 32) + 2 * ((H - 2 + 31) / 32));
    const int64_t ring = (strips + kUpNW - 1) / kUpNW * cog;
    const int64_t tiles = (int64_t)n * ((h + kUpLR - 1) / kUpLR) * ((wd + kUpLC - 1) / kUpLC) * cog;
    // the kernel addresses lo, skip and both weight arrays through 32-bit buffer offsets
    const int64_t lim = int64_t(1) << 31;
    if (ring + tiles > INT32_MAX || (int64_t)n * h * wd * cu * 2 >= lim || (int64_t)n * H * W * cs * 2 >= lim ||
        (int64_t)4 * cout * 9 * cu * 2 >= lim || (int64_t)cout * 9 * (cu + cs) * 2 >= lim)
        return LGI_ERR_UNSUPPORTED;
    // persistent grid: one block per CU (the kernel's 100 KB of LDS allows no more)
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
        return LGI_ERR_HIP;
    UpConvArgs a{};
    a.lo = (const bf16*)lo; a.skip = (const bf16*)skip; a.wp = (const bf16*)w_phase; a.wk = (const bf16*)w;
    a.bias = bias; a.out = (bf16*)out; a.Cu = cu; a.Cs = cs; a.Cout = cout; a.N = n; a.h = h; a.w = wd;
    a.relu = relu; a.tiles = (int)tiles; a.ring_groups = (int)ring;
    const int P = (int)std::min<int64_t>(cus, tiles + ring);
    hipLaunchKernelGGL(k_conv_up, dim3(P), dim3(kUpT), 0, (hipStream_t)stream, a);
    return hipGetLastError() == hipSuccess ? LGI_OK : LGI_ERR_HIP;
}

static lgi_status first_fwd(bool log_in, const float* x, const void* w, int32_t cin_w, const float* bias, int32_t n,
                            int32_t h, int32_t wd, int32_t relu, void* out, void* stream) {
    if (!x || !w || !out || cin_w < 3 || n <= 0 || h <= 0 || wd <= 0) return LGI_ERR_INVALID_ARG;
    const int64_t tiles = (int64_t)cdiv(wd, kTW) * cdiv(h, kTH) * n;
    if (tiles > INT32_MAX) return LGI_ERR_INVALID_ARG;
    // persistent: as many blocks as are resident at once, never more than tiles
    int dev = 0, cus = 256, per_cu = 1;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(
                                                                  log_in ? k_conv_first<true> : k_conv_first<false>),
                                                     kBlock, 0) != hipSuccess)
        return LGI_ERR_HIP;
    const int grid = (int)std::min<int64_t>(tiles, (int64_t)cus * std::max(1, per_cu));
    if (log_in)
        hipLaunchKernelGGL(k_conv_first<true>, dim3(grid), dim3(kBlock), 0, (hipStream_t)stream, x, (const bf16*)w,
                           cin_w, bias, n, h, wd, relu, (bf16*)out);
    else
        hipLaunchKernelGGL(k_conv_first<false>, dim3(grid), dim3(kBlock), 0, (hipStream_t)stream, x, (const bf16*)w,
                           cin_w, bias, n, h, wd, relu, (bf16*)out);
    return hipGetLastError() == hipSuccess ? LGI_OK : LGI_ERR_HIP;
}

lgi_status lgi_unet_conv3x3_first_fwd(const float* x, const void* w, int32_t cin_w, const float* bias, int32_t n,
                                      int32_t h, int32_t wd, int32_t relu, void* out, void* stream) {
    return first_fwd(false, x, w, cin_w, bias, n, h, wd, relu, out, stream);
}

lgi_status lgi_unet_conv3x3_first_radiance_fwd(const float* hdr, const void* w, int32_t cin_w, const float* bias,
                                               int32_t n, int32_t h, int32_t wd, int32_t relu, void* out,
                                               void* stream) {
    return first_fwd(true, hdr, w, cin_w, bias, n, h, wd, relu, out, stream);
}

lgi_status lgi_unet_conv3x3_bwd_data(const void* dy, const void* ymask, const void* w, int32_t cout, int32_t cin,
                                     int32_t cin_split, int32_t n, int32_t h, int32_t wd, void* dx0, void* dx1,
                                     void* wt_scratch, void* stream) {
    if (!dy || !w || !dx0 || !wt_scratch || cin_split <= 0 || cin_split > cin) return LGI_ERR_INVALID_ARG;
    hipStream_t st = (hipStream_t)stream;
    const int64_t total = (int64_t)cout * 9 * cin;
    hipLaunchKernelGGL(k_flip_weights, dim3(cdiv(total, kBlock)), dim3(kBlock), 0, st, (const bf16*)w, cout, cin, 9,
                       (bf16*)wt_scratch);
    ConvArgs a{};
    a.src0 = (const bf16*)dy; a.C0 = cout; a.src1 = nullptr; a.C1 = 0;
    a.mask0 = (const bf16*)ymask; a.mask1 = nullptr;
    a.wk = (const bf16*)wt_scratch; a.bias = nullptr;
    a.out0 = (bf16*)dx0; a.out1 = (bf16*)dx1; a.Cout = cin; a.Cs = cin_split;
    a.N = n; a.H = h; a.W = wd; a.relu = 0;
    return conv_launch(a, 9, st);
}

// Pixel splits of the weight gradient: enough 32 x 32-channel blocks to fill the chip (2048,
// about 8 per CU), at most one per 8x16 tile. Small layers at full resolution get many splits;
// the 512-channel layers few, which keeps their partial sums (splits x cout x 9 x cin floats)
// small.
static int64_t wgrad_splits(int cout, int cin, int64_t tiles) {
    const int64_t per_split = (int64_t)(cout / 32) * (cin / 32);
    return std::max<int64_t>(1, std::min<int64_t>(tiles, cdiv(2048, std::max<int64_t>(1, per_split))));
}

int64_t lgi_unet_wgrad_scratch_floats(int32_t cout, int32_t cin, int32_t n, int32_t h, int32_t wd) {
    if (cout <= 0 || cin <= 0 || n <= 0 || h <= 0 || wd <= 0) return 0;
    const int64_t tiles = (int64_t)cdiv(wd, kTW) * cdiv(h, kTH) * n;
    const int64_t npx = (int64_t)n * h * wd;
    return std::max<int64_t>(wgrad_splits(cout, cin, tiles) * cout * 9 * cin,
                             std::min<int64_t>(1024, cdiv(npx, 64)) * cout);
}

lgi_status lgi_unet_conv3x3_bwd_weight(const void* dy, const void* ymask, int32_t cout, const void* src0, int32_t c0,
                                       const void* src1, int32_t c1, int32_t n, int32_t h, int32_t wd, float* dw,
                                       float* db, float* scratch, int64_t scratch_floats, void* stream) {
    const int cin = c0 + (src1 ? c1 : 0);
    if (!dy || !src0 || !dw || !scratch || cout % 32 || c0 % 32 || (src1 && c1 % 32)) return LGI_ERR_INVALID_ARG;
    hipStream_t st = (hipStream_t)stream;
    const int tiles = cdiv(wd, kTW) * cdiv(h, kTH) * n;
    const int64_t wsize = (int64_t)cout * 9 * cin;
    // as many pixel splits as the scratch holds, at most one per tile, at most 64
    int splits = (int)std::min<int64_t>(wgrad_splits(cout, cin, tiles), scratch_floats / wsize);
    if (splits < 1 || (db && cout > 8 * kBlock)) return LGI_ERR_INVALID_ARG;
    const int per = cdiv(tiles, splits);
    splits = cdiv(tiles, per);
    hipLaunchKernelGGL(k_conv_wgrad, dim3(cout / 32, cin / 32, splits), dim3(kWgBlock), 0, st, (const bf16*)dy,
                       (const bf16*)ymask, cout, (const bf16*)src0, c0, (const bf16*)src1, src1 ? c1 : 0, n, h, wd, per,
                       scratch);
    reduce_splits(st, scratch, wsize, splits, dw, 0);
    if (db) {
        const int64_t npx = (int64_t)n * h * wd;
        const int bsplits = (int)std::min<int64_t>(std::min<int64_t>(1024, cdiv(npx, 64)), scratch_floats / cout);
        const int64_t bper = cdiv(npx, bsplits);
        const int bs = cdiv(npx, bper);
        hipLaunchKernelGGL(k_bias_grad, dim3(bs), dim3(kBlock), 0, st, (const bf16*)dy, (const bf16*)ymask, npx, cout,
                           bper, scratch);
        reduce_splits(st, scratch, (int64_t)cout, bs, db, 0);
    }
    return hipGetLastError() == hipSuccess ? LGI_OK : LGI_ERR_HIP;
}

lgi_status lgi_unet_conv_tuning(int32_t mode) {
    if (mode != 0 && mode != 1) return LGI_ERR_INVALID_ARG;
    g_tune_mode.store(mode);
    return LGI_OK;
}

lgi_status lgi_unet_conv_force_variant(int32_t variant) {
    if (variant < -1 || variant >= kNumVariants) return LGI_ERR_INVALID_ARG;
    g_force_variant.store(variant);
    return LGI_OK;
}

lgi_status lgi_unet_conv_tune_table(int32_t* rows, float* us, int32_t cap, int32_t* n) {
    if (!n || cap < 0 || (cap > 0 && (!rows || !us))) return LGI_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    int i = 0;
    for (const auto& kv : g_tune) {
        if (i < cap) {
            for (int k = 0; k < 9; ++k) rows[i * 10 + k] = kv.first[k];
            rows[i * 10 + 9] = kv.second.variant;
            for (int v = 0; v < kNumVariants; ++v) us[i * kNumVariants + v] = kv.second.us[v];
        }
        ++i;
    }
    *n = i;
    return LGI_OK;
}

lgi_status lgi_unet_maxpool_fwd(const void* in, int32_t n, int32_t h, int32_t w, int32_t c, void* out, void* stream) {
    if (!in || !out || h % 2 || w % 2 || c % 8 || !row_grid_fits(n, h / 2, w / 2, c / 8)) return LGI_ERR_INVALID_ARG;
    hipLaunchKernelGGL(k_pool_fwd, row_grid(n, h / 2, w / 2, c / 8), dim3(kBlock), 0, (hipStream_t)stream,
                       (const bf16*)in, n, h, w, c, (bf16*)out);
    return hipGetLastError() == hipSuccess ? LGI_OK : LGI_ERR_HIP;
}

lgi_status lgi_unet_maxpool_bwd(const void* in, const void* dout, int32_t n, int32_t h, int32_t w, int32_t c,
                                void* din, void* stream) {
    if (!in || !dout || !din || h % 2 || w % 2 || c % 8 || !row_grid_fits(n, h / 2, w / 2, c / 8))
        return LGI_ERR_INVALID_ARG;
    hipLaunchKernelGGL(k_pool_bwd, row_grid(n, h / 2, w / 2, c / 8), dim3(kBlock), 0, (hipStream_t)stream,
                       (const bf16*)in, (const bf16*)dout, n, h, w, c, (bf16*)din);
    return hipGetLastError() == hipSuccess ? LGI_OK : LGI_ERR_HIP;
}

lgi_status lgi_unet_upsample_fwd(const void* in, int32_t n, int32_t h, int32_t w, int32_t c, void* out, void* stream) {
    if (!in || !out || c % 8 || !row_grid_fits(n, 2 * (int64_t)h, 2 * (int64_t)w, c / 8)) return LGI_ERR_INVALID_ARG;
    hipLaunchKernelGGL(k_up_fwd, row_grid(n, h, w, c / 8), dim3(kBlock), 0, (hipStream_t)stream,
                       (const bf16*)in, n, h, w, c, (bf16*)out);
    return hipGetLastError() == hipSuccess ? LGI_OK : LGI_ERR_HIP;
}

lgi_status lgi_unet_upsample_bwd(const void* dout, int32_t n, int32_t h, int32_t w, int32_t c, void* din, void* stream) {
    if (!dout || !din || c % 8 || !row_grid_fits(n, h, w, c / 8)) return LGI_ERR_INVALID_ARG;
    hipLaunchKernelGGL(k_up_bwd, row_grid(n, h, w, c / 8), dim3(kBlock), 0, (hipStream_t)stream, (const bf16*)dout, n,
                       h, w, c, (bf16*)din);
    return hipGetLastError() == hipSuccess ? LGI_OK : LGI_ERR_HIP;
}

lgi_status lgi_unet_pack_input(const float* x, int64_t npx, void* out, void* stream) {
    if (!x || !out) return LGI_ERR_INVALID_ARG;
    hipLaunchKernelGGL(k_pack_input, dim3(cdiv(npx, kBlock)), dim3(kBlock), 0, (hipStream_t)stream, x, npx,
                       (bf16*)out);
    return hipGetLastError() == hipSuccess ? LGI_OK : LGI_ERR_HIP;
}

lgi_status lgi_unet_tail_fwd(const void* feat, const float* x, const float* w, const float* b, int64_t npx, float* out,
                             void* stream) {
    if (!feat || !x || !w || !b || !out || npx <= 0) return LGI_ERR_INVALID_ARG;
    hipLaunchKernelGGL(k_tail_fwd<false>, dim3(cdiv(npx, kBlock)), dim3(kBlock), 0, (hipStream_t)stream,
                       (const bf16*)feat, x, w, b, npx, out);
    return hipGetLastError() == hipSuccess ? LGI_OK : LGI_ERR_HIP;
}

lgi_status lgi_unet_tail_radiance_fwd(const void* feat, const float* hdr, const float* w, const float* b, int64_t npx,
                                      float* out, void* stream) {
    if (!feat || !hdr || !w || !b || !out || npx <= 0) return LGI_ERR_INVALID_ARG;
    hipLaunchKernelGGL(k_tail_fwd<true>, dim3(cdiv(npx, kBlock)), dim3(kBlock), 0, (hipStream_t)stream,
                       (const bf16*)feat, hdr, w, b, npx, out);
    return hipGetLastError() == hipSuccess ? LGI_OK : LGI_ERR_HIP;
}

lgi_status lgi_unet_tail_bwd(const void* feat, const float* dout, const float* w, int64_t npx, void* dfeat, float* dw,
                             float* db, float* scratch, int64_t scratch_floats, void* stream) {
    if (!feat || !dout || !w || !dfeat || !dw || !db || !scratch) return LGI_ERR_INVALID_ARG;
    const int blocks = (int)std::min<int64_t>(kTailBwdBlocks, cdiv(npx, kBlock));
    if ((int64_t)blocks * 99 > scratch_floats) return LGI_ERR_INVALID_ARG;
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(k_tail_bwd, dim3(blocks), dim3(kBlock), 0, st, (const bf16*)feat, dout, w, npx, (bf16*)dfeat,
                       scratch);
    hipLaunchKernelGGL(k_tail_reduce, dim3(2), dim3(kBlock), 0, st, scratch, blocks, dw, db);
    return hipGetLastError() == hipSuccess ? LGI_OK : LGI_ERR_HIP;
}

}  // extern "C"
