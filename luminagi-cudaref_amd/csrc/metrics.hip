// metrics.hip — image-wide FP64 reductions on the GPU: MSE (for PSNR), mean absolute
// difference and SSIM of two 8-bit RGB images, as tools/diff.py:25-29 and :100-101 compute
// them on the host (PSNR = 20 log10(255 / sqrt(MSE)); SSIM with 7x7 uniform windows, sample
// covariance, K1 = 0.01, K2 = 0.03, data range 255, mean over the interior where the window
// fits, averaged over channels). Deterministic: fixed grid, fixed-order two-stage sums.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "lgi.h"

namespace {

constexpr int kBlock = 256;
constexpr int kGrid = 1024;

__device__ double block_sum(double v, double* s) {
    s[threadIdx.x] = v;
    __syncthreads();
    for (int o = kBlock / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
        __syncthreads();
    }
    double r = s[0];
    __syncthreads();
    return r;
}

__global__ __launch_bounds__(kBlock) void k_err(const uint8_t* a, const uint8_t* b, int64_t n, double* partial) {
    __shared__ double s[kBlock];
    double sq = 0.0, ab = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
        const double d = (double)a[i] - (double)b[i];
        sq += d * d;
        ab += fabs(d);
    }
    sq = block_sum(sq, s);
    ab = block_sum(ab, s);
    if (threadIdx.x == 0) {
        partial[2 * blockIdx.x] = sq;
        partial[2 * blockIdx.x + 1] = ab;
    }
}

// SSIM map value at interior pixel (x, y), channel c (window fully inside the image).
__device__ double ssim_at(const uint8_t* a, const uint8_t* b, int W, int x, int y, int c) {
    double sa = 0, sb = 0, saa = 0, sbb = 0, sab = 0;
    for (int dy = -3; dy <= 3; ++dy)
        for (int dx = -3; dx <= 3; ++dx) {
            const int64_t i = ((int64_t)(y + dy) * W + (x + dx)) * 3 + c;
            const double va = a[i], vb = b[i];
            sa += va; sb += vb; saa += va * va; sbb += vb * vb; sab += va * vb;
        }
    const double np = 49.0, cov_norm = np / (np - 1.0);
    const double ux = sa / np, uy = sb / np;
    const double vx = cov_norm * (saa / np - ux * ux);
    const double vy = cov_norm * (sbb / np - uy * uy);
    const double vxy = cov_norm * (sab / np - ux * uy);
    const double C1 = (0.01 * 255.0) * (0.01 * 255.0), C2 = (0.03 * 255.0) * (0.03 * 255.0);
    return ((2.0 * ux * uy + C1) * (2.0 * vxy + C2)) / ((ux * ux + uy * uy + C1) * (vx + vy + C2));
}

__global__ __launch_bounds__(kBlock) void k_ssim(const uint8_t* a, const uint8_t* b, int H, int W, double* partial) {
    __shared__ double s[kBlock];
    const int iw = W - 6, ih = H - 6;
    const int64_t n = (int64_t)iw * ih * 3;
    double acc = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
        const int c = (int)(i % 3);
        const int64_t p = i / 3;
        acc += ssim_at(a, b, W, (int)(p % iw) + 3, (int)(p / iw) + 3, c);
    }
    acc = block_sum(acc, s);
    if (threadIdx.x == 0) partial[blockIdx.x] = acc;
}

__global__ void k_final(const double* perr, const double* pssim, int nb, int64_t n, int64_t nssim, double* out) {
    if (threadIdx.x != 0) return;
    double sq = 0, ab = 0, ss = 0;
    for (int i = 0; i < nb; ++i) {
        sq += perr[2 * i];
        ab += perr[2 * i + 1];
        ss += pssim[i];
    }
    out[0] = sq / (double)n;
    out[1] = nssim > 0 ? ss / (double)nssim : 1.0;
    out[2] = ab / (double)n;
}

// Heat colour table held in LDS; one pixel per thread.
struct HeatLut {
    uint8_t rgb[768];
};

__global__ __launch_bounds__(kBlock) void k_heatmap(const uint8_t* a, const uint8_t* b, int64_t npx, double gain,
                                                    HeatLut lut, uint8_t* out) {
    __shared__ uint8_t t[768];
    for (int i = threadIdx.x; i < 768; i += kBlock) t[i] = lut.rgb[i];
    __syncthreads();
    for (int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x; p < npx; p += (int64_t)gridDim.x * kBlock) {
        const uint8_t* pa = a + 3 * p;
        const uint8_t* pb = b + 3 * p;
        const int s = abs((int)pa[0] - (int)pb[0]) + abs((int)pa[1] - (int)pb[1]) + abs((int)pa[2] - (int)pb[2]);
        // numpy: mean over channels in float64 (sum / 3), times gain, clip to [0, 255], truncate
        double v = (double)s / 3.0 * gain;
        v = v > 0.0 ? (v < 255.0 ? v : 255.0) : 0.0;
        const int k = 3 * (int)v;
        out[3 * p] = t[k];
        out[3 * p + 1] = t[k + 1];
        out[3 * p + 2] = t[k + 2];
    }
}

}  // namespace

/* The heatmap's colour table (tools/diff.py:32-38): for x = i * (1/255) (x = 1 at i = 255, as
 * numpy.linspace(0, 1, 256) spaces it), three quadratics clipped to [0, 1], times 255,
 * truncated. Evaluated left to right in double precision as numpy does, no contraction. */
extern "C" lgi_status lgi_heat_lut(uint8_t* rgb768) {
    if (!rgb768) return LGI_ERR_INVALID_ARG;
    const double step = 1.0 / 255.0;
    auto chan = [](double v) {
        v = v < 0.0 ? 0.0 : (v > 1.0 ? 1.0 : v);
        return (uint8_t)(v * 255.0);
    };
    for (int i = 0; i < 256; ++i) {
        const double x = i == 255 ? 1.0 : (double)i * step;
        const double x2 = x * x;
        rgb768[3 * i] = chan(-0.2 + 2.4 * x - 1.3 * x2);
        rgb768[3 * i + 1] = chan(0.1 + 0.8 * x + 0.1 * x2);
        rgb768[3 * i + 2] = chan(0.4 + 0.8 * x - 1.1 * x2);
    }
    return LGI_OK;
}

extern "C" lgi_status lgi_diff_heatmap(const uint8_t* d_a, const uint8_t* d_b, int32_t h, int32_t w, double gain,
                                       uint8_t* d_out, void* stream) {
    if (!d_a || !d_b || !d_out || h <= 0 || w <= 0 || !(gain == gain)) return LGI_ERR_INVALID_ARG;
    HeatLut lut;
    lgi_heat_lut(lut.rgb);
    const int64_t npx = (int64_t)h * w;
    const int64_t blocks = (npx + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_heatmap, dim3((unsigned)(blocks < 4096 ? blocks : 4096)), dim3(kBlock), 0,
                       (hipStream_t)stream, d_a, d_b, npx, gain, lut, d_out);
    return hipGetLastError() == hipSuccess ? LGI_OK : LGI_ERR_HIP;
}

extern "C" lgi_status lgi_image_metrics(const uint8_t* d_a, const uint8_t* d_b, int32_t h, int32_t w, double* d_scratch,
                                        double* d_out, void* stream) {
    if (!d_a || !d_b || !d_scratch || !d_out || h < 7 || w < 7) return LGI_ERR_INVALID_ARG;
    hipStream_t st = (hipStream_t)stream;
    const int64_t n = (int64_t)h * w * 3;
    hipLaunchKernelGGL(k_err, dim3(kGrid), dim3(kBlock), 0, st, d_a, d_b, n, d_scratch);
    hipLaunchKernelGGL(k_ssim, dim3(kGrid), dim3(kBlock), 0, st, d_a, d_b, h, w, d_scratch + 2 * kGrid);
    hipLaunchKernelGGL(k_final, dim3(1), dim3(64), 0, st, d_scratch, d_scratch + 2 * kGrid, kGrid, n,
                       (int64_t)(w - 6) * (h - 6) * 3, d_out);
    return hipGetLastError() == hipSuccess ? LGI_OK : LGI_ERR_HIP;
}
