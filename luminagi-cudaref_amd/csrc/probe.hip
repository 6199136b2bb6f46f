// probe.hip — read-bandwidth probe of the memory hierarchy, for the tracer's roofline: the
// traversal's working set (tens of MB of nodes and vertices) is larger than one XCD's 4 MB L2
// and smaller than the 256 MB MALL, so its achieved rate is compared with what each level
// delivers on this chip, measured with the same kind of loads.
//
// Every thread reads 16-byte words i = t, t + T, t + 2T, ... (T = grid threads) of the buffer,
// `reps` times. Workgroups are dispatched round-robin over the XCDs, so each XCD re-reads the
// same 1/8 of the buffer on every pass: a buffer of up to 8 x 4 MB stays in the L2s, a buffer
// of 128 MB is served by the MALL and one of a few GB by HBM.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "lgi.h"

namespace {

constexpr int kBlock = 256;

__global__ __launch_bounds__(kBlock) void k_read_probe(const uint4* __restrict__ buf, uint64_t n, int reps,
                                                       uint32_t* out) {
    const uint64_t T = (uint64_t)gridDim.x * kBlock;
    const uint64_t t = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    uint32_t acc = 0;
    for (int r = 0; r < reps; ++r) {
        uint64_t i = t;
        // four independent loads in flight per thread
        for (; i + 3 * T < n; i += 4 * T) {
            const uint4 a = buf[i], b = buf[i + T], c = buf[i + 2 * T], d = buf[i + 3 * T];
            acc ^= a.x ^ a.y ^ a.z ^ a.w ^ b.x ^ b.y ^ b.z ^ b.w ^ c.x ^ c.y ^ c.z ^ c.w ^ d.x ^ d.y ^ d.z ^ d.w;
        }
        for (; i < n; i += T) {
            const uint4 a = buf[i];
            acc ^= a.x ^ a.y ^ a.z ^ a.w;
        }
        acc = acc * 0x9E3779B1u + (uint32_t)r;  // keeps the passes distinct to the compiler
    }
    if (acc == 0x7F4A7C15u) out[0] = acc;  // data-dependent store: the loads cannot be dropped
}

}  // namespace

extern "C" lgi_status lgi_probe_read_bw(const void* d_buf, uint64_t bytes, int32_t reps, int32_t blocks,
                                        uint32_t* d_sink, void* stream, float* ms) {
    if (!d_buf || !d_sink || bytes < 16 || reps <= 0 || blocks <= 0 || ((uintptr_t)d_buf & 15u))
        return LGI_ERR_INVALID_ARG;
    hipStream_t st = (hipStream_t)stream;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (ms) {
        if (hipEventCreate(&e0) != hipSuccess) return LGI_ERR_HIP;
        if (hipEventCreate(&e1) != hipSuccess) {
            (void)hipEventDestroy(e0);
            return LGI_ERR_HIP;
        }
        (void)hipEventRecord(e0, st);
    }
    hipLaunchKernelGGL(k_read_probe, dim3(blocks), dim3(kBlock), 0, st, (const uint4*)d_buf, bytes / 16, reps, d_sink);
    lgi_status rc = hipGetLastError() == hipSuccess ? LGI_OK : LGI_ERR_HIP;
    if (ms) {
        if (rc == LGI_OK && (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                             hipEventElapsedTime(ms, e0, e1) != hipSuccess))
            rc = LGI_ERR_HIP;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    }
    return rc;
}
