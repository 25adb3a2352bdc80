// Device-to-device copy with 16-byte vector loads and stores.
//
// Used by callers that keep operands resident (restaging a block without a
// host round trip) and as the known-size reference kernel when hardware
// byte counters are converted into HBM bytes.
#include "internal.h"

namespace jlc {
namespace {

typedef double d2 __attribute__((ext_vector_type(2)));

__global__ __launch_bounds__(256) void copy16(d2* __restrict__ dst, const d2* __restrict__ src, int64_t n16) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n16; i += stride) dst[i] = __builtin_nontemporal_load(&src[i]);
}

__global__ __launch_bounds__(256) void copy1(unsigned char* __restrict__ dst, const unsigned char* __restrict__ src,
                                             int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) dst[i] = src[i];
}

// out[0] = max(0, a[0..m)), out[1] = max(0, b[0..n)) with the comparison `v > cur` (NaN never taken), the
// order-independent maxima the split forms' count choosers need, so the host reads two floats instead of
// one statistic per row and column. One workgroup.
__global__ __launch_bounds__(1024) void max2_nonneg(const float* __restrict__ a, int64_t m, const float* __restrict__ b,
                                                   int64_t n, float* __restrict__ out) {
    __shared__ float red[2][16];
    float ma = 0.0f, mb = 0.0f;
    for (int64_t i = threadIdx.x; i < m; i += 1024) ma = a[i] > ma ? a[i] : ma;
    for (int64_t j = threadIdx.x; j < n; j += 1024) mb = b[j] > mb ? b[j] : mb;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const float oa = __shfl_xor(ma, off, 64), ob = __shfl_xor(mb, off, 64);
        ma = oa > ma ? oa : ma;
        mb = ob > mb ? ob : mb;
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[0][w] = ma;
        red[1][w] = mb;
    }
    __syncthreads();
    if (threadIdx.x < 2) {
        float v = 0.0f;
        for (int q = 0; q < 16; q++) v = red[threadIdx.x][q] > v ? red[threadIdx.x][q] : v;
        out[threadIdx.x] = v;
    }
}

}  // namespace

hipError_t launch_max2_nonneg(const float* a, int64_t m, const float* b, int64_t n, float* out, hipStream_t stream) {
    hipLaunchKernelGGL(max2_nonneg, dim3(1), dim3(1024), 0, stream, a, m, b, n, out);
    return hipGetLastError();
}

hipError_t launch_copy(void* dst, const void* src, size_t bytes, hipStream_t stream) {
    if (bytes == 0) return hipSuccess;
    const bool vec = ((reinterpret_cast<uintptr_t>(dst) | reinterpret_cast<uintptr_t>(src) | bytes) % 16) == 0;
    if (vec) {
        int64_t n16 = (int64_t)(bytes / 16);
        int64_t blocks = (n16 + 255) / 256;
        if (blocks > 256 * 64) blocks = 256 * 64;
        hipLaunchKernelGGL(copy16, dim3((unsigned)blocks), dim3(256), 0, stream, static_cast<d2*>(dst),
                           static_cast<const d2*>(src), n16);
    } else {
        int64_t blocks = ((int64_t)bytes + 255) / 256;
        if (blocks > 256 * 64) blocks = 256 * 64;
        hipLaunchKernelGGL(copy1, dim3((unsigned)blocks), dim3(256), 0, stream, static_cast<unsigned char*>(dst),
                           static_cast<const unsigned char*>(src), (int64_t)bytes);
    }
    return hipGetLastError();
}

}  // namespace jlc
