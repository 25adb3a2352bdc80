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

}  // namespace

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
