// libjlc_hip C ABI: device detection, per-thread streams and staging
// buffers, argument validation and the host-pointer / device-pointer entry
// points declared in include/jlc_hip.h.
//
// Threading model. The reference calls its GEMM from many threads at once
// (ThreadLocal workspaces, GemmWorkspace.java:8; ForkJoin tasks,
// OptimizedBLAS3.java:314-345). Each calling host thread here owns a HIP
// stream, a pair of timing events and grow-only device staging buffers, so
// no lock is held while work runs; a process-wide mutex guards only device
// selection at initialisation.
#include <hip/hip_runtime.h>

#include <dlfcn.h>

#include <atomic>
#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <cmath>
#include <string>
#include <algorithm>
#include <thread>
#include <vector>

#include "../../include/jlc_hip.h"
#include "internal.h"

namespace {

std::mutex g_mutex;
std::atomic<int> g_device{-1};
std::atomic<int> g_cus{256};
std::atomic<bool> g_alive{true};

thread_local std::string t_error;

// Scratch of the no-allocation entry point (jlc_dgemm_dev_noalloc): the split-K partial products of
// the FP64 MFMA kernel. Every device the library runs on holds its own copy for the life of the
// process; host code never uses the array's name, only the device address jlc_hip_init resolves per
// device with hipGetSymbolAddress (g_scratch_ptr). 256 MiB holds every split-K call the shipped
// tuning (4 tiles per CU, 256 CUs, at most 64 slices) can make: slices * m * n <= ceil(1024 / t) * t *
// 128^2 doubles for t output tiles, at most 1533 * 2^14 doubles (201 MB) at t = 511.
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)256 << 20;
constexpr int MAX_DEVICES = 64;
static __device__ __align__(256) double scratch[NOALLOC_SCRATCH_BYTES / sizeof(double)];
std::atomic<void*> g_scratch_ptr[MAX_DEVICES];

std::mutex g_tune_mutex;
jlc::TuneTable g_tune;
std::once_flag g_tune_autoload;

int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    t_error = buf;
    return code;
}

int hip_fail(hipError_t e, const char* what) {
    return fail(JLC_ERR_HIP, "%s failed: %s", what, hipGetErrorString(e));
}

#define JLC_HIP_CHECK(expr)                                    \
    do {                                                       \
        hipError_t e_ = (expr);                                \
        if (e_ != hipSuccess) return hip_fail(e_, #expr);      \
    } while (0)

bool env_flag(const char* name, bool dflt) {
    const char* v = std::getenv(name);
    if (!v || !*v) return dflt;
    if (!std::strcmp(v, "0") || !strcasecmp(v, "false") || !strcasecmp(v, "off") || !strcasecmp(v, "no"))
        return false;
    return true;
}

struct ThreadCtx {
    int device = -1;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr, ev3 = nullptr;
    // Staging: 0 = A, 1 = B, 2 = C; split workspace: 3 = exponents/spreads, 4 = A slices, 5 = B slices;
    // 6 = split-K partial products; 7 = residue planes of C (CRT form); 8, 9 = the W and W2 products of a
    // compact-WY update.
    static constexpr int NBUF = 10;
    void* buf[NBUF] = {};
    size_t cap[NBUF] = {};
    // Pinned host staging for the host-pointer entry points.
    void* pin = nullptr;
    size_t pin_cap = 0;
    // The split-form, split-K and WY workspace (buf[3..9]) is shared by every call of this thread,
    // whatever stream the call runs on: ws_ev marks the end of the last use, on ws_stream, and a call
    // on another stream waits for it before touching the buffers. A use on the thread's own stream is
    // marked lazily (ws_pending), by the first later call on another stream: each event record costs a
    // gap of several microseconds between the kernels around it, which a stream of calls on one stream
    // never needs.
    hipEvent_t ws_ev = nullptr;
    hipStream_t ws_stream = nullptr;
    bool ws_pending = false;
    // Blocked QR: pinned host copy of a CholeskyQR2 panel's flag and the event that marks its arrival.
    int* qr_flag = nullptr;
    hipEvent_t qr_ev = nullptr;
    // Split forms' count choosers: the two statistic maxima, on the device and their pinned host copy.
    float* choose_dev = nullptr;
    float* choose_pin = nullptr;

    ~ThreadCtx() {
        if (!g_alive.load() || device < 0) return;
        if (pin) (void)hipHostFree(pin);
        if (qr_flag) (void)hipHostFree(qr_flag);
        if (choose_pin) (void)hipHostFree(choose_pin);
        if (choose_dev) (void)hipFree(choose_dev);
        if (qr_ev) (void)hipEventDestroy(qr_ev);
        for (int i = 0; i < NBUF; i++)
            if (buf[i]) (void)hipFree(buf[i]);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (ev2) (void)hipEventDestroy(ev2);
        if (ev3) (void)hipEventDestroy(ev3);
        if (ws_ev) (void)hipEventDestroy(ws_ev);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

thread_local ThreadCtx t_ctx;

void mark_dead() { g_alive.store(false); }

int check_arch(int device, char* name, size_t name_len) {
    hipDeviceProp_t prop;
    JLC_HIP_CHECK(hipGetDeviceProperties(&prop, device));
    if (name && name_len) {
        std::snprintf(name, name_len, "%s (%s)", prop.name, prop.gcnArchName);
    }
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0 && !env_flag("JLC_HIP_FORCE", false))
        return fail(JLC_ERR_NO_DEVICE, "device %d is %s, this library is built for gfx950", device,
                    prop.gcnArchName);
    return JLC_OK;
}

int ensure_init() {
    if (!env_flag("JLC_HIP_ENABLED", true)) return fail(JLC_ERR_DISABLED, "JLC_HIP_ENABLED disables the HIP path");
    if (g_device.load() < 0) {
        int rc = jlc_hip_init(0);
        if (rc != JLC_OK) return rc;
    }
    ThreadCtx& c = t_ctx;
    int dev = g_device.load();
    if (c.device != dev) {
        JLC_HIP_CHECK(hipSetDevice(dev));
        JLC_HIP_CHECK(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
        JLC_HIP_CHECK(hipEventCreate(&c.ev0));
        JLC_HIP_CHECK(hipEventCreate(&c.ev1));
        JLC_HIP_CHECK(hipEventCreate(&c.ev2));
        JLC_HIP_CHECK(hipEventCreate(&c.ev3));
        JLC_HIP_CHECK(hipEventCreateWithFlags(&c.ws_ev, hipEventDisableTiming));
        c.device = dev;
    } else {
        JLC_HIP_CHECK(hipSetDevice(dev));
    }
    return JLC_OK;
}

// Orders work on stream s after every earlier launch of this thread that used the shared
// workspace on another stream (a no-op when the last user ran on s: a stream is ordered).
int ws_acquire(hipStream_t s) {
    ThreadCtx& c = t_ctx;
    if (c.ws_stream && c.ws_stream != s) {
        if (c.ws_pending) {
            // the last use ran on the thread's own stream (which lives as long as the thread): mark it now
            JLC_HIP_CHECK(hipEventRecord(c.ws_ev, c.ws_stream));
            c.ws_pending = false;
        }
        JLC_HIP_CHECK(hipStreamWaitEvent(s, c.ws_ev, 0));
    }
    return JLC_OK;
}

// Marks the end of this call's workspace use on stream s: at once on a caller's stream (it may be
// destroyed before the next call), lazily on the thread's own stream.
int ws_release(hipStream_t s) {
    ThreadCtx& c = t_ctx;
    if (s == c.stream) {
        c.ws_pending = true;
    } else {
        JLC_HIP_CHECK(hipEventRecord(c.ws_ev, s));
        c.ws_pending = false;
    }
    c.ws_stream = s;
    return JLC_OK;
}

// Scope of one call's workspace use: acquire() before the first launch that touches buf[3..7],
// the release recorded when the call's launches have all been queued.
struct WorkspaceUse {
    hipStream_t s;
    bool used = false;
    explicit WorkspaceUse(hipStream_t s_) : s(s_) {}
    int acquire() {
        if (used) return JLC_OK;
        used = true;
        return ws_acquire(s);
    }
    ~WorkspaceUse() {
        if (used) (void)ws_release(s);
    }
};

int ensure_buf(int slot, size_t bytes) {
    ThreadCtx& c = t_ctx;
    if (bytes <= c.cap[slot]) return JLC_OK;
    if (c.buf[slot]) {
        JLC_HIP_CHECK(hipFree(c.buf[slot]));
        c.buf[slot] = nullptr;
        c.cap[slot] = 0;
    }
    size_t want = bytes < 256 ? 256 : bytes;
    hipError_t e = hipMalloc(&c.buf[slot], want);
    if (e != hipSuccess) {
        c.buf[slot] = nullptr;
        return fail(JLC_ERR_ALLOC, "hipMalloc(%zu) failed: %s", want, hipGetErrorString(e));
    }
    c.cap[slot] = want;
    return JLC_OK;
}

int ensure_pin(size_t bytes) {
    ThreadCtx& c = t_ctx;
    if (bytes <= c.pin_cap) return JLC_OK;
    if (c.pin) {
        JLC_HIP_CHECK(hipHostFree(c.pin));
        c.pin = nullptr;
        c.pin_cap = 0;
    }
    size_t want = bytes < 4096 ? 4096 : bytes;
    hipError_t e = hipHostMalloc(&c.pin, want, hipHostMallocDefault);
    if (e != hipSuccess) {
        c.pin = nullptr;
        return fail(JLC_ERR_ALLOC, "hipHostMalloc(%zu) failed: %s", want, hipGetErrorString(e));
    }
    c.pin_cap = want;
    return JLC_OK;
}

// Packs a rows x cols block (row pitch spitch elements) into pinned memory
// and queues one copy to the device block (row pitch dpitch elements).
// Host copies of pitched row blocks, split over up to 8 threads above 16 MiB (one thread of a
// host core moves ~10 GB/s, well below what the copy engines take over PCIe).
void copy_rows(double* dst, int64_t dpitch, const double* src, int64_t spitch, int64_t rows, int64_t cols) {
    auto run = [=](int64_t r0, int64_t r1) {
        if (dpitch == cols && spitch == cols) {
            std::memcpy(dst + r0 * cols, src + r0 * cols, (size_t)((r1 - r0) * cols) * sizeof(double));
        } else {
            for (int64_t r = r0; r < r1; r++)
                std::memcpy(dst + r * dpitch, src + r * spitch, (size_t)cols * sizeof(double));
        }
    };
    const size_t bytes = (size_t)(rows * cols) * sizeof(double);
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const int64_t nt = std::min<int64_t>({8, (int64_t)hw, rows, (int64_t)(bytes >> 24)});
    if (nt <= 1) {
        run(0, rows);
        return;
    }
    std::vector<std::thread> pool;
    for (int64_t t = 1; t < nt; t++) pool.emplace_back(run, rows * t / nt, rows * (t + 1) / nt);
    run(0, rows / nt);
    for (auto& th : pool) th.join();
}

// Rows x cols of a pitched host array to the device through the pinned region `pin` (rows * cols
// doubles, contiguous): in blocks of ~128 MiB, each block is copied into its own part of the pinned
// region and its DMA queued at once, so the copy engine moves block i while the host fills block i+1.
int stage_to_device(double* dst, int64_t dpitch, const double* src, int64_t spitch, int64_t rows, int64_t cols,
                    double* pin, hipStream_t s) {
    if (rows <= 0 || cols <= 0) return JLC_OK;
    const int64_t block = std::max<int64_t>(1, ((int64_t)128 << 20) / (int64_t)sizeof(double) / cols);
    for (int64_t r0 = 0; r0 < rows; r0 += block) {
        const int64_t nr = std::min(block, rows - r0);
        double* p = pin + r0 * cols;
        copy_rows(p, cols, src + r0 * spitch, spitch, nr, cols);
        if (dpitch == cols)
            JLC_HIP_CHECK(hipMemcpyAsync(dst + r0 * dpitch, p, (size_t)(nr * cols) * sizeof(double),
                                         hipMemcpyHostToDevice, s));
        else
            JLC_HIP_CHECK(hipMemcpy2DAsync(dst + r0 * dpitch, dpitch * sizeof(double), p, cols * sizeof(double),
                                           cols * sizeof(double), nr, hipMemcpyHostToDevice, s));
    }
    return JLC_OK;
}

// Rows x cols of a contiguous device array back into a pitched host array through `pin`: the block
// DMAs are queued together, each followed by an event, and the host copies block i out as soon as its
// event completes while the copy engine moves the next ones.
int stage_from_device(double* dst, int64_t dpitch, const double* src, int64_t rows, int64_t cols, double* pin,
                      hipStream_t s) {
    if (rows <= 0 || cols <= 0) return JLC_OK;
    const int64_t block = std::max<int64_t>(1, ((int64_t)128 << 20) / (int64_t)sizeof(double) / cols);
    const int64_t nblk = (rows + block - 1) / block;
    std::vector<hipEvent_t> done((size_t)nblk, nullptr);
    int rc = JLC_OK;
    for (int64_t i = 0; i < nblk && rc == JLC_OK; i++) {
        const int64_t r0 = i * block, nr = std::min(block, rows - r0);
        hipError_t e = hipEventCreateWithFlags(&done[(size_t)i], hipEventDisableTiming);
        if (e == hipSuccess)
            e = hipMemcpyAsync(pin + r0 * cols, src + r0 * cols, (size_t)(nr * cols) * sizeof(double),
                               hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipEventRecord(done[(size_t)i], s);
        if (e != hipSuccess) rc = fail(JLC_ERR_HIP, "device-to-host staging failed: %s", hipGetErrorString(e));
    }
    for (int64_t i = 0; i < nblk && rc == JLC_OK; i++) {
        const int64_t r0 = i * block, nr = std::min(block, rows - r0);
        hipError_t e = hipEventSynchronize(done[(size_t)i]);
        if (e != hipSuccess) {
            rc = fail(JLC_ERR_HIP, "device-to-host staging failed: %s", hipGetErrorString(e));
            break;
        }
        copy_rows(dst + r0 * dpitch, dpitch, pin + r0 * cols, cols, nr, cols);
    }
    if (rc != JLC_OK) (void)hipStreamSynchronize(s);  // nothing may still write the pinned region
    for (hipEvent_t ev : done)
        if (ev) (void)hipEventDestroy(ev);
    return rc;
}

double now_ms() {
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

// Validates a GEMM request (shapes, leading dimensions, layouts, pointers).
int validate_gemm(int la, int lb, int64_t m, int64_t n, int64_t k, double alpha, const void* a, int64_t lda,
                  const void* b, int64_t ldb, const void* c, int64_t ldc, int mode) {
    if (la != JLC_ROW && la != JLC_COL) return fail(JLC_ERR_INVALID_ARG, "layout_a must be JLC_ROW or JLC_COL");
    if (lb != JLC_ROW && lb != JLC_COL) return fail(JLC_ERR_INVALID_ARG, "layout_b must be JLC_ROW or JLC_COL");
    if (m < 0 || n < 0 || k < 0) return fail(JLC_ERR_INVALID_ARG, "negative dimension m=%lld n=%lld k=%lld",
                                             (long long)m, (long long)n, (long long)k);
    if (m > INT32_MAX || n > INT32_MAX || k > INT32_MAX)
        return fail(JLC_ERR_INVALID_ARG, "dimension above 2^31-1");
    int form = mode & 0xff;
    if (form != JLC_MODE_FP64 && form != JLC_MODE_SPLIT_INT8 && form != JLC_MODE_SPLIT_BF16 &&
        form != JLC_MODE_SPLIT_CRT)
        return fail(JLC_ERR_INVALID_ARG, "unknown mode %d", mode);
    int slices = (mode >> 8) & 0xff;
    if (form == JLC_MODE_FP64 && slices != 0) return fail(JLC_ERR_INVALID_ARG, "slice count given for FP64 mode");
    if (form == JLC_MODE_SPLIT_INT8 && slices != 0 &&
        (slices < jlc::SPLIT_MIN_SLICES || slices > jlc::SPLIT_MAX_SLICES))
        return fail(JLC_ERR_INVALID_ARG, "INT8 slice count %d outside [%d, %d]", slices, jlc::SPLIT_MIN_SLICES,
                    jlc::SPLIT_MAX_SLICES);
    if (form == JLC_MODE_SPLIT_BF16 && slices != 0 &&
        (slices < jlc::SPLIT_MIN_SLICES || slices > jlc::SPLIT_BF16_MAX_SLICES))
        return fail(JLC_ERR_INVALID_ARG, "BF16 slice count %d outside [%d, %d]", slices, jlc::SPLIT_MIN_SLICES,
                    jlc::SPLIT_BF16_MAX_SLICES);
    if (form == JLC_MODE_SPLIT_CRT && slices != 0 &&
        (slices < jlc::CRT_MIN_MODULI || slices > jlc::CRT_MAX_MODULI))
        return fail(JLC_ERR_INVALID_ARG, "CRT moduli count %d outside [%d, %d]", slices, jlc::CRT_MIN_MODULI,
                    jlc::CRT_MAX_MODULI);
    if (m == 0 || n == 0) return JLC_OK;
    if (!c) return fail(JLC_ERR_INVALID_ARG, "C must not be null");
    if (ldc < n) return fail(JLC_ERR_INVALID_ARG, "ldc=%lld < n=%lld", (long long)ldc, (long long)n);
    if (k > 0 && alpha != 0.0) {
        if (!a || !b) return fail(JLC_ERR_INVALID_ARG, "A and B must not be null");
        int64_t need_a = la == JLC_ROW ? k : m;
        int64_t need_b = lb == JLC_ROW ? n : k;
        if (lda < need_a) return fail(JLC_ERR_INVALID_ARG, "lda=%lld < %lld", (long long)lda, (long long)need_a);
        if (ldb < need_b) return fail(JLC_ERR_INVALID_ARG, "ldb=%lld < %lld", (long long)ldb, (long long)need_b);
    }
    return JLC_OK;
}

// Slice count for the INT8 split form from the inputs' exponent spread. The
// truncation error of S 7-bit slices is about 2^(ea+eb-7S) per term while an
// FP64 product-sum errs by about 2^-53 of the terms' typical size; spread_a /
// spread_b are the largest row / column gaps (in bits) between the scale 2^e
// and the typical |x| (geometric mean), so matching FP64 needs
// 7S >= 53 + spread_a + spread_b - 4.
// Slice count from the largest row spread of op(A) (sa) and column spread of op(B) (sb), both >= 0.
int choose_slices(float sa, float sb, int digit_bits, int max_slices) {
    double bits = 53.0 + sa + sb - 4.0;
    int s = (int)std::ceil(bits / digit_bits);
    if (s < jlc::SPLIT_MIN_SLICES) s = jlc::SPLIT_MIN_SLICES;
    return s > max_slices ? 0 : s;  // 0: more slices than the form has
}

int ensure_buf(int slot, size_t bytes);

// max(0, a[0..m)) and max(0, b[0..n)) of device statistics (v > cur, so NaN is never taken) for the count
// choosers: reduced on the device and read back as 8 bytes of pinned memory, so the host waits for one
// small copy rather than copying and scanning one statistic per row and column.
int device_max2(const float* a, int64_t m, const float* b, int64_t n, hipStream_t s, float* ma, float* mb) {
    ThreadCtx& ctx = t_ctx;
    if (!ctx.choose_dev) JLC_HIP_CHECK(hipMalloc((void**)&ctx.choose_dev, 256));
    if (!ctx.choose_pin) JLC_HIP_CHECK(hipHostMalloc((void**)&ctx.choose_pin, 2 * sizeof(float), hipHostMallocDefault));
    JLC_HIP_CHECK(jlc::launch_max2_nonneg(a, m, b, n, ctx.choose_dev, s));
    JLC_HIP_CHECK(hipMemcpyAsync(ctx.choose_pin, ctx.choose_dev, 2 * sizeof(float), hipMemcpyDeviceToHost, s));
    JLC_HIP_CHECK(hipStreamSynchronize(s));
    *ma = ctx.choose_pin[0];
    *mb = ctx.choose_pin[1];
    return JLC_OK;
}

// Runs a slice form; in automatic mode *fallback is set (nothing launched
// beyond the statistics) when the spread needs more slices than the form has.
int run_split(bool bf16, int la, int lb, int64_t m, int64_t n, int64_t k, double alpha, const double* a,
              int64_t lda, const double* b, int64_t ldb, double beta, double* c, int64_t ldc, int slices,
              hipStream_t stream, jlc::TileInfo* info, bool* fallback) {
    *fallback = false;
    ThreadCtx& ctx = t_ctx;
    jlc::SplitPlan p;
    const int max_slices = bf16 ? jlc::SPLIT_BF16_MAX_SLICES : jlc::SPLIT_MAX_SLICES;
    auto sizes = [&](int s_) {
        if (bf16) jlc::split_bf16_plan_sizes(m, n, k, s_, &p);
        else jlc::split_plan_sizes(m, n, k, s_, &p);
    };
    sizes(slices ? slices : max_slices);
    const size_t stat_bytes = ((size_t)(p.mpad + p.npad) * (sizeof(int) + sizeof(float)) + 255) / 256 * 256;
    const size_t part_bytes = (size_t)jlc::split_stats_part_elems(m, n, k) * 4 * sizeof(double);
    int rc = ensure_buf(3, stat_bytes + part_bytes);
    if (rc != JLC_OK) return rc;
    char* sb = static_cast<char*>(ctx.buf[3]);
    p.ea = reinterpret_cast<int*>(sb);
    p.eb = p.ea + p.mpad;
    p.spread_a = reinterpret_cast<float*>(p.eb + p.npad);
    p.spread_b = p.spread_a + p.mpad;
    p.stats_part = reinterpret_cast<double*>(sb + stat_bytes);
    JLC_HIP_CHECK(jlc::launch_split_stats(la, lb, a, lda, b, ldb, &p, stream));
    if (slices == 0) {
        float sa = 0.0f, sbm = 0.0f;
        if ((rc = device_max2(p.spread_a, m, p.spread_b, n, stream, &sa, &sbm)) != JLC_OK) return rc;
        slices = choose_slices(sa, sbm, bf16 ? 8 : 7, max_slices);
        if (slices == 0) {
            *fallback = true;
            return JLC_OK;
        }
        p.slices = slices;
    }
    if ((rc = ensure_buf(4, (size_t)(p.slices * p.a_slice_bytes))) != JLC_OK) return rc;
    if ((rc = ensure_buf(5, (size_t)(p.slices * p.b_slice_bytes))) != JLC_OK) return rc;
    p.a_pack = ctx.buf[4];
    p.b_pack = ctx.buf[5];
    info->slices = p.slices;
    if (bf16)
        JLC_HIP_CHECK(jlc::launch_split_bf16(la, lb, a, lda, b, ldb, p, alpha, beta, c, ldc, stream, info));
    else
        JLC_HIP_CHECK(jlc::launch_split_pack_and_gemm(la, lb, a, lda, b, ldb, p, alpha, beta, c, ldc, stream, info));
    return JLC_OK;
}

// Whether nmod moduli hold every C' of the CRT form with norm bits (na, nb):
// |C'_ij| <= ||a'_i|| ||b'_j|| <= (2^na + sqrt(k)/2)(2^nb + sqrt(k)/2) must
// stay below M/4 (the reconstruction's quotient estimate needs the margin).
bool crt_fits(int na, int nb, int64_t k, int nmod) {
    const double h = 0.5 * std::sqrt((double)k);
    const double need = std::log2(std::ldexp(1.0, na) + h) + std::log2(std::ldexp(1.0, nb) + h) + 2.0;
    return need <= jlc::crt_log2_product(nmod) - 1e-6;
}

// Explicit moduli count: the largest even split of the norm bits that fits.
bool crt_explicit_bits(int nmod, int64_t k, int* na, int* nb) {
    for (int total = 2 * jlc::CRT_MAX_BITS; total >= 4; total--) {
        *na = total / 2;
        *nb = total - *na;
        if (crt_fits(*na, *nb, k, nmod)) return true;
    }
    return false;
}

// Norm bits per operand and moduli count of the CRT form. Row i is scaled so
// that ||a'_i|| ~ 2^na; its rounding error relative to the row norm is then
// ~2^-na and sqrt(k) * 2^-na relative to an rms entry. The FP64 product the
// form is matched against errs more on wide (heavy-tailed) rows too, so the
// bits grow slowly with sig = log2(rms / geometric mean):
//   na = ceil(offset + log2(k)/2 + slope * max_i sig_i),
// with offset 49.4 and slope 0.43 set from the benchmark distributions
// (a = (u - 0.5) exp(phi g), k = 16384; max sig ~1.1, ~10 and ~23 at phi =
// 0.5, 2 and 4): the max relative and componentwise errors stay below the
// FP64 kernel's at phi = 0.5 (57 bits, 15 moduli) and phi = 1 and 2 (16
// moduli), and phi = 4 exceeds the 65-bit limit and runs on the FP64 kernel
// (62 bits had given 30x its componentwise error there).
int crt_choose(double sig_a, double sig_b, int64_t k, int* na, int* nb) {
    const jlc::TuneTable tt = jlc::tune_snapshot();
    const double base = tt.crt_bits_offset + 0.5 * std::log2((double)std::max<int64_t>(k, 1));
    *na = std::max(2, (int)std::ceil(base + tt.crt_bits_slope * sig_a));
    *nb = std::max(2, (int)std::ceil(base + tt.crt_bits_slope * sig_b));
    if (*na > jlc::CRT_MAX_BITS || *nb > jlc::CRT_MAX_BITS) return 0;
    for (int nm = jlc::CRT_MIN_MODULI; nm <= jlc::CRT_MAX_MODULI; nm++) {
        if (!crt_fits(*na, *nb, k, nm)) continue;
        // The products cost the same whatever bits the moduli hold: spend what is left of their
        // capacity on more bits, the narrower operand first.
        for (;;) {
            int* lo = *na <= *nb ? na : nb;
            if (*lo >= jlc::CRT_MAX_BITS) break;
            ++*lo;
            if (!crt_fits(*na, *nb, k, nm)) {
                --*lo;
                break;
            }
        }
        return nm;
    }
    return 0;
}

// Runs the CRT form; *fallback is set (and nothing launched beyond the
// statistics) when automatic mode finds the spread too wide for 20 moduli.
int run_crt(int la, int lb, int64_t m, int64_t n, int64_t k, double alpha, const double* a, int64_t lda,
            const double* b, int64_t ldb, double beta, double* c, int64_t ldc, int nmod, hipStream_t stream,
            jlc::TileInfo* info, bool* fallback) {
    ThreadCtx& ctx = t_ctx;
    *fallback = false;
    if (k >= (int64_t)1 << 22) {  // the norm sums hold k < 2^22 terms of at most 2^40 + 1 units
        if (nmod != 0) return fail(JLC_ERR_UNSUPPORTED, "CRT form: k = %lld exceeds 2^22 - 1", (long long)k);
        *fallback = true;
        return JLC_OK;
    }
    jlc::CrtPlan p;
    jlc::crt_plan_sizes(m, n, k, jlc::CRT_MAX_MODULI, &p);
    jlc::SplitPlan sp;
    sp.m = m;
    sp.n = n;
    sp.k = k;
    sp.mpad = p.mpad;
    sp.npad = p.npad;
    // ea, eb, spread_a, spread_b, then the norm exponents ta, tb and chooser statistics sig_a, sig_b
    const size_t stat_bytes = ((size_t)(p.mpad + p.npad) * 2 * (sizeof(int) + sizeof(float)) + 255) / 256 * 256;
    const size_t part_bytes = (size_t)jlc::split_stats_part_elems(m, n, k) * 4 * sizeof(double);
    int rc = ensure_buf(3, stat_bytes + part_bytes);
    if (rc != JLC_OK) return rc;
    char* sb = static_cast<char*>(ctx.buf[3]);
    sp.ea = reinterpret_cast<int*>(sb);
    sp.eb = sp.ea + p.mpad;
    sp.spread_a = reinterpret_cast<float*>(sp.eb + p.npad);
    sp.spread_b = sp.spread_a + p.mpad;
    int* ta = reinterpret_cast<int*>(sp.spread_b + p.npad);
    int* tb = ta + p.mpad;
    float* sig_a = reinterpret_cast<float*>(tb + p.npad);
    float* sig_b = sig_a + p.mpad;
    sp.stats_part = reinterpret_cast<double*>(sb + stat_bytes);
    JLC_HIP_CHECK(jlc::launch_crt_stats_norms(la, lb, a, lda, b, ldb, sp, ta, tb, sig_a, sig_b, stream));
    int pa, pb;
    if (nmod == 0) {
        float sa = 0.0f, sbm = 0.0f;
        if ((rc = device_max2(sig_a, m, sig_b, n, stream, &sa, &sbm)) != JLC_OK) return rc;
        nmod = crt_choose(sa, sbm, k, &pa, &pb);
        if (nmod == 0) {
            *fallback = true;
            return JLC_OK;
        }
    } else if (!crt_explicit_bits(nmod, k, &pa, &pb)) {
        return fail(JLC_ERR_INVALID_ARG, "%d moduli leave no integer bits for k=%lld", nmod, (long long)k);
    }
    jlc::crt_plan_sizes(m, n, k, nmod, &p);
    // Launch geometry is 32-bit: 1 KiB fragment blocks per plane of either operand and of C.
    if ((p.mpad / 32) * (p.kpad / 32) > INT32_MAX || (p.npad / 32) * (p.kpad / 32) > INT32_MAX ||
        (p.mpad / 32) * (p.npad / 32) * 64 > INT32_MAX)
        return fail(JLC_ERR_UNSUPPORTED, "CRT form: %lld x %lld x %lld exceeds the 32-bit block count",
                    (long long)m, (long long)n, (long long)k);
    p.pa = pa;
    p.pb = pb;
    p.ea = ta;
    p.eb = tb;
    if ((rc = ensure_buf(4, (size_t)(nmod * p.a_plane_bytes))) != JLC_OK) return rc;
    if ((rc = ensure_buf(5, (size_t)(nmod * p.b_plane_bytes))) != JLC_OK) return rc;
    if ((rc = ensure_buf(7, (size_t)(nmod * p.c_plane_bytes))) != JLC_OK) return rc;
    p.a_pack = ctx.buf[4];
    p.b_pack = ctx.buf[5];
    p.c_res = ctx.buf[7];
    info->bits_a = pa;
    info->bits_b = pb;
    JLC_HIP_CHECK(jlc::launch_crt(la, lb, a, lda, b, ldb, p, alpha, beta, c, ldc, stream, info));
    return JLC_OK;
}

// The FP64 kernels of a call with alpha != 0 and k > 0: the matvec when n == 1 and op(A) is
// row-major (the n == 1 regime of GemmDispatch.selectKernel, GemmDispatch.java:80-82; x is the B
// column, y the C column), else the FP64 MFMA kernel, split over K for long-K / few-tile shapes. The
// split-K partial products go to the thread's grow-only workspace (ordered across streams by ws), or,
// when `fixed` is given, to that caller-owned buffer of fixed_bytes; a call whose partials do not fit
// it fails before anything is launched.
// Largest workspace a wave-quantization split may claim (the grow-only per-thread buffer keeps it).
constexpr size_t QUANT_SPLIT_MAX_BYTES = (size_t)1 << 30;

int run_fp64(int la, int lb, int64_t m, int64_t n, int64_t k, double alpha, const double* a, int64_t lda,
             const double* b, int64_t ldb, double beta, double* c, int64_t ldc, hipStream_t stream,
             jlc::TileInfo* info, int* kernel, WorkspaceUse* ws, double* fixed, size_t fixed_bytes) {
    if (n == 1 && la == JLC_ROW) {
        *kernel = JLC_KERNEL_GEMV;
        int64_t incx = lb == JLC_ROW ? ldb : 1;
        JLC_HIP_CHECK(jlc::launch_dgemv(m, k, alpha, a, lda, b, incx, beta, c, ldc, stream, info));
        return JLC_OK;
    }
    *kernel = JLC_KERNEL_MFMA_F64;
    const jlc::TuneTable tt = jlc::tune_snapshot();
    int slices = jlc::dgemm_splitk_slices(m, n, k, g_cus.load(), tt.splitk_tiles_per_cu, tt.splitk_min_kslice,
                                          tt.splitk_quant != 0);
    if (slices == 1 && tt.splitk_quant)
        slices = jlc::dgemm_quant_slices(m, n, k, g_cus.load(),
                                         fixed ? (int64_t)fixed_bytes : (int64_t)QUANT_SPLIT_MAX_BYTES);
    if (slices > 1) {
        const size_t need = (size_t)slices * (size_t)(m * n) * sizeof(double);
        double* work = fixed;
        if (fixed) {
            if (need > fixed_bytes)
                return fail(JLC_ERR_UNSUPPORTED,
                            "split-K partials of %lld x %lld x %lld (%d slices) need %zu bytes, above the %zu-byte "
                            "scratch", (long long)m, (long long)n, (long long)k, slices, need, fixed_bytes);
        } else {
            int rc = ws->acquire();
            if (rc != JLC_OK) return rc;
            if ((rc = ensure_buf(6, need)) != JLC_OK) return rc;
            work = static_cast<double*>(t_ctx.buf[6]);
        }
        JLC_HIP_CHECK(jlc::launch_dgemm_f64_splitk(la, lb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, slices,
                                                   work, stream, info));
        return JLC_OK;
    }
    JLC_HIP_CHECK(jlc::launch_dgemm_f64(la, lb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, stream, info));
    return JLC_OK;
}

// Device-side dispatch shared by the host and device entry points.
int run_gemm_device(int la, int lb, int64_t m, int64_t n, int64_t k, double alpha, const double* a, int64_t lda,
                    const double* b, int64_t ldb, double beta, double* c, int64_t ldc, int mode,
                    hipStream_t stream, jlc::TileInfo* info, int* kernel) {
    if (alpha == 0.0 || k == 0) {
        *kernel = JLC_KERNEL_SCALE;
        JLC_HIP_CHECK(jlc::launch_scale(m, n, beta, c, ldc, stream, info));
        return JLC_OK;
    }
    WorkspaceUse ws(stream);
    int form = mode & 0xff;
    if (form != JLC_MODE_FP64) {
        int rc = ws.acquire();
        if (rc != JLC_OK) return rc;
    }
    if (form == JLC_MODE_SPLIT_CRT) {
        bool fallback = false;
        *kernel = JLC_KERNEL_SPLIT_CRT;
        int rc = run_crt(la, lb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, (mode >> 8) & 0xff, stream, info,
                         &fallback);
        if (rc != JLC_OK || !fallback) return rc;
        form = JLC_MODE_FP64;  // spread too wide for the moduli: the FP64 kernel serves the call
    }
    if (form == JLC_MODE_SPLIT_INT8 || form == JLC_MODE_SPLIT_BF16) {
        const bool bf16 = form == JLC_MODE_SPLIT_BF16;
        bool fallback = false;
        *kernel = bf16 ? JLC_KERNEL_SPLIT_BF16 : JLC_KERNEL_SPLIT_INT8;
        int rc = run_split(bf16, la, lb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, (mode >> 8) & 0xff, stream,
                           info, &fallback);
        if (rc != JLC_OK || !fallback) return rc;
        // spread too wide for the form's slices: the FP64 kernel serves the call
    }
    return run_fp64(la, lb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, stream, info, kernel, &ws, nullptr, 0);
}

// The three products of a compact-WY update on device pointers (jlc_wy_apply_dev). W and W2 live in
// the thread's workspace (slots 8 and 9); the large products run in `mode`, the small T product in FP64.
int run_wy_apply(int side, int trans_t, int64_t m, int64_t n, int64_t kb, const double* v, int64_t ldv,
                 const double* t, int64_t ldt, double* c, int64_t ldc, int mode, hipStream_t s,
                 jlc::TileInfo* info, int* kernel) {
    WorkspaceUse ws(s);
    int rc = ws.acquire();
    if (rc != JLC_OK) return rc;
    const jlc::TuneTable tt = jlc::tune_snapshot();
    const int wgs = side == JLC_LEFT ? tt.wy_left_wgs_per_cu : tt.wy_right_wgs_per_cu;
    if ((mode & 0xff) == JLC_MODE_FP64 && kb <= 64 && wgs > 0 && ((m + 63) / 64) * ((n + 47) / 32) < INT32_MAX) {
        // panels of at most 64 reflectors: the WY kernels (partial W, W2, update), C streamed twice
        const jlc::WyPlan plan = jlc::wy_plan(side, m, n, kb, c, ldc, g_cus.load(), wgs);
        if ((rc = ensure_buf(8, (size_t)plan.part_elems * sizeof(double))) != JLC_OK) return rc;
        if ((rc = ensure_buf(9, (size_t)plan.w2_elems * sizeof(double))) != JLC_OK) return rc;
        JLC_HIP_CHECK(jlc::launch_wy_update(side, trans_t, (int)m, (int)n, (int)kb, v, ldv, t, ldt, c, ldc, plan,
                                            static_cast<double*>(t_ctx.buf[8]), static_cast<double*>(t_ctx.buf[9]),
                                            s));
        *kernel = JLC_KERNEL_WY;
        info->tile_m = 64;
        info->tile_n = 64;
        info->tile_k = plan.kb_pad;
        info->slices = plan.slices;
        info->launches += 3;
        return JLC_OK;
    }
    const size_t wbytes = (size_t)(side == JLC_LEFT ? kb * n : m * kb) * sizeof(double);
    if ((rc = ensure_buf(8, wbytes)) != JLC_OK) return rc;
    if ((rc = ensure_buf(9, wbytes)) != JLC_OK) return rc;
    double* w = static_cast<double*>(t_ctx.buf[8]);
    double* w2 = static_cast<double*>(t_ctx.buf[9]);
    const int lt = trans_t ? JLC_COL : JLC_ROW;
    int k2 = 0;
    if (side == JLC_LEFT) {
        // W = V^T C (kb x n); W2 = op(T) W; C -= V W2.
        rc = run_gemm_device(JLC_COL, JLC_ROW, kb, n, m, 1.0, v, ldv, c, ldc, 0.0, w, n, mode, s, info, kernel);
        if (rc == JLC_OK)
            rc = run_gemm_device(lt, JLC_ROW, kb, n, kb, 1.0, t, ldt, w, n, 0.0, w2, n, JLC_MODE_FP64, s, info, &k2);
        if (rc == JLC_OK)
            rc = run_gemm_device(JLC_ROW, JLC_ROW, m, n, kb, -1.0, v, ldv, w2, n, 1.0, c, ldc, mode, s, info, &k2);
    } else {
        // W = C V (m x kb); W2 = W op(T); C -= W2 V^T.
        rc = run_gemm_device(JLC_ROW, JLC_ROW, m, kb, n, 1.0, c, ldc, v, ldv, 0.0, w, kb, mode, s, info, kernel);
        if (rc == JLC_OK)
            rc = run_gemm_device(JLC_ROW, lt, m, kb, kb, 1.0, w, kb, t, ldt, 0.0, w2, kb, JLC_MODE_FP64, s, info, &k2);
        if (rc == JLC_OK)
            rc = run_gemm_device(JLC_ROW, JLC_COL, m, n, kb, -1.0, w2, kb, v, ldv, 1.0, c, ldc, mode, s, info, &k2);
    }
    return rc;
}

// Argument checks shared by jlc_wy_apply and jlc_wy_apply_dev.
int validate_wy(int side, int trans_t, int64_t m, int64_t n, int64_t kb, const double* v, int64_t ldv,
                const double* t, int64_t ldt, const double* c, int64_t ldc, int mode) {
    if (side != JLC_LEFT && side != JLC_RIGHT) return fail(JLC_ERR_INVALID_ARG, "side must be JLC_LEFT or JLC_RIGHT");
    if (trans_t != 0 && trans_t != 1) return fail(JLC_ERR_INVALID_ARG, "trans_t must be 0 or 1");
    if (m < 0 || n < 0 || kb < 0) return fail(JLC_ERR_INVALID_ARG, "negative dimension");
    if (m > INT32_MAX || n > INT32_MAX || kb > INT32_MAX) return fail(JLC_ERR_INVALID_ARG, "dimension above 2^31-1");
    int rc = validate_gemm(JLC_ROW, JLC_ROW, 0, 0, 0, 1.0, nullptr, 0, nullptr, 0, nullptr, 0, mode);
    if (rc != JLC_OK) return rc;
    if (m == 0 || n == 0 || kb == 0) return JLC_OK;
    if (!v || !t || !c) return fail(JLC_ERR_INVALID_ARG, "V, T and C must not be null");
    if (ldv < kb || ldt < kb || ldc < n)
        return fail(JLC_ERR_INVALID_ARG, "need ldv >= kb, ldt >= kb, ldc >= n (ldv=%lld ldt=%lld ldc=%lld)",
                    (long long)ldv, (long long)ldt, (long long)ldc);
    return JLC_OK;
}

void fill_stats(jlc_stats* s, const jlc::TileInfo& info, int kernel, double kms, double tms) {
    if (!s) return;
    s->kernel_ms = kms;
    s->total_ms = tms;
    s->kernel = kernel;
    s->tile_m = info.tile_m;
    s->tile_n = info.tile_n;
    s->tile_k = info.tile_k;
    s->slices = info.slices;
    s->bits_a = info.bits_a;
    s->bits_b = info.bits_b;
    s->main_ms = info.main_ms;
    s->launches = info.launches;
}

}  // namespace

namespace {
bool tune_field(jlc::TuneTable& t, const char* key, double* get, const double* set);

// Parses "key value" lines into t; JLC_OK or an error naming the file and line.
int parse_tune_file(const char* path, jlc::TuneTable& t) {
    FILE* f = std::fopen(path, "r");
    if (!f) return fail(JLC_ERR_INVALID_ARG, "cannot open tuning table %s", path);
    char line[512];
    int lineno = 0;
    while (std::fgets(line, sizeof line, f)) {
        lineno++;
        char* p = line;
        while (*p == ' ' || *p == '\t') p++;
        if (*p == '#' || *p == '\n' || *p == '\0') continue;
        char key[128];
        double v;
        if (std::sscanf(p, "%127s %lf", key, &v) != 2 || !tune_field(t, key, nullptr, &v)) {
            std::fclose(f);
            return fail(JLC_ERR_INVALID_ARG, "%s:%d: bad tuning entry", path, lineno);
        }
    }
    std::fclose(f);
    return JLC_OK;
}

// First use of the table: load $JLC_TUNE_TABLE if set, else the table shipped
// beside the library (<libdir>/../tuning/mi355x.tune), so every host (FFM, C++,
// Python) sees the same tuned values. A missing shipped table keeps the
// built-in defaults; a malformed one is ignored the same way.
void tune_autoload() {
    std::call_once(g_tune_autoload, [] {
        std::string path;
        if (const char* env = std::getenv("JLC_TUNE_TABLE")) {
            path = env;
        } else {
            Dl_info info{};
            if (dladdr(reinterpret_cast<void*>(&tune_autoload), &info) && info.dli_fname) {
                path = info.dli_fname;
                size_t slash = path.rfind('/');
                path = (slash == std::string::npos ? std::string(".") : path.substr(0, slash)) +
                       "/../tuning/mi355x.tune";
            }
        }
        if (path.empty()) return;
        jlc::TuneTable t;
        std::string saved = t_error;
        if (parse_tune_file(path.c_str(), t) == JLC_OK) {
            std::lock_guard<std::mutex> lock(g_tune_mutex);
            g_tune = t;
        }
        t_error = saved;
    });
}
}  // namespace

namespace jlc {
TuneTable tune_snapshot() {
    tune_autoload();
    std::lock_guard<std::mutex> lock(g_tune_mutex);
    return g_tune;
}
}  // namespace jlc

namespace {
// Reads or writes one named entry of a table (false for an unknown key).
bool tune_field(jlc::TuneTable& t, const char* key, double* get, const double* set) {
    struct F {
        const char* name;
        double jlc::TuneTable::*d;
        int jlc::TuneTable::*i;
    };
    static const F fields[] = {
        {"hip_min_flops", &jlc::TuneTable::hip_min_flops, nullptr},
        {"splitk_tiles_per_cu", nullptr, &jlc::TuneTable::splitk_tiles_per_cu},
        {"splitk_min_kslice", nullptr, &jlc::TuneTable::splitk_min_kslice},
        {"grouped_waves_per_simd", nullptr, &jlc::TuneTable::grouped_waves_per_simd},
        {"grouped_xcds", nullptr, &jlc::TuneTable::grouped_xcds},
        {"splitk_quant", nullptr, &jlc::TuneTable::splitk_quant},
        {"grouped_small_tile", nullptr, &jlc::TuneTable::grouped_small_tile},
        {"grouped_chunk_tiles", nullptr, &jlc::TuneTable::grouped_chunk_tiles},
        {"crt_bits_offset", &jlc::TuneTable::crt_bits_offset, nullptr},
        {"crt_bits_slope", &jlc::TuneTable::crt_bits_slope, nullptr},
        {"crt_waves", nullptr, &jlc::TuneTable::crt_waves},
        {"crt_group", nullptr, &jlc::TuneTable::crt_group},
        {"crt_xcd", nullptr, &jlc::TuneTable::crt_xcd},
        {"qr_wave_max_lds", nullptr, &jlc::TuneTable::qr_wave_max_lds},
        {"qr_block_threads", nullptr, &jlc::TuneTable::qr_block_threads},
        {"qr_reg_max_m", nullptr, &jlc::TuneTable::qr_reg_max_m},
        {"qr_cqr_min_rows", nullptr, &jlc::TuneTable::qr_cqr_min_rows},
        {"wy_left_wgs_per_cu", nullptr, &jlc::TuneTable::wy_left_wgs_per_cu},
        {"wy_right_wgs_per_cu", nullptr, &jlc::TuneTable::wy_right_wgs_per_cu},
    };
    for (const F& f : fields) {
        if (std::strcmp(f.name, key) != 0) continue;
        if (set) {
            if (f.d) t.*(f.d) = *set;
            else t.*(f.i) = (int)*set;
        }
        if (get) *get = f.d ? t.*(f.d) : (double)(t.*(f.i));
        return true;
    }
    return false;
}
}  // namespace

extern "C" {

int jlc_tune_get(const char* key, double* value) {
    if (!key || !value) return fail(JLC_ERR_INVALID_ARG, "key and value must not be null");
    tune_autoload();
    std::lock_guard<std::mutex> lock(g_tune_mutex);
    if (!tune_field(g_tune, key, value, nullptr)) return fail(JLC_ERR_INVALID_ARG, "unknown tuning key '%s'", key);
    return JLC_OK;
}

int jlc_tune_set(const char* key, double value) {
    if (!key) return fail(JLC_ERR_INVALID_ARG, "key must not be null");
    tune_autoload();
    std::lock_guard<std::mutex> lock(g_tune_mutex);
    if (!tune_field(g_tune, key, nullptr, &value)) return fail(JLC_ERR_INVALID_ARG, "unknown tuning key '%s'", key);
    return JLC_OK;
}

int jlc_tune_load(const char* path) {
    if (!path) return fail(JLC_ERR_INVALID_ARG, "path must not be null");
    jlc::TuneTable t = jlc::tune_snapshot();
    int rc = parse_tune_file(path, t);
    if (rc != JLC_OK) return rc;
    std::lock_guard<std::mutex> lock(g_tune_mutex);
    g_tune = t;
    return JLC_OK;
}

int jlc_tune_write(const char* path) {
    if (!path) return fail(JLC_ERR_INVALID_ARG, "path must not be null");
    const jlc::TuneTable t = jlc::tune_snapshot();
    FILE* f = std::fopen(path, "w");
    if (!f) return fail(JLC_ERR_INVALID_ARG, "cannot write %s", path);
    std::fprintf(f, "# libjlc_hip tuning table (key value); written by jlc_amd/tune.py\n");
    std::fprintf(f, "hip_min_flops %.17g\n", t.hip_min_flops);
    std::fprintf(f, "splitk_tiles_per_cu %d\n", t.splitk_tiles_per_cu);
    std::fprintf(f, "splitk_min_kslice %d\n", t.splitk_min_kslice);
    std::fprintf(f, "grouped_waves_per_simd %d\n", t.grouped_waves_per_simd);
    std::fprintf(f, "grouped_xcds %d\n", t.grouped_xcds);
    std::fprintf(f, "splitk_quant %d\n", t.splitk_quant);
    std::fprintf(f, "grouped_small_tile %d\n", t.grouped_small_tile);
    std::fprintf(f, "grouped_chunk_tiles %d\n", t.grouped_chunk_tiles);
    std::fprintf(f, "crt_bits_offset %.17g\n", t.crt_bits_offset);
    std::fprintf(f, "crt_bits_slope %.17g\n", t.crt_bits_slope);
    std::fprintf(f, "crt_waves %d\n", t.crt_waves);
    std::fprintf(f, "crt_group %d\n", t.crt_group);
    std::fprintf(f, "crt_xcd %d\n", t.crt_xcd);
    std::fprintf(f, "qr_wave_max_lds %d\n", t.qr_wave_max_lds);
    std::fprintf(f, "qr_block_threads %d\n", t.qr_block_threads);
    std::fprintf(f, "qr_reg_max_m %d\n", t.qr_reg_max_m);
    std::fprintf(f, "qr_cqr_min_rows %d\n", t.qr_cqr_min_rows);
    std::fprintf(f, "wy_left_wgs_per_cu %d\n", t.wy_left_wgs_per_cu);
    std::fprintf(f, "wy_right_wgs_per_cu %d\n", t.wy_right_wgs_per_cu);
    std::fclose(f);
    return JLC_OK;
}

const char* jlc_version(void) { return "libjlc_hip 0.1 gfx950"; }

const char* jlc_last_error(void) { return t_error.c_str(); }

int jlc_hip_probe(int* device_count, char* name, size_t name_len) {
    if (device_count) *device_count = 0;
    if (name && name_len) name[0] = '\0';
    if (!env_flag("JLC_HIP_ENABLED", true)) return fail(JLC_ERR_DISABLED, "JLC_HIP_ENABLED disables the HIP path");
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count <= 0) return fail(JLC_ERR_NO_DEVICE, "no HIP device visible");
    if (device_count) *device_count = count;
    return check_arch(0, name, name_len);
}

int jlc_hip_init(int device) {
    std::lock_guard<std::mutex> lock(g_mutex);
    if (!env_flag("JLC_HIP_ENABLED", true)) return fail(JLC_ERR_DISABLED, "JLC_HIP_ENABLED disables the HIP path");
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count <= 0) return fail(JLC_ERR_NO_DEVICE, "no HIP device visible");
    if (device < 0 || device >= count) return fail(JLC_ERR_INVALID_ARG, "device %d out of range [0,%d)", device, count);
    int rc = check_arch(device, nullptr, 0);
    if (rc != JLC_OK) return rc;
    JLC_HIP_CHECK(hipSetDevice(device));
    {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
            g_cus.store(prop.multiProcessorCount);
    }
    // Once per device, with the device current: the device address of this device's scratch copy.
    if (device < MAX_DEVICES && !g_scratch_ptr[device].load()) {
        void* ptr = nullptr;
        JLC_HIP_CHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) return fail(JLC_ERR_HIP, "hipGetSymbolAddress returned no address for the scratch array");
        g_scratch_ptr[device].store(ptr);
    }
    static bool registered = false;
    if (!registered) {
        std::atexit(mark_dead);
        registered = true;
    }
    g_device.store(device);
    return JLC_OK;
}

int jlc_hip_shutdown(void) {
    ThreadCtx& c = t_ctx;
    if (c.device >= 0) {
        for (int i = 0; i < ThreadCtx::NBUF; i++) {
            if (c.buf[i]) (void)hipFree(c.buf[i]);
            c.buf[i] = nullptr;
            c.cap[i] = 0;
        }
        if (c.stream) (void)hipStreamSynchronize(c.stream);
    }
    return JLC_OK;
}

int jlc_dgemm_dev(int layout_a, int layout_b, int64_t m, int64_t n, int64_t k, double alpha, const double* a,
                  int64_t lda, const double* b, int64_t ldb, double beta, double* c, int64_t ldc, int mode,
                  void* stream, jlc_stats* stats) {
    double t0 = now_ms();
    int rc = validate_gemm(layout_a, layout_b, m, n, k, alpha, a, lda, b, ldb, c, ldc, mode);
    if (rc != JLC_OK) return rc;
    jlc::TileInfo info;
    if (m == 0 || n == 0) {
        fill_stats(stats, info, JLC_KERNEL_NONE, 0.0, now_ms() - t0);
        return JLC_OK;
    }
    rc = ensure_init();
    if (rc != JLC_OK) return rc;
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : t_ctx.stream;
    int kernel = JLC_KERNEL_NONE;
    if (stats) {
        JLC_HIP_CHECK(hipEventRecord(t_ctx.ev0, s));
        info.main_begin = t_ctx.ev2;
        info.main_end = t_ctx.ev3;
    }
    rc = run_gemm_device(layout_a, layout_b, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, mode, s, &info, &kernel);
    if (rc != JLC_OK) return rc;
    double kms = 0.0;
    if (stats) {
        JLC_HIP_CHECK(hipEventRecord(t_ctx.ev1, s));
        JLC_HIP_CHECK(hipEventSynchronize(t_ctx.ev1));
        float f = 0.0f;
        JLC_HIP_CHECK(hipEventElapsedTime(&f, t_ctx.ev0, t_ctx.ev1));
        kms = f;
        if (kernel != JLC_KERNEL_SCALE) {
            JLC_HIP_CHECK(hipEventElapsedTime(&f, t_ctx.ev2, t_ctx.ev3));
            info.main_ms = f;
        }
    }
    fill_stats(stats, info, kernel, kms, now_ms() - t0);
    return JLC_OK;
}

int jlc_wy_apply_dev(int side, int trans_t, int64_t m, int64_t n, int64_t kb, const double* v, int64_t ldv,
                     const double* t, int64_t ldt, double* c, int64_t ldc, int mode, void* stream, jlc_stats* stats) {
    double t0 = now_ms();
    int rc = validate_wy(side, trans_t, m, n, kb, v, ldv, t, ldt, c, ldc, mode);
    if (rc != JLC_OK) return rc;
    jlc::TileInfo info;
    if (m == 0 || n == 0 || kb == 0) {
        fill_stats(stats, info, JLC_KERNEL_NONE, 0.0, now_ms() - t0);
        return JLC_OK;
    }
    if ((rc = ensure_init()) != JLC_OK) return rc;
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : t_ctx.stream;
    if (stats) JLC_HIP_CHECK(hipEventRecord(t_ctx.ev0, s));
    int kernel = JLC_KERNEL_NONE;
    if ((rc = run_wy_apply(side, trans_t, m, n, kb, v, ldv, t, ldt, c, ldc, mode, s, &info, &kernel)) != JLC_OK)
        return rc;
    double kms = 0.0;
    if (stats) {
        JLC_HIP_CHECK(hipEventRecord(t_ctx.ev1, s));
        JLC_HIP_CHECK(hipEventSynchronize(t_ctx.ev1));
        float f = 0.0f;
        JLC_HIP_CHECK(hipEventElapsedTime(&f, t_ctx.ev0, t_ctx.ev1));
        kms = f;
        info.main_ms = 0.0;
    }
    fill_stats(stats, info, kernel, kms, now_ms() - t0);
    return JLC_OK;
}

int jlc_wy_apply(int side, int trans_t, int64_t m, int64_t n, int64_t kb, const double* v, int64_t ldv,
                 const double* t, int64_t ldt, double* c, int64_t ldc, int mode, jlc_stats* stats) {
    double t0 = now_ms();
    int rc = validate_wy(side, trans_t, m, n, kb, v, ldv, t, ldt, c, ldc, mode);
    if (rc != JLC_OK) return rc;
    jlc::TileInfo info;
    if (m == 0 || n == 0 || kb == 0) {
        fill_stats(stats, info, JLC_KERNEL_NONE, 0.0, now_ms() - t0);
        return JLC_OK;
    }
    if ((rc = ensure_init()) != JLC_OK) return rc;
    ThreadCtx& ctx = t_ctx;
    hipStream_t s = ctx.stream;
    // Only the blocks the update touches cross PCIe, each packed with an even pitch: V (rows x kb),
    // T (kb x kb) and the m x n block of C, which comes back the same way.
    const int64_t vrows = side == JLC_LEFT ? m : n;
    const int64_t dk = (kb + 1) & ~int64_t(1), dn = n;
    if ((rc = ensure_buf(0, (size_t)(vrows * dk) * sizeof(double))) != JLC_OK) return rc;
    if ((rc = ensure_buf(1, (size_t)(kb * dk) * sizeof(double))) != JLC_OK) return rc;
    if ((rc = ensure_buf(2, (size_t)(m * dn) * sizeof(double))) != JLC_OK) return rc;
    double* dv = static_cast<double*>(ctx.buf[0]);
    double* dt = static_cast<double*>(ctx.buf[1]);
    double* dc = static_cast<double*>(ctx.buf[2]);
    const int64_t v_el = vrows * kb, t_el = kb * kb, c_el = m * n;
    if ((rc = ensure_pin((size_t)(v_el + t_el + c_el) * sizeof(double))) != JLC_OK) return rc;
    double* pin = static_cast<double*>(ctx.pin);
    if ((rc = stage_to_device(dv, dk, v, ldv, vrows, kb, pin, s)) != JLC_OK) return rc;
    if ((rc = stage_to_device(dt, dk, t, ldt, kb, kb, pin + v_el, s)) != JLC_OK) return rc;
    if ((rc = stage_to_device(dc, dn, c, ldc, m, n, pin + v_el + t_el, s)) != JLC_OK) return rc;
    JLC_HIP_CHECK(hipEventRecord(ctx.ev0, s));
    int kernel = JLC_KERNEL_NONE;
    if ((rc = run_wy_apply(side, trans_t, m, n, kb, dv, dk, dt, dk, dc, dn, mode, s, &info, &kernel)) != JLC_OK)
        return rc;
    JLC_HIP_CHECK(hipEventRecord(ctx.ev1, s));
    if ((rc = stage_from_device(c, ldc, dc, m, n, pin + v_el + t_el, s)) != JLC_OK) return rc;
    JLC_HIP_CHECK(hipStreamSynchronize(s));
    float f = 0.0f;
    JLC_HIP_CHECK(hipEventElapsedTime(&f, ctx.ev0, ctx.ev1));
    fill_stats(stats, info, kernel, f, now_ms() - t0);
    return JLC_OK;
}

int jlc_dgemm_dev_noalloc(int layout_a, int layout_b, int64_t m, int64_t n, int64_t k, double alpha,
                          const double* a, int64_t lda, const double* b, int64_t ldb, double beta, double* c,
                          int64_t ldc, void* stream) {
    int rc = validate_gemm(layout_a, layout_b, m, n, k, alpha, a, lda, b, ldb, c, ldc, JLC_MODE_FP64);
    if (rc != JLC_OK) return rc;
    if (m == 0 || n == 0) return JLC_OK;
    if (!env_flag("JLC_HIP_ENABLED", true)) return fail(JLC_ERR_DISABLED, "JLC_HIP_ENABLED disables the HIP path");
    if (g_device.load() < 0 && (rc = jlc_hip_init(0)) != JLC_OK) return rc;
    const int dev = g_device.load();
    void* scr = dev < MAX_DEVICES ? g_scratch_ptr[dev].load() : nullptr;
    if (!scr) return fail(JLC_ERR_UNSUPPORTED, "device %d has no resolved scratch (device index above %d?)", dev,
                          MAX_DEVICES - 1);
    JLC_HIP_CHECK(hipSetDevice(dev));
    hipStream_t s = static_cast<hipStream_t>(stream);
    jlc::TileInfo info;
    int kernel = JLC_KERNEL_NONE;
    if (alpha == 0.0 || k == 0) {
        JLC_HIP_CHECK(jlc::launch_scale(m, n, beta, c, ldc, s, &info));
        return JLC_OK;
    }
    return run_fp64(layout_a, layout_b, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, s, &info, &kernel, nullptr,
                    static_cast<double*>(scr), NOALLOC_SCRATCH_BYTES);
}

int jlc_dgemm(int layout_a, int layout_b, int64_t m, int64_t n, int64_t k, double alpha, const double* a,
              int64_t lda, const double* b, int64_t ldb, double beta, double* c, int64_t ldc, int mode,
              jlc_stats* stats) {
    double t0 = now_ms();
    int rc = validate_gemm(layout_a, layout_b, m, n, k, alpha, a, lda, b, ldb, c, ldc, mode);
    if (rc != JLC_OK) return rc;
    jlc::TileInfo info;
    if (m == 0 || n == 0) {
        fill_stats(stats, info, JLC_KERNEL_NONE, 0.0, now_ms() - t0);
        return JLC_OK;
    }
    rc = ensure_init();
    if (rc != JLC_OK) return rc;
    ThreadCtx& ctx = t_ctx;
    hipStream_t s = ctx.stream;
    const bool use_ab = k > 0 && alpha != 0.0;
    // Stage operands compactly: op(A) rows (ROW) or columns (COL) become
    // contiguous records, which is the reference's pack step done by the copy engine.
    int64_t a_rows = layout_a == JLC_ROW ? m : k, a_cols = layout_a == JLC_ROW ? k : m;
    int64_t b_rows = layout_b == JLC_ROW ? k : n, b_cols = layout_b == JLC_ROW ? n : k;
    // Round each device leading dimension up to even so 16-byte loads stay aligned.
    int64_t dlda = (a_cols + 1) & ~int64_t(1), dldb = (b_cols + 1) & ~int64_t(1), dldc = n;
    if (use_ab) {
        if ((rc = ensure_buf(0, (size_t)(a_rows * dlda) * sizeof(double))) != JLC_OK) return rc;
        if ((rc = ensure_buf(1, (size_t)(b_rows * dldb) * sizeof(double))) != JLC_OK) return rc;
    }
    if ((rc = ensure_buf(2, (size_t)(m * dldc) * sizeof(double))) != JLC_OK) return rc;
    double* da = static_cast<double*>(ctx.buf[0]);
    double* db = static_cast<double*>(ctx.buf[1]);
    double* dc = static_cast<double*>(ctx.buf[2]);
    // Pinned staging: A, B, C regions packed contiguously, one DMA copy each.
    const int64_t a_el = use_ab ? a_rows * a_cols : 0, b_el = use_ab ? b_rows * b_cols : 0, c_el = m * n;
    if ((rc = ensure_pin((size_t)(a_el + b_el + c_el) * sizeof(double))) != JLC_OK) return rc;
    double* pin = static_cast<double*>(ctx.pin);
    double* pin_a = pin;
    double* pin_b = pin + a_el;
    double* pin_c = pin + a_el + b_el;
    if (use_ab) {
        if ((rc = stage_to_device(da, dlda, a, lda, a_rows, a_cols, pin_a, s)) != JLC_OK) return rc;
        if ((rc = stage_to_device(db, dldb, b, ldb, b_rows, b_cols, pin_b, s)) != JLC_OK) return rc;
    }
    if (beta != 0.0 && (rc = stage_to_device(dc, dldc, c, ldc, m, n, pin_c, s)) != JLC_OK) return rc;
    int kernel = JLC_KERNEL_NONE;
    JLC_HIP_CHECK(hipEventRecord(ctx.ev0, s));
    rc = run_gemm_device(layout_a, layout_b, m, n, k, alpha, da, dlda, db, dldb, beta, dc, dldc, mode, s, &info,
                         &kernel);
    if (rc != JLC_OK) return rc;
    JLC_HIP_CHECK(hipEventRecord(ctx.ev1, s));
    if ((rc = stage_from_device(c, ldc, dc, m, n, pin_c, s)) != JLC_OK) return rc;
    JLC_HIP_CHECK(hipStreamSynchronize(s));
    float f = 0.0f;
    JLC_HIP_CHECK(hipEventElapsedTime(&f, ctx.ev0, ctx.ev1));
    fill_stats(stats, info, kernel, f, now_ms() - t0);
    return JLC_OK;
}

int jlc_dgemv_dev(int64_t m, int64_t k, double alpha, const double* a, int64_t lda, const double* x,
                  int64_t incx, double beta, double* y, int64_t incy, void* stream, jlc_stats* stats) {
    double t0 = now_ms();
    if (m < 0 || k < 0) return fail(JLC_ERR_INVALID_ARG, "negative dimension");
    if (m > INT32_MAX || k > INT32_MAX) return fail(JLC_ERR_INVALID_ARG, "dimension above 2^31-1");
    jlc::TileInfo info;
    if (m == 0) {
        fill_stats(stats, info, JLC_KERNEL_NONE, 0.0, now_ms() - t0);
        return JLC_OK;
    }
    if (!y || incy < 1) return fail(JLC_ERR_INVALID_ARG, "y must not be null and incy >= 1");
    if (k > 0 && alpha != 0.0 && (!a || !x || incx < 1 || lda < k))
        return fail(JLC_ERR_INVALID_ARG, "A/x must not be null, incx >= 1, lda >= k");
    int rc = ensure_init();
    if (rc != JLC_OK) return rc;
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : t_ctx.stream;
    if (stats) JLC_HIP_CHECK(hipEventRecord(t_ctx.ev0, s));
    JLC_HIP_CHECK(jlc::launch_dgemv(m, k, alpha, a, lda, x, incx, beta, y, incy, s, &info));
    double kms = 0.0;
    if (stats) {
        JLC_HIP_CHECK(hipEventRecord(t_ctx.ev1, s));
        JLC_HIP_CHECK(hipEventSynchronize(t_ctx.ev1));
        float f = 0.0f;
        JLC_HIP_CHECK(hipEventElapsedTime(&f, t_ctx.ev0, t_ctx.ev1));
        kms = f;
    }
    fill_stats(stats, info, JLC_KERNEL_GEMV, kms, now_ms() - t0);
    return JLC_OK;
}

int jlc_dgemv(int64_t m, int64_t k, double alpha, const double* a, int64_t lda, const double* x, int64_t incx,
              double beta, double* y, int64_t incy, jlc_stats* stats) {
    double t0 = now_ms();
    if (m < 0 || k < 0) return fail(JLC_ERR_INVALID_ARG, "negative dimension");
    jlc::TileInfo info;
    if (m == 0) {
        fill_stats(stats, info, JLC_KERNEL_NONE, 0.0, now_ms() - t0);
        return JLC_OK;
    }
    if (!y || incy < 1) return fail(JLC_ERR_INVALID_ARG, "y must not be null and incy >= 1");
    const bool use_ax = k > 0 && alpha != 0.0;
    if (use_ax && (!a || !x || incx < 1 || lda < k))
        return fail(JLC_ERR_INVALID_ARG, "A/x must not be null, incx >= 1, lda >= k");
    int rc = ensure_init();
    if (rc != JLC_OK) return rc;
    ThreadCtx& ctx = t_ctx;
    hipStream_t s = ctx.stream;
    int64_t dlda = (k + 1) & ~int64_t(1);
    if (use_ax) {
        if ((rc = ensure_buf(0, (size_t)(m * dlda) * sizeof(double))) != JLC_OK) return rc;
        if ((rc = ensure_buf(1, (size_t)k * sizeof(double))) != JLC_OK) return rc;
    }
    if ((rc = ensure_buf(2, (size_t)m * sizeof(double))) != JLC_OK) return rc;
    double* da = static_cast<double*>(ctx.buf[0]);
    double* dx = static_cast<double*>(ctx.buf[1]);
    double* dy = static_cast<double*>(ctx.buf[2]);
    const int64_t a_el = use_ax ? m * k : 0, x_el = use_ax ? k : 0;
    if ((rc = ensure_pin((size_t)(a_el + x_el + m) * sizeof(double))) != JLC_OK) return rc;
    double* pin_a = static_cast<double*>(ctx.pin);
    double* pin_x = pin_a + a_el;
    double* pin_y = pin_x + x_el;
    if (use_ax) {
        if ((rc = stage_to_device(da, dlda, a, lda, m, k, pin_a, s)) != JLC_OK) return rc;
        for (int64_t i = 0; i < k; i++) pin_x[i] = x[i * incx];
        JLC_HIP_CHECK(hipMemcpyAsync(dx, pin_x, (size_t)k * sizeof(double), hipMemcpyHostToDevice, s));
    }
    if (beta != 0.0) {
        for (int64_t i = 0; i < m; i++) pin_y[i] = y[i * incy];
        JLC_HIP_CHECK(hipMemcpyAsync(dy, pin_y, (size_t)m * sizeof(double), hipMemcpyHostToDevice, s));
    }
    JLC_HIP_CHECK(hipEventRecord(ctx.ev0, s));
    JLC_HIP_CHECK(jlc::launch_dgemv(m, k, alpha, da, dlda, dx, 1, beta, dy, 1, s, &info));
    JLC_HIP_CHECK(hipEventRecord(ctx.ev1, s));
    JLC_HIP_CHECK(hipMemcpyAsync(pin_y, dy, (size_t)m * sizeof(double), hipMemcpyDeviceToHost, s));
    JLC_HIP_CHECK(hipStreamSynchronize(s));
    for (int64_t i = 0; i < m; i++) y[i * incy] = pin_y[i];
    JLC_HIP_CHECK(hipStreamSynchronize(s));
    float f = 0.0f;
    JLC_HIP_CHECK(hipEventElapsedTime(&f, ctx.ev0, ctx.ev1));
    fill_stats(stats, info, JLC_KERNEL_GEMV, f, now_ms() - t0);
    return JLC_OK;
}

int jlc_dev_alloc(void** ptr, size_t bytes) {
    if (!ptr) return fail(JLC_ERR_INVALID_ARG, "ptr must not be null");
    *ptr = nullptr;
    int rc = ensure_init();
    if (rc != JLC_OK) return rc;
    hipError_t e = hipMalloc(ptr, bytes ? bytes : 1);
    if (e != hipSuccess) return fail(JLC_ERR_ALLOC, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    return JLC_OK;
}

int jlc_dev_free(void* ptr) {
    if (!ptr) return JLC_OK;
    JLC_HIP_CHECK(hipFree(ptr));
    return JLC_OK;
}

int jlc_memcpy_h2d(void* dst, const void* src, size_t bytes) {
    int rc = ensure_init();
    if (rc != JLC_OK) return rc;
    JLC_HIP_CHECK(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return JLC_OK;
}

int jlc_memcpy_d2h(void* dst, const void* src, size_t bytes) {
    int rc = ensure_init();
    if (rc != JLC_OK) return rc;
    JLC_HIP_CHECK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return JLC_OK;
}

int jlc_dev_synchronize(void) {
    int rc = ensure_init();
    if (rc != JLC_OK) return rc;
    JLC_HIP_CHECK(hipStreamSynchronize(t_ctx.stream));
    JLC_HIP_CHECK(hipDeviceSynchronize());
    return JLC_OK;
}

int jlc_dev_copy(void* dst, const void* src, size_t bytes, void* stream) {
    if (bytes && (!dst || !src)) return fail(JLC_ERR_INVALID_ARG, "dst and src must not be null");
    if (bytes == 0) return JLC_OK;
    int rc = ensure_init();
    if (rc != JLC_OK) return rc;
    JLC_HIP_CHECK(jlc::launch_copy(dst, src, bytes, stream ? static_cast<hipStream_t>(stream) : t_ctx.stream));
    return JLC_OK;
}

int jlc_stream_get(void** stream) {
    if (!stream) return fail(JLC_ERR_INVALID_ARG, "stream must not be null");
    int rc = ensure_init();
    if (rc != JLC_OK) return rc;
    *stream = t_ctx.stream;
    return JLC_OK;
}

int jlc_stream_create(void** stream) {
    if (!stream) return fail(JLC_ERR_INVALID_ARG, "stream must not be null");
    *stream = nullptr;
    int rc = ensure_init();
    if (rc != JLC_OK) return rc;
    hipStream_t s = nullptr;
    JLC_HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = s;
    return JLC_OK;
}

int jlc_stream_destroy(void* stream) {
    if (!stream) return JLC_OK;
    int rc = ensure_init();
    if (rc != JLC_OK) return rc;
    if (stream == t_ctx.stream) return fail(JLC_ERR_INVALID_ARG, "the thread's library stream is not the caller's to destroy");
    JLC_HIP_CHECK(hipStreamDestroy(static_cast<hipStream_t>(stream)));
    return JLC_OK;
}

int jlc_stream_synchronize(void* stream) {
    int rc = ensure_init();
    if (rc != JLC_OK) return rc;
    JLC_HIP_CHECK(hipStreamSynchronize(stream ? static_cast<hipStream_t>(stream) : t_ctx.stream));
    return JLC_OK;
}

static_assert(sizeof(jlc_gemm_desc) == sizeof(jlc::GroupedDesc), "descriptor layouts must match");

struct jlc_grouped_plan {
    int device = -1;
    int64_t count = 0;
    int ntiles = 0;
    int grid = 0;
    int slices = 0;  // 0: FP64 MFMA tiles; 3..9: the INT8 slice form with that many slices
    jlc::GroupedDesc* d_descs = nullptr;
    uint64_t* d_tiles = nullptr;
    // INT8 slice form: per-problem workspace offsets, pack jobs, packed fragments and exponents.
    jlc::GroupedPack* d_packs = nullptr;
    uint64_t* d_jobs = nullptr;
    int64_t njobs = 0;
    uint64_t* d_sjobs = nullptr;
    int64_t nsjobs = 0;
    void* d_frag = nullptr;
    int* d_expo = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

static int stats_for_stream(hipStream_t s, jlc_stats* stats, double t0, int slices, bool start) {
    if (!stats) return JLC_OK;
    if (start) {
        JLC_HIP_CHECK(hipEventRecord(t_ctx.ev0, s));
        return JLC_OK;
    }
    JLC_HIP_CHECK(hipEventRecord(t_ctx.ev1, s));
    JLC_HIP_CHECK(hipEventSynchronize(t_ctx.ev1));
    float f = 0.0f;
    JLC_HIP_CHECK(hipEventElapsedTime(&f, t_ctx.ev0, t_ctx.ev1));
    jlc::TileInfo info;
    info.tile_m = 32;
    info.tile_n = 32;
    info.tile_k = slices ? 32 : 16;
    info.slices = slices;
    info.launches = slices ? 3 : 1;  // the slice form packs its operands in two launches of their own
    fill_stats(stats, info, JLC_KERNEL_GROUPED, f, now_ms() - t0);
    return JLC_OK;
}

int jlc_grouped_plan_create(int64_t count, const jlc_gemm_desc* descs, int mode, jlc_grouped_plan** plan) {
    if (!plan) return fail(JLC_ERR_INVALID_ARG, "plan must not be null");
    *plan = nullptr;
    if (count < 0 || (count > 0 && !descs)) return fail(JLC_ERR_INVALID_ARG, "bad descriptor array");
    if (count >= (int64_t)1 << 31) return fail(JLC_ERR_INVALID_ARG, "more than 2^31-1 problems");
    // FP64 MFMA tiles, or the INT8 slice form with an explicit slice count (the same slices for every
    // problem: a count chosen from the data would need the data on the host at plan time).
    const int form = mode & 0xff, slices = (mode >> 8) & 0xff;
    if (form == JLC_MODE_FP64) {
        if (slices != 0) return fail(JLC_ERR_INVALID_ARG, "JLC_MODE_FP64 takes no slice count");
    } else if (form == JLC_MODE_SPLIT_INT8) {
        if (slices < jlc::SPLIT_MIN_SLICES || slices > jlc::SPLIT_MAX_SLICES)
            return fail(JLC_ERR_INVALID_ARG, "grouped JLC_MODE_SPLIT_INT8 needs an explicit slice count in [%d, %d]",
                        jlc::SPLIT_MIN_SLICES, jlc::SPLIT_MAX_SLICES);
    } else {
        return fail(JLC_ERR_UNSUPPORTED, "grouped launches serve JLC_MODE_FP64 and JLC_MODE_SPLIT_INT8");
    }
    // Problems longest K first (stable); each cut into chunks of whole tile rows of at most
    // grouped_chunk_tiles tiles (at least one tile row).
    for (int64_t p = 0; p < count; p++) {
        const jlc_gemm_desc& d = descs[p];
        int rc = validate_gemm(d.layout_a, d.layout_b, d.m, d.n, d.k, d.alpha, d.a, d.lda, d.b, d.ldb, d.c, d.ldc,
                               JLC_MODE_FP64);
        if (rc != JLC_OK) {
            std::string msg = t_error;
            return fail(rc, "descriptor %lld: %s", (long long)p, msg.c_str());
        }
        if (d.m > 32 * 32767 || d.n > 32 * 32767) return fail(JLC_ERR_INVALID_ARG, "descriptor %lld too large", (long long)p);
    }
    std::vector<int64_t> probs((size_t)count);
    for (int64_t p = 0; p < count; p++) probs[(size_t)p] = p;
    std::stable_sort(probs.begin(), probs.end(), [&](int64_t x, int64_t y) { return descs[x].k > descs[y].k; });
    int rc = ensure_init();
    if (rc != JLC_OK) return rc;
    hipDeviceProp_t prop;
    const int dev = g_device.load();
    const int cus = hipGetDeviceProperties(&prop, dev) == hipSuccess ? prop.multiProcessorCount : 256;
    const int wps = jlc::tune_snapshot().grouped_waves_per_simd;
    const int grid = form == JLC_MODE_FP64 ? jlc::grouped_grid(cus, wps) : jlc::grouped_split_grid(slices, cus, wps);
    const int64_t W = (int64_t)grid * jlc::GROUPED_WAVES_PER_GROUP;
    // Workgroups are dispatched round-robin over the XCDs (workgroup b on XCD b % X), so XCD x owns
    // wavefronts 4b + i for b = x (mod X). Each chunk goes to the XCD with the least work so far
    // (work = tiles * max(k, 16)), and each XCD deals its sequence over its own wavefronts.
    int X = std::max(1, jlc::tune_snapshot().grouped_xcds);
    if (grid % X != 0) X = 1;
    const int64_t Wx = W / X;
    // The FP64 form cuts a problem with m (n) up to grouped_small_tile into 16-row (16-column) tiles, so a
    // small problem does not pad its MFMA work to 32; the split form keeps 32 x 32 tiles.
    const int small = form == JLC_MODE_FP64 ? jlc::tune_snapshot().grouped_small_tile : 0;
    const int chunk = std::max(1, jlc::tune_snapshot().grouped_chunk_tiles);
    std::vector<std::vector<uint64_t>> seq((size_t)X);
    std::vector<double> work((size_t)X, 0.0);
    for (int64_t p : probs) {
        const jlc_gemm_desc& d = descs[p];
        // (tile indices hold 15 bits: 16-wide tiles only up to 16 * 32767)
        const bool rows16 = d.m <= small && d.m <= 16 * 32767, cols16 = d.n <= small && d.n <= 16 * 32767;
        const int TM = rows16 ? 16 : 32, TN = cols16 ? 16 : 32;
        const uint64_t shape = (uint64_t)((rows16 ? 2 : 0) | (cols16 ? 1 : 0)) << jlc::GROUPED_SHAPE_SHIFT;
        const int tm = (int)((d.m + TM - 1) / TM), tn = (int)((d.n + TN - 1) / TN);
        const int rows = std::max(1, chunk / std::max(1, tn));
        const double blocks = (double)(TM / 16) * (TN / 16);  // 16 x 16 MFMA blocks per tile
        for (int r0 = 0; r0 < tm; r0 += rows) {
            const int r1 = std::min(tm, r0 + rows);
            const size_t x = (size_t)(std::min_element(work.begin(), work.end()) - work.begin());
            work[x] += (double)(r1 - r0) * tn * blocks * std::max<int64_t>(d.k, 16);
            for (int i = r0; i < r1; i++)
                for (int j = 0; j < tn; j++)
                    seq[x].push_back(((uint64_t)p << 32) | shape | ((uint64_t)i << 15) | (uint64_t)j);
        }
    }
    size_t rounds = 0;
    for (const auto& q : seq) rounds = std::max(rounds, (size_t)((q.size() + Wx - 1) / Wx));
    if (rounds * (size_t)W >= ((size_t)1 << 31)) return fail(JLC_ERR_INVALID_ARG, "too many tiles");
    std::vector<uint64_t> tiles(rounds * (size_t)W, jlc::GROUPED_SKIP);
    size_t used = 0;
    for (int x = 0; x < X; x++)
        for (size_t i = 0; i < seq[(size_t)x].size(); i++) {
            const int64_t j = (int64_t)(i % (size_t)Wx), r = (int64_t)(i / (size_t)Wx);
            const int64_t b = x + (int64_t)X * (j / jlc::GROUPED_WAVES_PER_GROUP);
            const size_t t = (size_t)(r * W + b * jlc::GROUPED_WAVES_PER_GROUP + j % jlc::GROUPED_WAVES_PER_GROUP);
            tiles[t] = seq[(size_t)x][i];
            used = std::max(used, t + 1);
        }
    tiles.resize(used);
    // The INT8 slice form packs every operand block once per run: the workspace holds, per problem with
    // alpha != 0 and k > 0, S fragments of 1 KiB per (32-row block of op(A) or 32-column block of op(B),
    // 32-deep k block), and one exponent per padded row / column.
    std::vector<jlc::GroupedPack> packs;
    std::vector<uint64_t> jobs, sjobs;
    int64_t frag_blocks = 0, expo_count = 0;
    if (form != JLC_MODE_FP64) {
        packs.assign((size_t)count, jlc::GroupedPack{-1, -1, -1, -1});
        for (int64_t p = 0; p < count; p++) {
            const jlc_gemm_desc& d = descs[p];
            if (d.alpha == 0.0 || d.k <= 0 || d.m <= 0 || d.n <= 0) continue;
            const int64_t rb = (d.m + 31) / 32, cb = (d.n + 31) / 32, kb = (d.k + 31) / 32;
            jlc::GroupedPack& pk = packs[(size_t)p];
            pk.a_off = frag_blocks;
            frag_blocks += rb * kb * slices;
            pk.b_off = frag_blocks;
            frag_blocks += cb * kb * slices;
            pk.ea_off = expo_count;
            expo_count += rb * 32;
            pk.eb_off = expo_count;
            expo_count += cb * 32;
            if (kb >= ((int64_t)1 << 24)) return fail(JLC_ERR_INVALID_ARG, "descriptor %lld: k too large for the "
                                                                         "grouped slice form", (long long)p);
            for (int64_t i = 0; i < rb + cb; i++) {
                const uint64_t job = ((uint64_t)p << 32) | (i < rb ? (uint64_t)i : (1ull << 31) | (uint64_t)(i - rb));
                for (int64_t q = 0; q < kb; q++) sjobs.push_back(((uint64_t)jobs.size() << 24) | (uint64_t)q);
                jobs.push_back(job);
            }
        }
        if ((int64_t)sjobs.size() > (int64_t)INT32_MAX * jlc::GROUPED_WAVES_PER_GROUP)
            return fail(JLC_ERR_INVALID_ARG, "too many operand blocks");
    }
    jlc_grouped_plan* pl = new jlc_grouped_plan;
    pl->device = dev;
    pl->count = count;
    pl->ntiles = (int)tiles.size();
    pl->grid = grid;
    pl->slices = form == JLC_MODE_FP64 ? 0 : slices;
    pl->njobs = (int64_t)jobs.size();
    pl->nsjobs = (int64_t)sjobs.size();
    hipError_t e = hipMalloc(&pl->d_descs, std::max<size_t>(1, (size_t)count) * sizeof(jlc::GroupedDesc));
    if (e == hipSuccess) e = hipMalloc(&pl->d_tiles, std::max<size_t>(1, tiles.size()) * sizeof(uint64_t));
    if (e == hipSuccess && count)
        e = hipMemcpy(pl->d_descs, descs, (size_t)count * sizeof(jlc::GroupedDesc), hipMemcpyHostToDevice);
    if (e == hipSuccess && !tiles.empty())
        e = hipMemcpy(pl->d_tiles, tiles.data(), tiles.size() * sizeof(uint64_t), hipMemcpyHostToDevice);
    if (e == hipSuccess && form != JLC_MODE_FP64) {
        e = hipMalloc(&pl->d_packs, std::max<size_t>(1, packs.size()) * sizeof(jlc::GroupedPack));
        if (e == hipSuccess) e = hipMalloc(&pl->d_jobs, std::max<size_t>(1, jobs.size()) * sizeof(uint64_t));
        if (e == hipSuccess) e = hipMalloc(&pl->d_frag, std::max<size_t>(1, (size_t)frag_blocks) * 1024);
        if (e == hipSuccess) e = hipMalloc(&pl->d_expo, std::max<size_t>(1, (size_t)expo_count) * sizeof(int));
        if (e == hipSuccess && !packs.empty())
            e = hipMemcpy(pl->d_packs, packs.data(), packs.size() * sizeof(jlc::GroupedPack), hipMemcpyHostToDevice);
        if (e == hipSuccess && !jobs.empty())
            e = hipMemcpy(pl->d_jobs, jobs.data(), jobs.size() * sizeof(uint64_t), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMalloc(&pl->d_sjobs, std::max<size_t>(1, sjobs.size()) * sizeof(uint64_t));
        if (e == hipSuccess && !sjobs.empty())
            e = hipMemcpy(pl->d_sjobs, sjobs.data(), sjobs.size() * sizeof(uint64_t), hipMemcpyHostToDevice);
    }
    if (e != hipSuccess) {
        jlc_grouped_plan_destroy(pl);
        return fail(JLC_ERR_ALLOC, "grouped plan upload failed: %s", hipGetErrorString(e));
    }
    *plan = pl;
    return JLC_OK;
}

static hipError_t launch_plan(const jlc_grouped_plan* plan, hipStream_t s) {
    return plan->slices ? jlc::launch_grouped_split_i8(plan->slices, plan->d_descs, plan->d_packs, plan->d_jobs,
                                                       plan->njobs, plan->d_sjobs, plan->nsjobs, plan->d_tiles,
                                                       plan->ntiles, plan->grid, plan->d_frag, plan->d_expo, s)
                        : jlc::launch_grouped_f64(plan->d_descs, plan->d_tiles, plan->ntiles, plan->grid, s);
}

int jlc_grouped_plan_run(jlc_grouped_plan* plan, void* stream, jlc_stats* stats) {
    if (!plan) return fail(JLC_ERR_INVALID_ARG, "plan must not be null");
    double t0 = now_ms();
    int rc = ensure_init();
    if (rc != JLC_OK) return rc;
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : t_ctx.stream;
    if ((rc = stats_for_stream(s, stats, t0, plan->slices, true)) != JLC_OK) return rc;
    JLC_HIP_CHECK(launch_plan(plan, s));
    return stats_for_stream(s, stats, t0, plan->slices, false);
}

int jlc_grouped_plan_replay(jlc_grouped_plan* plan, void* stream, jlc_stats* stats) {
    if (!plan) return fail(JLC_ERR_INVALID_ARG, "plan must not be null");
    double t0 = now_ms();
    int rc = ensure_init();
    if (rc != JLC_OK) return rc;
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : t_ctx.stream;
    if (!plan->exec) {
        // Capture on a private stream so the caller's stream state is never in capture mode.
        hipStream_t cap;
        JLC_HIP_CHECK(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
        hipError_t e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
        if (e == hipSuccess) e = launch_plan(plan, cap);
        hipGraph_t graph = nullptr;
        hipError_t e2 = hipStreamEndCapture(cap, &graph);
        if (e == hipSuccess) e = e2;
        if (e == hipSuccess) e = hipGraphInstantiate(&plan->exec, graph, nullptr, nullptr, 0);
        (void)hipStreamDestroy(cap);
        if (e != hipSuccess) {
            if (graph) (void)hipGraphDestroy(graph);
            plan->exec = nullptr;
            return hip_fail(e, "grouped graph capture");
        }
        plan->graph = graph;
    }
    if ((rc = stats_for_stream(s, stats, t0, plan->slices, true)) != JLC_OK) return rc;
    JLC_HIP_CHECK(hipGraphLaunch(plan->exec, s));
    return stats_for_stream(s, stats, t0, plan->slices, false);
}

int jlc_grouped_plan_destroy(jlc_grouped_plan* plan) {
    if (!plan) return JLC_OK;
    if (g_alive.load()) {
        if (plan->exec) (void)hipGraphExecDestroy(plan->exec);
        if (plan->graph) (void)hipGraphDestroy(plan->graph);
        if (plan->d_descs) (void)hipFree(plan->d_descs);
        if (plan->d_tiles) (void)hipFree(plan->d_tiles);
        if (plan->d_packs) (void)hipFree(plan->d_packs);
        if (plan->d_jobs) (void)hipFree(plan->d_jobs);
        if (plan->d_sjobs) (void)hipFree(plan->d_sjobs);
        if (plan->d_frag) (void)hipFree(plan->d_frag);
        if (plan->d_expo) (void)hipFree(plan->d_expo);
    }
    delete plan;
    return JLC_OK;
}

int jlc_dgemm_grouped_dev(int64_t count, const jlc_gemm_desc* descs, int mode, void* stream, jlc_stats* stats) {
    jlc_grouped_plan* plan = nullptr;
    int rc = jlc_grouped_plan_create(count, descs, mode, &plan);
    if (rc != JLC_OK) return rc;
    rc = jlc_grouped_plan_run(plan, stream, stats);
    if (rc == JLC_OK) {
        hipStream_t s = stream ? static_cast<hipStream_t>(stream) : t_ctx.stream;
        hipError_t e = hipStreamSynchronize(s);
        if (e != hipSuccess) rc = hip_fail(e, "hipStreamSynchronize");
    }
    jlc_grouped_plan_destroy(plan);
    return rc;
}

// ---------------------------------------------------------------------------
// Blocked Householder QR (HouseholderQR.factorizeBlocked / updateTrailingMatrix,
// src/main/java/net/faulj/decomposition/qr/HouseholderQR.java:133-209).

namespace {

struct DevBuf {
    void* p = nullptr;
    ~DevBuf() {
        if (p && g_alive.load()) (void)hipFree(p);
    }
    hipError_t alloc(size_t bytes) { return hipMalloc(&p, bytes ? bytes : 8); }
    double* d() { return static_cast<double*>(p); }
};

int gemm_on(int la, int lb, int64_t m, int64_t n, int64_t k, double alpha, const double* a, int64_t lda,
            const double* b, int64_t ldb, double beta, double* c, int64_t ldc, int mode, hipStream_t s,
            jlc::TileInfo* info) {
    int kernel = 0;
    return run_gemm_device(la, lb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, mode, s, info, &kernel);
}

// V (as Vc = V^T, nb x mk) and T of the panel [k0, k0+nb) of a factored matrix; gpart holds
// gram_work_doubles(m - k0) doubles.
int panel_wy(const double* a, int64_t lda, int64_t m, int64_t k0, int64_t nb, const double* tau, double* vc,
             double* g, double* t, double* gpart, hipStream_t s, jlc::TileInfo* info) {
    const int64_t mk = m - k0;
    JLC_HIP_CHECK(jlc::launch_build_vc(a, lda, (int)m, (int)k0, (int)nb, vc, s));
    // G = V^T V, with Vc = V^T row-major, i.e. V column-major (ld mk).
    JLC_HIP_CHECK(jlc::launch_gram(vc, mk, (int)mk, (int)nb, gpart, g, s));
    JLC_HIP_CHECK(jlc::launch_wy_t(g, (int)nb, tau + k0, t, (int)nb, s));
    info->launches += 4;
    return JLC_OK;
}

// Workspace of a CholeskyQR2 panel: b x b matrices (row-major, b <= nb), two m x nb column-major panels
// (Q1, Q), the breakdown flag and the tau = 0 candidates.
struct CqrWork {
    double *g, *r1, *x1, *r, *x2, *ui, *q1, *q, *gpart;
    int *flag, *spec;
};

// Panel [k0, k0 + b) of the column-major m-row matrix by CholeskyQR2 and Householder reconstruction
// (qr_cqr.hip): G1 = Ap^T Ap, R1 = chol(G1), Q1 = Ap R1^-1, G2 = Q1^T Q1, R2 = chol(G2), then (when
// neither Cholesky reported a breakdown or a loss of orthogonality, read back here before anything is
// written to the matrix) Q = Q1 R2^-1, the modified LU of Q's top block (Y1, R_H = D R2 R1, tau, U^-1),
// Y2 = Q_bottom U^-1 into the panel, and the tau = 0 fix-up. *done = false leaves the panel untouched
// for the column kernels.
int panel_cqr(double* a, int64_t lda, int64_t m, int64_t k0, int64_t b, double* tau, const CqrWork& w,
              hipStream_t s, jlc::TileInfo* info, bool* done) {
    *done = false;
    const int64_t mp = m - k0;
    double* ap = a + k0 * lda + k0;
    const int bi = (int)b;
    int rc;
    JLC_HIP_CHECK(jlc::launch_gram(ap, lda, (int)mp, bi, w.gpart, w.g, s));
    JLC_HIP_CHECK(jlc::launch_cqr_chol(w.g, bi, 1, nullptr, w.r1, w.x1, w.flag, s));
    // Q1^T (b x mp, row-major) = X1^T Ap^T
    if ((rc = gemm_on(JLC_COL, JLC_ROW, b, mp, b, 1.0, w.x1, b, ap, lda, 0.0, w.q1, mp, JLC_MODE_FP64, s, info)) !=
        JLC_OK)
        return rc;
    JLC_HIP_CHECK(jlc::launch_gram(w.q1, mp, (int)mp, bi, w.gpart, w.g, s));
    JLC_HIP_CHECK(jlc::launch_cqr_chol(w.g, bi, 2, w.r1, w.r, w.x2, w.flag, s));
    info->launches += 6;
    // The flag is read back while Q = Q1 R2^-1 (workspace only, harmless if the panel falls back) runs.
    ThreadCtx& ctx = t_ctx;
    if (!ctx.qr_flag) JLC_HIP_CHECK(hipHostMalloc((void**)&ctx.qr_flag, sizeof(int), hipHostMallocDefault));
    if (!ctx.qr_ev) JLC_HIP_CHECK(hipEventCreateWithFlags(&ctx.qr_ev, hipEventDisableTiming));
    *ctx.qr_flag = -1;
    JLC_HIP_CHECK(hipMemcpyAsync(ctx.qr_flag, w.flag, sizeof(int), hipMemcpyDeviceToHost, s));
    JLC_HIP_CHECK(hipEventRecord(ctx.qr_ev, s));
    if ((rc = gemm_on(JLC_COL, JLC_ROW, b, mp, b, 1.0, w.x2, b, w.q1, mp, 0.0, w.q, mp, JLC_MODE_FP64, s, info)) !=
        JLC_OK)
        return rc;
    JLC_HIP_CHECK(hipEventSynchronize(ctx.qr_ev));
    if (*ctx.qr_flag != 0) return JLC_OK;
    JLC_HIP_CHECK(jlc::launch_cqr_lu(w.q, mp, bi, w.r, ap, lda, tau + k0, w.ui, w.spec, s));
    // Y2^T (b x (mp - b), rows of ap below the top block, ld lda) = U^-T Q_bottom^T
    if ((rc = gemm_on(JLC_COL, JLC_ROW, b, mp - b, b, 1.0, w.ui, b, w.q + b, mp, 0.0, ap + b, lda, JLC_MODE_FP64, s,
                      info)) != JLC_OK)
        return rc;
    JLC_HIP_CHECK(jlc::launch_cqr_special(ap, lda, (int)mp, bi, tau + k0, w.spec, s));
    info->launches += 2;
    *done = true;
    return JLC_OK;
}

}  // namespace

int jlc_dgeqrf_dev(int64_t m, int64_t n, double* a, int64_t lda, double* tau, int64_t nb, int mode, void* stream,
                   jlc_stats* stats) {
    double t0 = now_ms();
    if (m < 0 || n < 0 || m > INT32_MAX || n > INT32_MAX) return fail(JLC_ERR_INVALID_ARG, "bad dimensions");
    if (nb < 1 || nb > 128) return fail(JLC_ERR_INVALID_ARG, "panel width nb must be in [1, 128]");
    const int form = mode & 0xff;
    if (form != JLC_MODE_FP64 && form != JLC_MODE_SPLIT_INT8 && form != JLC_MODE_SPLIT_CRT)
        return fail(JLC_ERR_INVALID_ARG, "trailing-update mode must be FP64, SPLIT_INT8 or SPLIT_CRT");
    const int64_t kmax = m < n ? m : n;
    if (kmax == 0) return JLC_OK;
    if (!a || !tau || lda < m) return fail(JLC_ERR_INVALID_ARG, "a/tau must not be null and lda >= m");
    int rc = ensure_init();
    if (rc != JLC_OK) return rc;
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : t_ctx.stream;
    const int64_t G = (m + jlc::qr_panel_chunk() - 1) / jlc::qr_panel_chunk();
    DevBuf part, dots, vtmp, vc, g, t, w1, w2;
    // Two sets of panel partials (norm partials + pivot + pivot row), dot partials and column copies.
    hipError_t e = part.alloc((size_t)2 * (2 * G + 1 + 128) * 8);
    if (e == hipSuccess) e = vtmp.alloc((size_t)2 * m * 8);
    if (e == hipSuccess) e = dots.alloc((size_t)2 * G * 128 * 8);
    // Panels go in pairs: V and T of a pair (2nb reflectors) update the rest of the matrix in one pass.
    if (e == hipSuccess) e = vc.alloc((size_t)2 * nb * m * 8);
    if (e == hipSuccess) e = g.alloc((size_t)3 * nb * nb * 8);     // G of a panel, V_A^T V_B, G_AB T_B
    if (e == hipSuccess) e = t.alloc((size_t)4 * nb * nb * 8);     // T of the pair (2nb x 2nb)
    if (e == hipSuccess) e = w1.alloc((size_t)n * 2 * nb * 8);
    if (e == hipSuccess) e = w2.alloc((size_t)n * 2 * nb * 8);
    // CholeskyQR2 panels: those with at least qr_cqr_min_rows rows and twice their width.
    const int64_t cqr_min = jlc::tune_snapshot().qr_cqr_min_rows;
    const auto cqr_panel = [&](int64_t k0, int64_t b) {
        return cqr_min > 0 && m - k0 >= cqr_min && m - k0 >= 2 * b;
    };
    DevBuf cqp, cqs, gpart;
    CqrWork cw{};
    if (e == hipSuccess) e = gpart.alloc(std::max(jlc::gram_work_doubles(m), jlc::gram_cross_work_doubles(m)) * 8);
    if (e == hipSuccess && cqr_panel(0, std::min(nb, kmax))) {
        e = cqp.alloc((size_t)2 * m * nb * 8);
        if (e == hipSuccess) e = cqs.alloc((size_t)6 * nb * nb * 8 + (size_t)(nb + 1) * sizeof(int));
        if (e == hipSuccess) {
            cw.q1 = cqp.d();
            cw.q = cqp.d() + m * nb;
            double* sm = cqs.d();
            cw.g = sm;
            cw.r1 = sm + nb * nb;
            cw.x1 = sm + 2 * nb * nb;
            cw.r = sm + 3 * nb * nb;
            cw.x2 = sm + 4 * nb * nb;
            cw.ui = sm + 5 * nb * nb;
            cw.flag = reinterpret_cast<int*>(sm + 6 * nb * nb);
            cw.spec = cw.flag + 1;
            cw.gpart = gpart.d();
        }
    }
    if (e != hipSuccess) return fail(JLC_ERR_ALLOC, "QR workspace allocation failed: %s", hipGetErrorString(e));
    jlc::TileInfo info;
    // With stats, an event pair around each panel's last trailing-update product (C^T -= W2^T V^T, the
    // dominant kernel): main_ms is its device time summed over the panels.
    struct Events {
        std::vector<hipEvent_t> ev;
        ~Events() {
            if (g_alive.load())
                for (hipEvent_t x : ev) (void)hipEventDestroy(x);
        }
    } pev;
    if (stats) {
        const int64_t panels = (kmax + nb - 1) / nb;
        pev.ev.reserve((size_t)(2 * panels));
        for (int64_t q = 0; q < 2 * panels; q++) {
            hipEvent_t x;
            JLC_HIP_CHECK(hipEventCreate(&x));
            pev.ev.push_back(x);
        }
        JLC_HIP_CHECK(hipEventRecord(t_ctx.ev0, s));
    }
    int cqr_panels = 0, col_panels = 0;
    std::vector<size_t> timed;  // event pairs recorded (pairs with a trailing update)
    const int64_t reg_rows = std::min<int64_t>(jlc::tune_snapshot().qr_reg_max_m, 192);
    // Factors panel [k0, k1) in place: CholeskyQR2 where it applies, else column by column (on one
    // workgroup, the batched QR's register kernel, when the panel has at most qr_reg_max_m rows; else the
    // column kernels).
    const auto factor = [&](int64_t k0, int64_t k1) -> int {
        bool done = false;
        int r = JLC_OK;
        if (cw.q && cqr_panel(k0, k1 - k0) && (r = panel_cqr(a, lda, m, k0, k1 - k0, tau, cw, s, &info, &done)) != JLC_OK)
            return r;
        if (done) {
            cqr_panels++;
        } else if (m - k0 <= reg_rows && k1 - k0 <= 128) {
            JLC_HIP_CHECK(jlc::launch_qr_panel_reg(a + k0 * lda + k0, lda, (int)(m - k0), (int)(k1 - k0), tau + k0, s));
            info.launches += 1;
            col_panels++;
        } else {
            JLC_HIP_CHECK(jlc::launch_qr_panel_fused(a, lda, (int)m, (int)k0, (int)k1, tau, part.d(), dots.d(),
                                                     vtmp.d(), s));
            info.launches += (int)(k1 - k0) + 1;
            col_panels++;
        }
        return JLC_OK;
    };
    // C <- (I - V T^T V^T) C for C = A[k0:m, c0:c0+nc] held as C^T (nc x mk, row-major, ld = lda), V as Vc
    // (kb x mk, row stride mk) and T (kb x kb, row stride ldt): W^T = C^T V; W^T <- W^T T; C^T -= W^T V^T.
    const auto update = [&](int64_t k0, int64_t c0, int64_t nc, int64_t kb, const double* tm, int64_t ldt,
                            int umode, bool timed_product, size_t pi) -> int {
        const int64_t mk = m - k0;
        double* ct = a + c0 * lda + k0;
        int r;
        if (nc <= 128 && kb <= 128 && (umode & 0xff) == JLC_MODE_FP64) {
            // a panel-wide C: W^T = C^T V by the Gram kernel (the split-K GEMM has one output tile to spread)
            JLC_HIP_CHECK(jlc::launch_gram_cross(ct, lda, (int)nc, vc.d(), mk, (int)kb, (int)mk, gpart.d(), w1.d(),
                                                 (int)kb, s));
            info.launches += 2;
        } else if ((r = gemm_on(JLC_ROW, JLC_COL, nc, kb, mk, 1.0, ct, lda, vc.d(), mk, 0.0, w1.d(), kb, umode, s,
                                &info)) != JLC_OK) {
            return r;
        }
        if ((r = gemm_on(JLC_ROW, JLC_ROW, nc, kb, kb, 1.0, w1.d(), kb, tm, ldt, 0.0, w2.d(), kb, JLC_MODE_FP64, s,
                         &info)) != JLC_OK)
            return r;
        if (timed_product) {
            JLC_HIP_CHECK(hipEventRecord(pev.ev[pi], s));
            timed.push_back(pi);
        }
        if ((r = gemm_on(JLC_ROW, JLC_ROW, nc, mk, kb, -1.0, w2.d(), kb, vc.d(), mk, 1.0, ct, lda, umode, s, &info)) !=
            JLC_OK)
            return r;
        if (timed_product) JLC_HIP_CHECK(hipEventRecord(pev.ev[pi + 1], s));
        return JLC_OK;
    };
    double* const ga = g.d();                // G of one panel (nb x nb)
    double* const gab = g.d() + nb * nb;     // V_A^T V_B
    double* const xab = g.d() + 2 * nb * nb; // G_AB T_B
    for (int64_t k0 = 0; k0 < kmax; k0 += 2 * nb) {
        // Pair of panels A = [k0, ka), B = [ka, kb1) (B may be absent); T of the pair is 2nb x 2nb.
        const int64_t ka = std::min(k0 + nb, kmax), kb1 = std::min(k0 + 2 * nb, kmax);
        const int64_t nbA = ka - k0, nbB = kb1 - ka, mk = m - k0, kc = nbA + nbB;
        const int64_t ldt = kc;
        const size_t pi = (size_t)(2 * (k0 / nb));
        if ((rc = factor(k0, ka)) != JLC_OK) return rc;
        if (n - ka <= 0) continue;
        // V_A, G_A = V_A^T V_A, T_A into the pair's T.
        JLC_HIP_CHECK(jlc::launch_build_vc(a, lda, (int)m, (int)k0, (int)nbA, vc.d(), mk, 0, s));
        JLC_HIP_CHECK(jlc::launch_gram(vc.d(), mk, (int)mk, (int)nbA, gpart.d(), ga, s));
        JLC_HIP_CHECK(jlc::launch_wy_t(ga, (int)nbA, tau + k0, t.d(), (int)ldt, s));
        info.launches += 4;
        if (nbB == 0) {
            if ((rc = update(k0, ka, n - ka, nbA, t.d(), ldt, mode, stats != nullptr, pi)) != JLC_OK) return rc;
            continue;
        }
        // Panel B's columns take A's reflectors, then B is factored.
        if ((rc = update(k0, ka, nbB, nbA, t.d(), ldt, mode, false, 0)) != JLC_OK) return rc;
        if ((rc = factor(ka, kb1)) != JLC_OK) return rc;
        if (n - kb1 <= 0) continue;
        // V_B below V_A in Vc (its first nbA entries zero), G_B and T_B, then T_AB = -T_A (V_A^T V_B) T_B.
        double* vcb = vc.d() + nbA * mk;
        JLC_HIP_CHECK(jlc::launch_build_vc(a, lda, (int)m, (int)ka, (int)nbB, vcb, mk, (int)nbA, s));
        JLC_HIP_CHECK(jlc::launch_gram(vcb + nbA, mk, (int)(mk - nbA), (int)nbB, gpart.d(), ga, s));
        double* tb = t.d() + nbA * ldt + nbA;
        JLC_HIP_CHECK(jlc::launch_wy_t(ga, (int)nbB, tau + ka, tb, (int)ldt, s));
        JLC_HIP_CHECK(jlc::launch_gram_cross(vc.d() + nbA, mk, (int)nbA, vcb + nbA, mk, (int)nbB, (int)(mk - nbA),
                                             gpart.d(), gab, (int)nbB, s));
        JLC_HIP_CHECK(jlc::launch_fill_eye(t.d() + nbA * ldt, ldt, (int)nbB, (int)nbA, 0.0, s));
        info.launches += 7;
        if ((rc = gemm_on(JLC_ROW, JLC_ROW, nbA, nbB, nbB, 1.0, gab, nbB, tb, ldt, 0.0, xab, nbB, JLC_MODE_FP64, s,
                          &info)) != JLC_OK)
            return rc;
        if ((rc = gemm_on(JLC_ROW, JLC_ROW, nbA, nbB, nbA, -1.0, t.d(), ldt, xab, nbB, 0.0, t.d() + nbA, ldt,
                          JLC_MODE_FP64, s, &info)) != JLC_OK)
            return rc;
        if ((rc = update(k0, kb1, n - kb1, kc, t.d(), ldt, mode, stats != nullptr, pi)) != JLC_OK) return rc;
    }
    double kms = 0.0;
    if (stats) {
        JLC_HIP_CHECK(hipEventRecord(t_ctx.ev1, s));
        JLC_HIP_CHECK(hipEventSynchronize(t_ctx.ev1));
        float f = 0.0f;
        JLC_HIP_CHECK(hipEventElapsedTime(&f, t_ctx.ev0, t_ctx.ev1));
        kms = f;
        for (size_t q : timed) {
            float pf = 0.0f;
            JLC_HIP_CHECK(hipEventElapsedTime(&pf, pev.ev[q], pev.ev[q + 1]));
            info.main_ms += pf;
        }
    } else {
        JLC_HIP_CHECK(hipStreamSynchronize(s));  // workspace is freed on return
    }
    info.tile_m = (int)nb;
    info.tile_n = cqr_panels;
    info.tile_k = col_panels;
    fill_stats(stats, info, JLC_KERNEL_QR, kms, now_ms() - t0);
    return JLC_OK;
}

namespace {
// X <- H_0 H_1 ... H_{kk-1} X for the reflectors below the diagonal of the column-major factored matrix
// a (m rows, lda, tau), X column-major m x nx (ldx): panels of nb applied last to first as compact-WY
// updates on the library GEMM (buildQ's backward accumulation, HouseholderQR.java:390-444). When X
// starts as [I 0] (identity_start), rows k0.. of columns < k0 are still zero when panel k0 is applied,
// so only columns k0.. are updated.
int apply_q_panels(int64_t m, int64_t kk, const double* a, int64_t lda, const double* tau, int64_t nb, double* x,
                   int64_t ldx, int64_t nx, bool identity_start, hipStream_t s) {
    if (kk <= 0 || nx <= 0) return JLC_OK;
    DevBuf vc, g, t, y1, y2, gpart;
    hipError_t e = vc.alloc((size_t)nb * m * 8);
    if (e == hipSuccess) e = gpart.alloc(jlc::gram_work_doubles(m) * 8);
    if (e == hipSuccess) e = g.alloc((size_t)nb * nb * 8);
    if (e == hipSuccess) e = t.alloc((size_t)nb * nb * 8);
    if (e == hipSuccess) e = y1.alloc((size_t)nx * nb * 8);
    if (e == hipSuccess) e = y2.alloc((size_t)nx * nb * 8);
    if (e != hipSuccess) return fail(JLC_ERR_ALLOC, "Q workspace allocation failed: %s", hipGetErrorString(e));
    jlc::TileInfo info;
    int rc = JLC_OK;
    const int64_t last = ((kk - 1) / nb) * nb;
    for (int64_t k0 = last; k0 >= 0; k0 -= nb) {
        const int64_t k1 = k0 + nb < kk ? k0 + nb : kk;
        const int64_t nbk = k1 - k0, mk = m - k0;
        const int64_t c0 = identity_start ? k0 : 0, nq = nx - c0;
        if (nq <= 0) continue;
        if ((rc = panel_wy(a, lda, m, k0, nbk, tau, vc.d(), g.d(), t.d(), gpart.d(), s, &info)) != JLC_OK) return rc;
        // X <- (I - V T V^T) X on rows k0..m-1:  X^T -= ((X^T V) T^T) V^T  (X^T row-major, ld = ldx).
        double* xt = x + c0 * ldx + k0;
        if ((rc = gemm_on(JLC_ROW, JLC_COL, nq, nbk, mk, 1.0, xt, ldx, vc.d(), mk, 0.0, y1.d(), nbk, JLC_MODE_FP64, s,
                          &info)) != JLC_OK)
            return rc;
        if ((rc = gemm_on(JLC_ROW, JLC_COL, nq, nbk, nbk, 1.0, y1.d(), nbk, t.d(), nbk, 0.0, y2.d(), nbk,
                          JLC_MODE_FP64, s, &info)) != JLC_OK)
            return rc;
        if ((rc = gemm_on(JLC_ROW, JLC_ROW, nq, mk, nbk, -1.0, y2.d(), nbk, vc.d(), mk, 1.0, xt, ldx, JLC_MODE_FP64,
                          s, &info)) != JLC_OK)
            return rc;
    }
    JLC_HIP_CHECK(hipStreamSynchronize(s));  // the workspace is freed on return
    return JLC_OK;
}
}  // namespace

int jlc_dorgqr_dev(int64_t m, int64_t n, const double* a, int64_t lda, const double* tau, int64_t nb, double* q,
                   int64_t ldq, void* stream) {
    if (m < 0 || n < 0 || n > m || m > INT32_MAX) return fail(JLC_ERR_INVALID_ARG, "need 0 <= n <= m");
    if (nb < 1 || nb > 128) return fail(JLC_ERR_INVALID_ARG, "panel width nb must be in [1, 128]");
    if (n == 0) return JLC_OK;
    if (!a || !tau || !q || lda < m || ldq < m) return fail(JLC_ERR_INVALID_ARG, "bad pointers or leading dims");
    int rc = ensure_init();
    if (rc != JLC_OK) return rc;
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : t_ctx.stream;
    // Q^T (n x m, row-major, ld = ldq) starts as [I 0].
    JLC_HIP_CHECK(jlc::launch_fill_eye(q, ldq, (int)n, (int)m, 1.0, s));
    return apply_q_panels(m, n, a, lda, tau, nb, q, ldq, n, true, s);
}

int jlc_dtsqr_dev(int64_t m, int64_t n, const double* a, int64_t lda, double* q, int64_t ldq, double* r, int64_t ldr,
                  void* stream, jlc_stats* stats) {
    double t0 = now_ms();
    if (m < 0 || n < 0 || m > INT32_MAX || n > INT32_MAX) return fail(JLC_ERR_INVALID_ARG, "bad dimensions");
    const int64_t k = m < n ? m : n;
    jlc::TileInfo info;
    if (k == 0) {
        fill_stats(stats, info, JLC_KERNEL_NONE, 0.0, now_ms() - t0);
        return JLC_OK;
    }
    if (!a || !q || !r || lda < m || ldq < m || ldr < k)
        return fail(JLC_ERR_INVALID_ARG, "a, q, r must not be null; need lda >= m, ldq >= m, ldr >= min(m, n)");
    int rc = ensure_init();
    if (rc != JLC_OK) return rc;
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : t_ctx.stream;
    if (stats) JLC_HIP_CHECK(hipEventRecord(t_ctx.ev0, s));
    const int64_t nbq = n < 128 ? n : 128;
    DevBuf w, tb;
    hipError_t e = w.alloc((size_t)m * n * 8);
    if (e != hipSuccess) return fail(JLC_ERR_ALLOC, "TSQR workspace: %s", hipGetErrorString(e));
    JLC_HIP_CHECK(hipMemcpy2DAsync(w.d(), (size_t)m * 8, a, (size_t)lda * 8, (size_t)m * 8, (size_t)n,
                                   hipMemcpyDeviceToDevice, s));
    if (m <= n) {
        // HouseholderQR.decomposeThin (TallSkinnyQR.java:34-36): Q m x m, R m x n.
        if ((e = tb.alloc((size_t)k * 8)) != hipSuccess) return fail(JLC_ERR_ALLOC, "TSQR tau: %s", hipGetErrorString(e));
        if ((rc = jlc_dgeqrf_dev(m, n, w.d(), m, tb.d(), nbq, JLC_MODE_FP64, s, nullptr)) != JLC_OK) return rc;
        if ((rc = jlc_dorgqr_dev(m, k, w.d(), m, tb.d(), k < 128 ? k : 128, q, ldq, s)) != JLC_OK) return rc;
        JLC_HIP_CHECK(jlc::launch_copy_upper_blocks(w.d(), m, 0, (int)k, (int)k, (int)n, r, ldr, s));
    } else {
        // Row blocks of max(n, 64) rows, a thin QR of each, a thin QR of the R factors stacked in block
        // order, Q's block rows = the block's Q times its rows of the stack's Q (TallSkinnyQR.java:38-88).
        const int64_t b = n > 64 ? n : 64, nblk = (m + b - 1) / b;
        const int64_t rows_last = m - (nblk - 1) * b;
        const int64_t srows = (nblk - 1) * n + (rows_last < n ? rows_last : n);
        DevBuf stk, tau2, q2, dd;
        e = tb.alloc((size_t)nblk * n * 8);
        if (e == hipSuccess) e = stk.alloc((size_t)srows * n * 8);
        if (e == hipSuccess) e = tau2.alloc((size_t)n * 8);
        if (e == hipSuccess) e = q2.alloc((size_t)srows * n * 8);
        if (e != hipSuccess) return fail(JLC_ERR_ALLOC, "TSQR workspace: %s", hipGetErrorString(e));
        const bool batched = b <= 256 && b * n <= jlc::QR_BATCH_MAX_MN && b * n <= jlc::QR_APPLY_MAX_MNX;
        if (batched) {
            std::vector<jlc_qr_desc> qd((size_t)nblk);
            for (int64_t t = 0; t < nblk; t++) {
                const int64_t rows = t + 1 < nblk ? b : rows_last;
                qd[(size_t)t] = jlc_qr_desc{(int32_t)rows, (int32_t)n, m, w.d() + t * b, tb.d() + t * n};
            }
            if ((rc = jlc_dgeqrf_batched_dev(nblk, qd.data(), s, nullptr)) != JLC_OK) return rc;
        } else {
            for (int64_t t = 0; t < nblk; t++) {
                const int64_t rows = t + 1 < nblk ? b : rows_last;
                if ((rc = jlc_dgeqrf_dev(rows, n, w.d() + t * b, m, tb.d() + t * n, nbq, JLC_MODE_FP64, s, nullptr)) !=
                    JLC_OK)
                    return rc;
            }
        }
        JLC_HIP_CHECK(jlc::launch_copy_upper_blocks(w.d(), m, b, (int)n, (int)srows, (int)n, stk.d(), srows, s));
        if ((rc = jlc_dgeqrf_dev(srows, n, stk.d(), srows, tau2.d(), nbq, JLC_MODE_FP64, s, nullptr)) != JLC_OK)
            return rc;
        JLC_HIP_CHECK(jlc::launch_copy_upper_blocks(stk.d(), srows, 0, (int)n, (int)n, (int)n, r, ldr, s));
        if ((rc = jlc_dorgqr_dev(srows, n, stk.d(), srows, tau2.d(), nbq, q2.d(), srows, s)) != JLC_OK) return rc;
        // Q's rows of block t start as [Q2 rows t*n.. of the block's R; 0], then X <- Q_t X.
        JLC_HIP_CHECK(jlc::launch_spread_blocks(q2.d(), srows, (int)n, (int)b, (int)m, (int)n, (int)srows, q, ldq, s));
        if (batched) {
            std::vector<jlc::QrApplyDesc> ad((size_t)nblk);
            for (int64_t t = 0; t < nblk; t++) {
                const int64_t rows = t + 1 < nblk ? b : rows_last;
                ad[(size_t)t] = jlc::QrApplyDesc{(int32_t)rows, (int32_t)n, (int32_t)(rows < n ? rows : n), 0, m, ldq,
                                                 w.d() + t * b, tb.d() + t * n, q + t * b};
            }
            if ((e = dd.alloc(ad.size() * sizeof(jlc::QrApplyDesc))) != hipSuccess)
                return fail(JLC_ERR_ALLOC, "TSQR descriptors: %s", hipGetErrorString(e));
            JLC_HIP_CHECK(hipMemcpyAsync(dd.p, ad.data(), ad.size() * sizeof(jlc::QrApplyDesc), hipMemcpyHostToDevice, s));
            JLC_HIP_CHECK(jlc::launch_qr_batched_apply_q(static_cast<const jlc::QrApplyDesc*>(dd.p), (int)nblk,
                                                         (int)(b * n), s));
        } else {
            for (int64_t t = 0; t < nblk; t++) {
                const int64_t rows = t + 1 < nblk ? b : rows_last;
                if ((rc = apply_q_panels(rows, rows < n ? rows : n, w.d() + t * b, m, tb.d() + t * n, nbq, q + t * b,
                                         ldq, n, false, s)) != JLC_OK)
                    return rc;
            }
        }
    }
    double kms = 0.0;
    if (stats) {
        JLC_HIP_CHECK(hipEventRecord(t_ctx.ev1, s));
        JLC_HIP_CHECK(hipEventSynchronize(t_ctx.ev1));
        float f = 0.0f;
        JLC_HIP_CHECK(hipEventElapsedTime(&f, t_ctx.ev0, t_ctx.ev1));
        kms = f;
    } else {
        JLC_HIP_CHECK(hipStreamSynchronize(s));  // the workspace is freed on return
    }
    fill_stats(stats, info, JLC_KERNEL_QR, kms, now_ms() - t0);
    return JLC_OK;
}

int jlc_dtsqr(int64_t m, int64_t n, const double* a, int64_t lda, double* q, int64_t ldq, double* r, int64_t ldr,
              jlc_stats* stats) {
    // Host row-major operands (the reference's Matrix storage): A (m x n, lda >= n), Q (m x k, ldq >= k),
    // R (k x n, ldr >= n), k = min(m, n); transposed into column-major device copies and back.
    double t0 = now_ms();
    if (m < 0 || n < 0 || m > INT32_MAX || n > INT32_MAX) return fail(JLC_ERR_INVALID_ARG, "bad dimensions");
    const int64_t k = m < n ? m : n;
    jlc::TileInfo info;
    if (k == 0) {
        fill_stats(stats, info, JLC_KERNEL_NONE, 0.0, now_ms() - t0);
        return JLC_OK;
    }
    if (!a || !q || !r || lda < n || ldq < k || ldr < n)
        return fail(JLC_ERR_INVALID_ARG, "a, q, r must not be null; need lda >= n, ldq >= min(m, n), ldr >= n");
    int rc = ensure_init();
    if (rc != JLC_OK) return rc;
    std::vector<double> h((size_t)(m * n));
    for (int64_t i = 0; i < m; i++)
        for (int64_t j = 0; j < n; j++) h[(size_t)(j * m + i)] = a[i * lda + j];
    DevBuf da, dq, dr;
    hipError_t e = da.alloc((size_t)m * n * 8);
    if (e == hipSuccess) e = dq.alloc((size_t)m * k * 8);
    if (e == hipSuccess) e = dr.alloc((size_t)k * n * 8);
    if (e != hipSuccess) return fail(JLC_ERR_ALLOC, "TSQR buffers: %s", hipGetErrorString(e));
    JLC_HIP_CHECK(hipMemcpy(da.p, h.data(), h.size() * 8, hipMemcpyHostToDevice));
    jlc_stats st{};
    if ((rc = jlc_dtsqr_dev(m, n, da.d(), m, dq.d(), m, dr.d(), k, t_ctx.stream, &st)) != JLC_OK) return rc;
    std::vector<double> hq((size_t)(m * k)), hr((size_t)(k * n));
    JLC_HIP_CHECK(hipMemcpy(hq.data(), dq.p, hq.size() * 8, hipMemcpyDeviceToHost));
    JLC_HIP_CHECK(hipMemcpy(hr.data(), dr.p, hr.size() * 8, hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < m; i++)
        for (int64_t j = 0; j < k; j++) q[i * ldq + j] = hq[(size_t)(j * m + i)];
    for (int64_t i = 0; i < k; i++)
        for (int64_t j = 0; j < n; j++) r[i * ldr + j] = hr[(size_t)(j * k + i)];
    fill_stats(stats, info, JLC_KERNEL_QR, st.kernel_ms, now_ms() - t0);
    return JLC_OK;
}

namespace {
int fro_norm(const double* x, int64_t ld, int64_t rows, int64_t cols, hipStream_t s, double* out) {
    const int parts = 1024;
    DevBuf p;
    if (p.alloc(parts * 8) != hipSuccess) return fail(JLC_ERR_ALLOC, "norm workspace");
    JLC_HIP_CHECK(jlc::launch_sumsq_partial(x, ld, (int)rows, (int)cols, p.d(), parts, s));
    std::vector<double> h(parts);
    JLC_HIP_CHECK(hipMemcpyAsync(h.data(), p.d(), parts * 8, hipMemcpyDeviceToHost, s));
    JLC_HIP_CHECK(hipStreamSynchronize(s));
    double sum = 0.0;
    for (double v : h) sum += v;
    *out = std::sqrt(sum);
    return JLC_OK;
}
}  // namespace

int jlc_dqr_errors_dev(int64_t m, int64_t n, const double* a_orig, int64_t lda_orig, const double* qr, int64_t lda,
                       const double* tau, int64_t nb, double* recon_err, double* orth_err) {
    if (!recon_err || !orth_err) return fail(JLC_ERR_INVALID_ARG, "outputs must not be null");
    if (n > m || n <= 0) return fail(JLC_ERR_INVALID_ARG, "need 0 < n <= m");
    int rc = ensure_init();
    if (rc != JLC_OK) return rc;
    hipStream_t s = t_ctx.stream;
    DevBuf q, r, d, e;
    hipError_t he = q.alloc((size_t)m * n * 8);
    if (he == hipSuccess) he = r.alloc((size_t)n * n * 8);
    if (he == hipSuccess) he = d.alloc((size_t)m * n * 8);
    if (he == hipSuccess) he = e.alloc((size_t)n * n * 8);
    if (he != hipSuccess) return fail(JLC_ERR_ALLOC, "error-check workspace: %s", hipGetErrorString(he));
    if ((rc = jlc_dorgqr_dev(m, n, qr, lda, tau, nb, q.d(), m, s)) != JLC_OK) return rc;
    JLC_HIP_CHECK(jlc::launch_extract_r(qr, lda, (int)n, r.d(), s));
    // D^T = A^T - R^T Q^T  (all n x m row-major views of column-major m x n matrices).
    JLC_HIP_CHECK(hipMemcpy2DAsync(d.d(), (size_t)m * 8, a_orig, (size_t)lda_orig * 8, (size_t)m * 8, (size_t)n,
                                   hipMemcpyDeviceToDevice, s));
    jlc::TileInfo info;
    if ((rc = gemm_on(JLC_COL, JLC_ROW, n, m, n, -1.0, r.d(), n, q.d(), m, 1.0, d.d(), m, JLC_MODE_FP64, s, &info)) !=
        JLC_OK)
        return rc;
    double dn = 0.0, an = 0.0, en = 0.0;
    if ((rc = fro_norm(d.d(), m, n, m, s, &dn)) != JLC_OK) return rc;
    if ((rc = fro_norm(a_orig, lda_orig, n, m, s, &an)) != JLC_OK) return rc;
    // E = Q^T Q - I.
    JLC_HIP_CHECK(jlc::launch_fill_eye(e.d(), n, (int)n, (int)n, -1.0, s));
    if ((rc = gemm_on(JLC_ROW, JLC_COL, n, n, m, 1.0, q.d(), m, q.d(), m, 1.0, e.d(), n, JLC_MODE_FP64, s, &info)) !=
        JLC_OK)
        return rc;
    if ((rc = fro_norm(e.d(), n, n, n, s, &en)) != JLC_OK) return rc;
    *recon_err = an > 0.0 ? dn / an : dn;
    *orth_err = en;
    return JLC_OK;
}

static_assert(sizeof(jlc_qr_desc) == sizeof(jlc::QrBatchDesc), "QR descriptor layouts must match");

// Matrices are grouped into size classes by the LDS their tile needs, one
// launch per class, so small matrices run several wavefronts per CU instead of
// being held to the occupancy of the largest matrix of the batch.
struct jlc_qr_batched_plan {
    static constexpr int NCLASS = 7;
    static constexpr int NREG = 6;  // register-kernel buckets: m <= 32, 64, ..., 192 (rows per thread 4 .. 24)
    int reg_off[NREG] = {}, reg_cnt[NREG] = {}, reg_n[NREG] = {};
    int count = 0;
    int max_mn = 0;
    int cls_off[NCLASS] = {}, cls_cnt[NCLASS] = {}, cls_mn[NCLASS] = {};
    int cls_threads[NCLASS] = {};  // threads per matrix: 64 (one wavefront), 256, 512 or 1024
    jlc::QrBatchDesc* d_descs = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {
// LDS entries (doubles) of each class: 16, 8, 5, 4, 3, 2 and 1 wavefronts per CU by LDS.
constexpr int QR_CLASS_MN[jlc_qr_batched_plan::NCLASS] = {1280, 2560, 4096, 5120, 6826, 10240, jlc::QR_BATCH_MAX_MN};

hipError_t launch_qr_plan(const jlc_qr_batched_plan* p, hipStream_t s) {
    for (int b = 0; b < jlc_qr_batched_plan::NREG; b++) {
        if (!p->reg_cnt[b]) continue;
        hipError_t e = jlc::launch_qr_batched_reg(p->d_descs + p->reg_off[b], p->reg_cnt[b], 4 * (b + 1), p->reg_n[b], s);
        if (e != hipSuccess) return e;
    }
    for (int c = 0; c < jlc_qr_batched_plan::NCLASS; c++) {
        if (!p->cls_cnt[c]) continue;
        hipError_t e = jlc::launch_qr_batched_small(p->d_descs + p->cls_off[c], p->cls_cnt[c], p->cls_mn[c],
                                                    p->cls_threads[c], s);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}
}  // namespace

int jlc_qr_batched_plan_create(int64_t count, const jlc_qr_desc* descs, jlc_qr_batched_plan** plan) {
    if (!plan) return fail(JLC_ERR_INVALID_ARG, "plan must not be null");
    *plan = nullptr;
    if (count < 0 || count > INT32_MAX || (count && !descs)) return fail(JLC_ERR_INVALID_ARG, "bad descriptors");
    int max_mn = 1;
    for (int64_t i = 0; i < count; i++) {
        const jlc_qr_desc& d = descs[i];
        if (d.m < 1 || d.n < 1 || d.n > 256 || d.lda < d.m || !d.a || !d.tau)
            return fail(JLC_ERR_INVALID_ARG, "descriptor %lld: need m, n >= 1, n <= 256, lda >= m", (long long)i);
        if ((int64_t)d.m * d.n > jlc::QR_BATCH_MAX_MN)
            return fail(JLC_ERR_INVALID_ARG, "descriptor %lld: m*n = %lld exceeds the %d-entry LDS tile",
                        (long long)i, (long long)d.m * d.n, jlc::QR_BATCH_MAX_MN);
        max_mn = std::max(max_mn, d.m * d.n);
    }
    int rc = ensure_init();
    if (rc != JLC_OK) return rc;
    jlc_qr_batched_plan* p = new jlc_qr_batched_plan;
    p->count = (int)count;
    p->max_mn = max_mn;
    // Stable order by size class; each class launches with its own largest tile. Matrices that fit the
    // register kernel (m <= qr_reg_max_m <= 192, n <= 128) go first, bucketed by rows per thread.
    std::vector<jlc_qr_desc> sorted;
    sorted.reserve((size_t)count);
    const int reg_max_m = std::min(192, jlc::tune_snapshot().qr_reg_max_m);
    auto reg_bucket = [&](const jlc_qr_desc& d) { return d.m <= reg_max_m && d.n <= 128 ? (d.m - 1) / 32 : -1; };
    for (int b = 0; b < jlc_qr_batched_plan::NREG; b++) {
        p->reg_off[b] = (int)sorted.size();
        int mx = 0;
        for (int64_t i = 0; i < count; i++)
            if (reg_bucket(descs[i]) == b) {
                sorted.push_back(descs[i]);
                mx = std::max(mx, descs[i].n);
            }
        p->reg_cnt[b] = (int)sorted.size() - p->reg_off[b];
        p->reg_n[b] = mx;
    }
    for (int c = 0, off = (int)sorted.size(); c < jlc_qr_batched_plan::NCLASS; c++) {
        const int lo = c ? QR_CLASS_MN[c - 1] : 0, hi = QR_CLASS_MN[c];
        // Small classes run one wavefront per matrix, the large ones qr_block_threads threads (the
        // boundary and the width are the tuning table's qr_wave_max_lds and qr_block_threads).
        const jlc::TuneTable tt = jlc::tune_snapshot();
        int mx = 0;
        for (int64_t i = 0; i < count; i++) {
            const int mn = jlc::qr_batched_lds_entries(descs[i].m, descs[i].n);
            if (reg_bucket(descs[i]) < 0 && mn > lo && mn <= hi) {
                sorted.push_back(descs[i]);
                mx = std::max(mx, mn);
            }
        }
        p->cls_off[c] = off;
        p->cls_cnt[c] = (int)sorted.size() - off;
        p->cls_mn[c] = mx;
        p->cls_threads[c] = hi <= tt.qr_wave_max_lds ? 64 : tt.qr_block_threads;
        off = (int)sorted.size();
    }
    hipError_t e = hipMalloc(&p->d_descs, std::max<size_t>(1, (size_t)count) * sizeof(jlc::QrBatchDesc));
    if (e == hipSuccess && count)
        e = hipMemcpy(p->d_descs, sorted.data(), (size_t)count * sizeof(jlc::QrBatchDesc), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        jlc_qr_batched_plan_destroy(p);
        return fail(JLC_ERR_ALLOC, "QR plan upload failed: %s", hipGetErrorString(e));
    }
    *plan = p;
    return JLC_OK;
}

int jlc_qr_batched_plan_replay(jlc_qr_batched_plan* plan, void* stream, jlc_stats* stats) {
    if (!plan) return fail(JLC_ERR_INVALID_ARG, "plan must not be null");
    double t0 = now_ms();
    int rc = ensure_init();
    if (rc != JLC_OK) return rc;
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : t_ctx.stream;
    if (plan->count == 0) return JLC_OK;
    if (!plan->exec) {
        hipStream_t cap;
        JLC_HIP_CHECK(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
        hipError_t e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
        if (e == hipSuccess) e = launch_qr_plan(plan, cap);
        hipGraph_t graph = nullptr;
        hipError_t e2 = hipStreamEndCapture(cap, &graph);
        if (e == hipSuccess) e = e2;
        if (e == hipSuccess) e = hipGraphInstantiate(&plan->exec, graph, nullptr, nullptr, 0);
        (void)hipStreamDestroy(cap);
        if (e != hipSuccess) {
            if (graph) (void)hipGraphDestroy(graph);
            plan->exec = nullptr;
            return hip_fail(e, "batched QR graph capture");
        }
        plan->graph = graph;
    }
    if (stats) JLC_HIP_CHECK(hipEventRecord(t_ctx.ev0, s));
    JLC_HIP_CHECK(hipGraphLaunch(plan->exec, s));
    if (stats) {
        JLC_HIP_CHECK(hipEventRecord(t_ctx.ev1, s));
        JLC_HIP_CHECK(hipEventSynchronize(t_ctx.ev1));
        float f = 0.0f;
        JLC_HIP_CHECK(hipEventElapsedTime(&f, t_ctx.ev0, t_ctx.ev1));
        jlc::TileInfo info;
        info.launches = 1;
        fill_stats(stats, info, JLC_KERNEL_QR, f, now_ms() - t0);
    }
    return JLC_OK;
}

int jlc_qr_batched_plan_destroy(jlc_qr_batched_plan* plan) {
    if (!plan) return JLC_OK;
    if (g_alive.load()) {
        if (plan->exec) (void)hipGraphExecDestroy(plan->exec);
        if (plan->graph) (void)hipGraphDestroy(plan->graph);
        if (plan->d_descs) (void)hipFree(plan->d_descs);
    }
    delete plan;
    return JLC_OK;
}

int jlc_dgeqrf_batched_dev(int64_t count, const jlc_qr_desc* descs, void* stream, jlc_stats* stats) {
    jlc_qr_batched_plan* plan = nullptr;
    int rc = jlc_qr_batched_plan_create(count, descs, &plan);
    if (rc != JLC_OK) return rc;
    double t0 = now_ms();
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : t_ctx.stream;
    if (stats) {
        hipError_t e = hipEventRecord(t_ctx.ev0, s);
        if (e != hipSuccess) rc = hip_fail(e, "hipEventRecord");
    }
    if (rc == JLC_OK && count) {
        hipError_t e = launch_qr_plan(plan, s);
        if (e != hipSuccess) rc = hip_fail(e, "batched QR launch");
    }
    if (rc == JLC_OK) {
        float f = 0.0f;
        hipError_t e = hipSuccess;
        if (stats) {
            e = hipEventRecord(t_ctx.ev1, s);
            if (e == hipSuccess) e = hipEventSynchronize(t_ctx.ev1);
            if (e == hipSuccess) e = hipEventElapsedTime(&f, t_ctx.ev0, t_ctx.ev1);
        }
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) rc = hip_fail(e, "batched QR");
        jlc::TileInfo info;
        info.launches = 1;
        fill_stats(stats, info, JLC_KERNEL_QR, f, now_ms() - t0);
    }
    jlc_qr_batched_plan_destroy(plan);
    return rc;
}

struct jlc_timer {
    hipEvent_t start = nullptr, stop = nullptr;
};

int jlc_timer_create(jlc_timer** timer) {
    if (!timer) return fail(JLC_ERR_INVALID_ARG, "timer must not be null");
    int rc = ensure_init();
    if (rc != JLC_OK) return rc;
    jlc_timer* t = new jlc_timer;
    hipError_t e = hipEventCreate(&t->start);
    if (e == hipSuccess) e = hipEventCreate(&t->stop);
    if (e != hipSuccess) {
        if (t->start) (void)hipEventDestroy(t->start);
        delete t;
        return hip_fail(e, "hipEventCreate");
    }
    *timer = t;
    return JLC_OK;
}

int jlc_timer_start(jlc_timer* timer, void* stream) {
    if (!timer) return fail(JLC_ERR_INVALID_ARG, "timer must not be null");
    int rc = ensure_init();
    if (rc != JLC_OK) return rc;
    JLC_HIP_CHECK(hipEventRecord(timer->start, stream ? static_cast<hipStream_t>(stream) : t_ctx.stream));
    return JLC_OK;
}

int jlc_timer_stop(jlc_timer* timer, void* stream, double* elapsed_ms) {
    if (!timer || !elapsed_ms) return fail(JLC_ERR_INVALID_ARG, "timer and elapsed_ms must not be null");
    int rc = ensure_init();
    if (rc != JLC_OK) return rc;
    JLC_HIP_CHECK(hipEventRecord(timer->stop, stream ? static_cast<hipStream_t>(stream) : t_ctx.stream));
    JLC_HIP_CHECK(hipEventSynchronize(timer->stop));
    float f = 0.0f;
    JLC_HIP_CHECK(hipEventElapsedTime(&f, timer->start, timer->stop));
    *elapsed_ms = f;
    return JLC_OK;
}

int jlc_timer_destroy(jlc_timer* timer) {
    if (!timer) return JLC_OK;
    if (g_alive.load()) {
        (void)hipEventDestroy(timer->start);
        (void)hipEventDestroy(timer->stop);
    }
    delete timer;
    return JLC_OK;
}

}  // extern "C"
