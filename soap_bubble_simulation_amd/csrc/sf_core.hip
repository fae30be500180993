// sf_core.hip -- error state, device probing and version of the C ABI.
#include <cstdarg>
#include <cstddef>
#include <cstring>

#include "sf_common.hpp"

namespace sf {

static thread_local char g_last_error[512] = "";

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
    va_end(ap);
}

const char* last_error() { return g_last_error; }

void require_device() {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n < 1) {
        set_error("no HIP device available (%s); the soapfilm library runs only on gfx950 GPUs",
                  e == hipSuccess ? "0 devices" : hipGetErrorString(e));
        throw Error{SF_ERR_DEVICE};
    }
}

}  // namespace sf

namespace sf {

// Plain float4 copy: the reference kernel for achievable HBM bandwidth and
// for scaling the FETCH_SIZE / WRITE_SIZE counters of the other kernels.
// One block per contiguous 16 KiB chunk, four independent 16-byte loads in
// flight per thread, ordinary (cached) loads and stores like the stencil
// kernels'.  Of the copy shapes tried (grid-stride loops of 1-64 Ki blocks,
// unrolled or not, non-temporal or not) this is the fastest with cached
// accesses.
constexpr int kCopyUnroll = 4;
constexpr size_t kCopyChunk = kCopyUnroll * 256;

__global__ __launch_bounds__(256) void copy_probe_kernel(const float4* __restrict__ src,
                                                          float4* __restrict__ dst, size_t n4) {
    const size_t base = blockIdx.x * kCopyChunk + threadIdx.x;
    float4 v[kCopyUnroll];
#pragma unroll
    for (int u = 0; u < kCopyUnroll; ++u)
        if (base + u * 256 < n4) v[u] = src[base + u * 256];
#pragma unroll
    for (int u = 0; u < kCopyUnroll; ++u)
        if (base + u * 256 < n4) dst[base + u * 256] = v[u];
}

}  // namespace sf

static_assert(sizeof(sf_params) == 16 * 8 + 2 * 4, "sf_params layout");
static_assert(offsetof(sf_params, flags) == 128, "sf_params layout");
static_assert(sizeof(sf_stats) == 5 * 8 + 8 + 4 + 4 + 8, "sf_stats layout");
static_assert(sizeof(sf_camera) == 24, "sf_camera layout");

extern "C" int sf_copy_probe(size_t n_bytes, int32_t iters, double* avg_ms) {
    return sf::guard([&] {
        SF_REQUIRE(n_bytes >= 16 && iters >= 1, SF_ERR_INVALID, "bad arguments");
        sf::require_device();
        const size_t n4 = n_bytes / 16;
        void *a = nullptr, *b = nullptr;
        SF_HIP(hipMalloc(&a, n4 * 16));
        hipError_t e = hipMalloc(&b, n4 * 16);
        if (e != hipSuccess) {
            (void)hipFree(a);
            SF_HIP(e);
        }
        hipEvent_t t0, t1;
        (void)hipEventCreate(&t0);
        (void)hipEventCreate(&t1);
        (void)hipMemset(a, 0, n4 * 16);
        const unsigned blocks = static_cast<unsigned>((n4 + sf::kCopyChunk - 1) / sf::kCopyChunk);
        hipLaunchKernelGGL(sf::copy_probe_kernel, dim3(blocks), dim3(256), 0, nullptr,
                           static_cast<const float4*>(a), static_cast<float4*>(b), n4);
        (void)hipEventRecord(t0, nullptr);
        for (int k = 0; k < iters; ++k)
            hipLaunchKernelGGL(sf::copy_probe_kernel, dim3(blocks), dim3(256), 0, nullptr,
                               static_cast<const float4*>(a), static_cast<float4*>(b), n4);
        (void)hipEventRecord(t1, nullptr);
        e = hipEventSynchronize(t1);
        float ms = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, t0, t1);
        (void)hipEventDestroy(t0);
        (void)hipEventDestroy(t1);
        (void)hipFree(a);
        (void)hipFree(b);
        SF_HIP(e);
        if (avg_ms) *avg_ms = static_cast<double>(ms) / iters;
    });
}

extern "C" const char* sf_last_error(void) { return sf::last_error(); }

extern "C" const char* sf_version(void) { return "soapfilm-mi355x 0.1.0 (gfx950)"; }

extern "C" int sf_device_available(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n > 0 ? 1 : 0;
}
