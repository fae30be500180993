// bandwidth.hip -- streaming-bandwidth probes of the GPU (libsfprobe.so).
//
// Measurement tooling for the roofline figures, kept out of libsoapfilm.so:
// the achievable HBM bandwidth of plain streams against which the stencil
// kernels are judged, and the copy of known size that scales the
// FETCH_SIZE / WRITE_SIZE counters of rocprofv3.  Every stream moves float4
// elements; a block covers a contiguous chunk of U x 256 elements with U
// independent 16-byte accesses in flight per thread.
//
//   mode 0  copy            dst = src                   (cached loads and stores)
//   mode 1  copy, streaming dst = src                   (non-temporal loads and stores)
//   mode 2  read            per-block sum of src        (one float4 written per block)
//   mode 3  write           dst = constant              (non-temporal stores)
//   mode 4  3-read/2-write  d = a + b, e = b + c        (the stencil's read/write mix)
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <cstdio>

namespace {

typedef float f4v __attribute__((ext_vector_type(4)));

template <int U, bool NT>
__global__ __launch_bounds__(256) void copy_kernel(const f4v* __restrict__ src, f4v* __restrict__ dst, size_t n4) {
    const size_t base = blockIdx.x * static_cast<size_t>(U * 256) + threadIdx.x;
    f4v v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const size_t k = base + u * 256;
        if (k < n4) v[u] = NT ? __builtin_nontemporal_load(src + k) : src[k];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const size_t k = base + u * 256;
        if (k < n4) {
            if (NT)
                __builtin_nontemporal_store(v[u], dst + k);
            else
                dst[k] = v[u];
        }
    }
}

template <int U>
__global__ __launch_bounds__(256) void read_kernel(const f4v* __restrict__ src, f4v* __restrict__ sums, size_t n4) {
    const size_t base = blockIdx.x * static_cast<size_t>(U * 256) + threadIdx.x;
    f4v acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const size_t k = base + u * 256;
        if (k < n4) acc += src[k];
    }
    // The block's sum keeps every load live; the reduction is cheap next to
    // the stream.
    __shared__ f4v red[256];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (static_cast<int>(threadIdx.x) < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) sums[blockIdx.x] = red[0];
}

template <int U>
__global__ __launch_bounds__(256) void write_kernel(f4v* __restrict__ dst, size_t n4, float value) {
    const size_t base = blockIdx.x * static_cast<size_t>(U * 256) + threadIdx.x;
    const f4v v = {value, value, value, value};
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const size_t k = base + u * 256;
        if (k < n4) __builtin_nontemporal_store(v, dst + k);
    }
}

template <int U>
__global__ __launch_bounds__(256) void mix_kernel(const f4v* __restrict__ a, const f4v* __restrict__ b,
                                                  const f4v* __restrict__ c, f4v* __restrict__ d, f4v* __restrict__ e,
                                                  size_t n4) {
    const size_t base = blockIdx.x * static_cast<size_t>(U * 256) + threadIdx.x;
    f4v va[U], vb[U], vc[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const size_t k = base + u * 256;
        if (k < n4) {
            va[u] = a[k];
            vb[u] = b[k];
            vc[u] = c[k];
        }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const size_t k = base + u * 256;
        if (k < n4) {
            d[k] = va[u] + vb[u];
            e[k] = vb[u] + vc[u];
        }
    }
}

template <int U>
hipError_t launch(int mode, f4v* const* buf, size_t n4, hipStream_t s) {
    const unsigned blocks = static_cast<unsigned>((n4 + U * 256 - 1) / (U * 256));
    switch (mode) {
        case 0: hipLaunchKernelGGL((copy_kernel<U, false>), dim3(blocks), dim3(256), 0, s, buf[0], buf[1], n4); break;
        case 1: hipLaunchKernelGGL((copy_kernel<U, true>), dim3(blocks), dim3(256), 0, s, buf[0], buf[1], n4); break;
        case 2: hipLaunchKernelGGL(read_kernel<U>, dim3(blocks), dim3(256), 0, s, buf[0], buf[1], n4); break;
        case 3: hipLaunchKernelGGL(write_kernel<U>, dim3(blocks), dim3(256), 0, s, buf[1], n4, 1.f); break;
        case 4:
            hipLaunchKernelGGL(mix_kernel<U>, dim3(blocks), dim3(256), 0, s, buf[0], buf[1], buf[2], buf[3], buf[4], n4);
            break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_u(int mode, int unroll, f4v* const* buf, size_t n4, hipStream_t s) {
    switch (unroll) {
        case 4: return launch<4>(mode, buf, n4, s);
        case 8: return launch<8>(mode, buf, n4, s);
        case 16: return launch<16>(mode, buf, n4, s);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace

// Runs stream `mode` with `unroll` (4, 8 or 16) accesses in flight per
// thread over arrays of n_bytes each, `iters` times after one untimed
// launch; *avg_ms receives the average launch time (HIP events) and
// *bytes_per_launch the bytes the stream reads and writes per launch.
// Returns 0, or -1 on bad arguments, -2 on a HIP error (message on stderr).
extern "C" int sfp_stream(int32_t mode, int32_t unroll, size_t n_bytes, int32_t iters, double* avg_ms,
                          double* bytes_per_launch) {
    if (mode < 0 || mode > 4 || (unroll != 4 && unroll != 8 && unroll != 16) || n_bytes < 16 || iters < 1)
        return -1;
    const size_t n4 = n_bytes / 16;
    const int nbuf = mode == 4 ? 5 : 2;
    f4v* buf[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    hipError_t e = hipSuccess;
    for (int k = 0; k < nbuf && e == hipSuccess; ++k) {
        // the read stream's second buffer holds one float4 per block only
        const size_t bytes = (mode == 2 && k == 1) ? (n4 / (static_cast<size_t>(unroll) * 256) + 1) * 16 : n4 * 16;
        e = hipMalloc(reinterpret_cast<void**>(&buf[k]), bytes);
        if (e == hipSuccess && !(mode == 2 && k == 1)) e = hipMemset(buf[k], 0, bytes);
    }
    hipEvent_t t0 = nullptr, t1 = nullptr;
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventCreate(&t0);
    if (e == hipSuccess) e = hipEventCreate(&t1);
    if (e == hipSuccess) e = launch_u(mode, unroll, buf, n4, nullptr);
    if (e == hipSuccess) e = hipEventRecord(t0, nullptr);
    for (int k = 0; k < iters && e == hipSuccess; ++k) e = launch_u(mode, unroll, buf, n4, nullptr);
    if (e == hipSuccess) e = hipEventRecord(t1, nullptr);
    if (e == hipSuccess) e = hipEventSynchronize(t1);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, t0, t1);
    if (t0) (void)hipEventDestroy(t0);
    if (t1) (void)hipEventDestroy(t1);
    for (f4v* p : buf)
        if (p) (void)hipFree(p);
    if (e != hipSuccess) {
        std::fprintf(stderr, "sfp_stream: %s\n", hipGetErrorString(e));
        return -2;
    }
    if (avg_ms) *avg_ms = static_cast<double>(ms) / iters;
    if (bytes_per_launch) {
        const double b = static_cast<double>(n4) * 16.0;
        *bytes_per_launch = mode == 4 ? 5.0 * b : (mode == 2 || mode == 3 ? b : 2.0 * b);
    }
    return 0;
}
