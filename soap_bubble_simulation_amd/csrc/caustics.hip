// caustics.hip -- caustic-intensity map of the film thickness field.
//
// The consumer of the drainage output in the reference's frame
// (caustics.rs:441-457 dispatches caustics_compute.wgsl:117-126 on the
// thickness buffer): per grid point, a 9-point stencil with phi wrapped and
// theta clamped gives the gradient, the Laplacian and the Hessian
// determinant of the thickness (in index units), and
//   I = clamp((focusing * branch * refraction)^sharpness * intensity, 0, 10),
//   focusing = 1 / max(|1 + focal_length * lap|, 0.01),
//   branch = 1.5 if det < -branch_threshold else 1,
//   refraction = 1 + |grad| / 2      (caustics_compute.wgsl:89-114).
// FP32 like the shader, with its operation order and no contraction, so the
// branch decision matches any FP32 evaluation of the same expressions.
#include <hip/hip_runtime.h>

#include <cmath>

#include "drainage.hpp"

namespace sf {

struct CausticArgs {
    const float* hi;
    const float* lo;        // may be null (thickness = hi)
    int nt, np;
    float focal, intensity, sharpness, threshold;
    float* out;
};

__device__ inline float th_at(const CausticArgs& a, int x, int y) {
    int sx = x % a.np;
    if (sx < 0) sx += a.np;
    const int sy = min(max(y, 0), a.nt - 1);
    const size_t k = static_cast<size_t>(sy) * a.np + sx;
    return a.lo ? a.hi[k] + a.lo[k] : a.hi[k];
}

__global__ __launch_bounds__(256) void caustics_kernel(const CausticArgs a) {
#pragma clang fp contract(off)
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= a.np) return;
    const float c = th_at(a, x, y);
    const float l = th_at(a, x - 1, y), r = th_at(a, x + 1, y);
    const float u = th_at(a, x, y - 1), d = th_at(a, x, y + 1);
    const float ul = th_at(a, x - 1, y - 1), ur = th_at(a, x + 1, y - 1);
    const float dl = th_at(a, x - 1, y + 1), dr = th_at(a, x + 1, y + 1);
    const float gx = (r - l) / 2.0f;
    const float gy = (d - u) / 2.0f;
    const float lap = (l + r + u + d - 4.0f * c) / 1.0f;
    const float d2x = (l + r - 2.0f * c) / 1.0f;
    const float d2y = (u + d - 2.0f * c) / 1.0f;
    const float dxy = (dr - dl - ur + ul) / 4.0f;
    const float det = d2x * d2y - dxy * dxy;
    const float curv = -lap * a.focal;
    const float focusing = 1.0f / fmaxf(fabsf(1.0f - curv), 0.01f);
    const float branch = det < -a.threshold ? 1.5f : 1.0f;
    const float refraction = 1.0f + sqrtf(gx * gx + gy * gy) * 0.5f;
    float I = focusing * branch * refraction;
    I = powf(I, a.sharpness) * a.intensity;
    a.out[static_cast<size_t>(y) * a.np + x] = fminf(fmaxf(I, 0.0f), 10.0f);
}

static void launch_caustics(const CausticArgs& a, hipStream_t s) {
    hipLaunchKernelGGL(caustics_kernel, dim3((a.np + 255) / 256, a.nt), dim3(256), 0, s, a);
    SF_HIP(hipGetLastError());
}

static void check_caustic_params(const sf_caustic_params* p) {
    SF_REQUIRE(p != nullptr, SF_ERR_INVALID, "params is NULL");
    SF_REQUIRE(std::isfinite(p->focal_length) && std::isfinite(p->caustic_intensity) &&
                   std::isfinite(p->caustic_sharpness) && std::isfinite(p->branch_threshold),
               SF_ERR_INVALID, "caustic parameters must be finite");
}

}  // namespace sf

using namespace sf;

extern "C" int sf_caustic_params_default(sf_caustic_params* p) {
    return guard([&] {
        SF_REQUIRE(p != nullptr, SF_ERR_INVALID, "NULL argument");
        p->refractive_index = 1.33f;          // caustics.rs:56-78
        p->film_thickness_scale = 1e6f;
        p->light_dir[0] = 0.0f;
        p->light_dir[1] = -1.0f;
        p->light_dir[2] = 0.0f;
        p->light_intensity = 1.0f;
        p->focal_length = 0.1f;
        p->caustic_intensity = 2.0f;
        p->caustic_sharpness = 1.5f;
        p->branch_threshold = 0.001f;
        p->ground_y = -0.1f;
        p->bubble_radius = 0.025f;
        p->time = 0.0f;
    });
}

extern "C" int sf_caustics(sf_drainage* ctx, const sf_caustic_params* p, float* map_host) {
    return guard([&] {
        SF_REQUIRE(ctx != nullptr, SF_ERR_INVALID, "context is NULL");
        check_caustic_params(p);
        Drainage* d = reinterpret_cast<Drainage*>(ctx);
        d->materialize();
        if (d->caustic == nullptr) {
            void* q = nullptr;
            SF_HIP(hipMalloc(&q, d->n * sizeof(float)));
            d->caustic = static_cast<float*>(q);
        }
        CausticArgs a{d->hi[d->cur_hi], d->lo[d->cur], d->nt, d->np, p->focal_length, p->caustic_intensity,
                      p->caustic_sharpness, p->branch_threshold, d->caustic};
        launch_caustics(a, d->stream);
        if (map_host) {
            SF_HIP(hipMemcpyAsync(map_host, d->caustic, d->n * sizeof(float), hipMemcpyDeviceToHost, d->stream));
            SF_HIP(hipStreamSynchronize(d->stream));
        }
    });
}

extern "C" int sf_caustics_device_map(sf_drainage* ctx, const float** map) {
    return guard([&] {
        SF_REQUIRE(ctx != nullptr && map != nullptr, SF_ERR_INVALID, "NULL argument");
        Drainage* d = reinterpret_cast<Drainage*>(ctx);
        SF_REQUIRE(d->caustic != nullptr, SF_ERR_INVALID, "no caustic map computed yet");
        *map = d->caustic;
    });
}

extern "C" int sf_caustics_field(const sf_caustic_params* p, int32_t n_theta, int32_t n_phi, const float* thickness,
                                 float* map) {
    return guard([&] {
        check_caustic_params(p);
        SF_REQUIRE(n_theta >= 1 && n_phi >= 1 && thickness != nullptr && map != nullptr, SF_ERR_INVALID,
                   "bad arguments");
        require_device();
        const size_t n = static_cast<size_t>(n_theta) * n_phi;
        float *dh = nullptr, *dm = nullptr;
        SF_HIP(hipMalloc(&dh, n * sizeof(float)));
        hipError_t e = hipMalloc(&dm, n * sizeof(float));
        if (e != hipSuccess) {
            (void)hipFree(dh);
            SF_HIP(e);
        }
        try {
            SF_HIP(hipMemcpy(dh, thickness, n * sizeof(float), hipMemcpyHostToDevice));
            CausticArgs a{dh, nullptr, n_theta, n_phi, p->focal_length, p->caustic_intensity, p->caustic_sharpness,
                          p->branch_threshold, dm};
            launch_caustics(a, nullptr);
            SF_HIP(hipMemcpy(map, dm, n * sizeof(float), hipMemcpyDeviceToHost));
        } catch (...) {
            (void)hipFree(dh);
            (void)hipFree(dm);
            throw;
        }
        (void)hipFree(dh);
        (void)hipFree(dm);
    });
}
