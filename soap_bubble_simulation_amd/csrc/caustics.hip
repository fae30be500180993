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
// branch decision matches any FP32 evaluation of the same expressions; the
// power runs on the hardware log2/exp2 (caustic_pow).
#include <hip/hip_runtime.h>

#include <cmath>

#include "drainage.hpp"

namespace sf {

constexpr unsigned kCausticXcds = 8;   // MI355X: 8 accelerator complex dies

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

// I^s for the map's I >= 0 as exp2(s * log2 I) on the hardware log2 and exp2
// (v_log_f32, v_exp_f32: about one ulp each), the product split exactly with
// an FMA and its rounding error folded back to first order, so the result
// stays within about 1e-6 relative of the correctly rounded power; the shader
// evaluates pow() the same way.  s == 0 gives 1 (pow(x, 0) for every x).
__device__ inline float caustic_pow(float x, float s) {
    if (s == 0.0f) return 1.0f;
    const float l = __builtin_amdgcn_logf(x);
    const float p = s * l;
    const float pe = fmaf(s, l, -p);
    const float r = __builtin_amdgcn_exp2f(p);
    // l = -inf (x = 0) or r = inf leave pe or the correction non-finite
    return __builtin_isfinite(pe) && __builtin_isfinite(r) ? fmaf(r, pe * 0.6931471805599453f, r) : r;
}

// The intensity of one cell from its 3 x 3 neighbourhood (u: row above, d:
// row below; l/r: west/east), in the shader's operation order.
__device__ inline float caustic_cell(const CausticArgs& a, float c, float l, float r, float u, float d, float ul,
                                     float ur, float dl, float dr) {
#pragma clang fp contract(off)
    const float gx = (r - l) / 2.0f;
    const float gy = (d - u) / 2.0f;
    const float lap = (l + r + u + d - 4.0f * c) / 1.0f;
    const float d2x = (l + r - 2.0f * c) / 1.0f;
    const float d2y = (u + d - 2.0f * c) / 1.0f;
    const float dxy = (dr - dl - ur + ul) / 4.0f;
    const float det = d2x * d2y - dxy * dxy;
    const float curv = -lap * a.focal;
    // v_rcp_f32 and v_sqrt_f32 (about one ulp) in place of the correctly
    // rounded library forms: their error is below the power's, and the
    // branch decision does not depend on them.
    const float focusing = __builtin_amdgcn_rcpf(fmaxf(fabsf(1.0f - curv), 0.01f));
    const float branch = det < -a.threshold ? 1.5f : 1.0f;
    const float refraction = 1.0f + __builtin_amdgcn_sqrtf(gx * gx + gy * gy) * 0.5f;
    float I = focusing * branch * refraction;
    I = caustic_pow(I, a.sharpness) * a.intensity;
    return fminf(fmaxf(I, 0.0f), 10.0f);
}

// Any width: one cell per thread, nine gathered loads.
__global__ __launch_bounds__(256) void caustics_kernel(const CausticArgs a) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= a.np) return;
    a.out[static_cast<size_t>(y) * a.np + x] =
        caustic_cell(a, th_at(a, x, y), th_at(a, x - 1, y), th_at(a, x + 1, y), th_at(a, x, y - 1),
                     th_at(a, x, y + 1), th_at(a, x - 1, y - 1), th_at(a, x + 1, y - 1), th_at(a, x - 1, y + 1),
                     th_at(a, x + 1, y + 1));
}

// Thickness of four consecutive cells from column c (c % 4 == 0) of a row
// whose hi and lo start at rh and rl (rl null: thickness = hi); the row
// pointers are wave-uniform, so the loads take a scalar base and a 32-bit
// lane offset.
__device__ inline float4 th4_row(const float* rh, const float* rl, int c) {
    float4 t = *reinterpret_cast<const float4*>(rh + c);
    if (rl) {
        const float4 l = *reinterpret_cast<const float4*>(rl + c);
        t.x += l.x; t.y += l.y; t.z += l.z; t.w += l.w;
    }
    return t;
}

__device__ inline float th1_row(const float* rh, const float* rl, int c) {
    return rl ? rh[c] + rl[c] : rh[c];
}

// Widths divisible by 4: a lane owns four consecutive cells of one row and
// loads them as float4 from the rows above, at and below; the west and east
// neighbours come from the adjacent lanes, and only the wave's first lane
// and its last lane in the row load theirs (phi wrapped).  A block is four
// waves, 1024 columns of one row; blocks are renumbered so that each XCD
// walks a contiguous run of rows and the rows it shares with its neighbours
// stay in its own L2.
__global__ __launch_bounds__(256) void caustics_vec4_kernel(const CausticArgs a) {
    const int lane = threadIdx.x & 63;
    const unsigned nblk = gridDim.x * gridDim.y;
    const unsigned b = blockIdx.y * gridDim.x + blockIdx.x;
    const unsigned per_xcd = (nblk + kCausticXcds - 1) / kCausticXcds;
    unsigned v = (b % kCausticXcds) * per_xcd + b / kCausticXcds;
    if (nblk % kCausticXcds != 0) v = b;
    const int bx = static_cast<int>(v % gridDim.x), y = static_cast<int>(v / gridDim.x);
    const int wave0 = (bx * 4 + static_cast<int>(threadIdx.x >> 6)) * 256;
    if (wave0 >= a.np) return;                       // the whole wave is past the row's end
    const int c0 = wave0 + 4 * lane;
    const bool valid = c0 < a.np;
    const int cc = valid ? c0 : a.np - 4;
    const int cw = cc == 0 ? a.np - 1 : cc - 1;
    const int ce = cc + 4 == a.np ? 0 : cc + 4;
    const bool load_w = lane == 0;
    const bool load_e = lane == 63 || cc + 4 >= a.np;
    const int yu = max(y - 1, 0), yd = min(y + 1, a.nt - 1);
    const size_t np = static_cast<size_t>(a.np);
    const float *hu = a.hi + yu * np, *hm = a.hi + y * np, *hd = a.hi + yd * np;
    const float *lu = a.lo ? a.lo + yu * np : nullptr, *lm = a.lo ? a.lo + y * np : nullptr,
                *ld = a.lo ? a.lo + yd * np : nullptr;
    const float4 U = th4_row(hu, lu, cc), M = th4_row(hm, lm, cc), D = th4_row(hd, ld, cc);
    float uw = __shfl_up(U.w, 1), mw = __shfl_up(M.w, 1), dw = __shfl_up(D.w, 1);
    float ue = __shfl_down(U.x, 1), me = __shfl_down(M.x, 1), de = __shfl_down(D.x, 1);
    if (load_w) { uw = th1_row(hu, lu, cw); mw = th1_row(hm, lm, cw); dw = th1_row(hd, ld, cw); }
    if (load_e) { ue = th1_row(hu, lu, ce); me = th1_row(hm, lm, ce); de = th1_row(hd, ld, ce); }
    if (!valid) return;
    float4 o;
    o.x = caustic_cell(a, M.x, mw, M.y, U.x, D.x, uw, U.y, dw, D.y);
    o.y = caustic_cell(a, M.y, M.x, M.z, U.y, D.y, U.x, U.z, D.x, D.z);
    o.z = caustic_cell(a, M.z, M.y, M.w, U.z, D.z, U.y, U.w, D.y, D.w);
    o.w = caustic_cell(a, M.w, M.z, me, U.w, D.w, U.z, ue, D.z, de);
    *reinterpret_cast<float4*>(a.out + static_cast<size_t>(y) * a.np + c0) = o;
}

static void launch_caustics(const CausticArgs& a, hipStream_t s) {
    if (a.np % 4 == 0)
        hipLaunchKernelGGL(caustics_vec4_kernel, dim3((a.np + 1023) / 1024, a.nt), dim3(256), 0, s, a);
    else
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
