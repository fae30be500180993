// drainage.hip -- host side of the drainage path: the sf_drainage context,
// the marginal-regeneration (TFE) front schedule, and the C ABI entry points
// that replace DrainageSimulator (src/physics/drainage.rs) and
// GPUDrainageSimulator (src/render/gpu_drainage.rs).
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

#include <hip/hip_ext.h>

#include "drainage.hpp"

namespace sf {

// ---------------------------------------------------------------------------
// Small kernels around the step.

// Writes the pole rows from the ring partials of the last step (the ring
// mean of drainage.rs:480-503), so that the buffers hold the full state.
__global__ void materialize_poles_kernel(float* hi, float* lo, float* g, const float2* hpart,
                                         const float2* gpart, int nt, int np, int ntiles,
                                         int do_h, int do_g) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= np) return;
    const float np_f = static_cast<float>(np);
    if (do_h) {
        const float2 t = pole_mean(hpart, ntiles, np_f);
        const float2 b = pole_mean(hpart + ntiles, ntiles, np_f);
        hi[j] = t.x;
        lo[j] = t.y;
        hi[static_cast<size_t>(nt - 1) * np + j] = b.x;
        lo[static_cast<size_t>(nt - 1) * np + j] = b.y;
    }
    if (do_g) {
        const float2 t = pole_mean(gpart, ntiles, np_f);
        const float2 b = pole_mean(gpart + ntiles, ntiles, np_f);
        g[j] = t.x + t.y;
        g[static_cast<size_t>(nt - 1) * np + j] = b.x + b.y;
    }
}

struct StatsPart {
    double mn, mx, mass, lam;
    unsigned long long nrupt;
    unsigned long long pad;
};

// Per-block min / max / FP64 mass / FP64 CFL lambda / rupture count of
// h = hi + lo (lo may be null: an FP32 field).  Every block writes its own
// slot of `part`, so the partials need no clearing between calls.
__global__ __launch_bounds__(256) void stats_partial_kernel(const float* hi, const float* lo,
                                                            const RowD* rowd, int nt, int np,
                                                            float hc0, float hc1, float hc2,
                                                            double coef3, StatsPart* part) {
    const size_t n = static_cast<size_t>(nt) * np;
    double mn = INFINITY, mx = -INFINITY, mass = 0.0, lam = 0.0;
    unsigned long long nr = 0;
    for (size_t c = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; c < n;
         c += static_cast<size_t>(gridDim.x) * blockDim.x) {
        const int i = static_cast<int>(c / np);
        const float h = hi[c], l = lo ? lo[c] : 0.f;
        const double v = static_cast<double>(h) + static_cast<double>(l);
        mn = fmin(mn, v);
        mx = fmax(mx, v);
        const RowD rd = rowd[i];
        mass += v * rd.area;
        if (i > 0 && i < nt - 1) lam = fmax(lam, rd.lam_grid + coef3 * rd.sin_abs * v * v);
        nr += ff_less(h, l, hc0, hc1, hc2) ? 1ull : 0ull;
    }
    __shared__ double s[4][4];
    __shared__ unsigned long long sn[4];
    for (int off = 32; off > 0; off >>= 1) {
        mn = fmin(mn, __shfl_down(mn, off));
        mx = fmax(mx, __shfl_down(mx, off));
        mass += __shfl_down(mass, off);
        lam = fmax(lam, __shfl_down(lam, off));
        nr += __shfl_down(nr, off);
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
        s[0][wave] = mn; s[1][wave] = mx; s[2][wave] = mass; s[3][wave] = lam; sn[wave] = nr;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        StatsPart p{s[0][0], s[1][0], s[2][0], s[3][0], sn[0], 0};
        for (int w = 1; w < static_cast<int>(blockDim.x >> 6); ++w) {
            p.mn = fmin(p.mn, s[0][w]);
            p.mx = fmax(p.mx, s[1][w]);
            p.mass += s[2][w];
            p.lam = fmax(p.lam, s[3][w]);
            p.nrupt += sn[w];
        }
        part[blockIdx.x] = p;
    }
}

// One wave folds the per-block partials.
__global__ __launch_bounds__(64) void stats_final_kernel(const StatsPart* part, int nparts, StatsPart* out) {
    StatsPart r{INFINITY, -INFINITY, 0.0, 0.0, 0, 0};
    for (int b = threadIdx.x; b < nparts; b += 64) {
        r.mn = fmin(r.mn, part[b].mn);
        r.mx = fmax(r.mx, part[b].mx);
        r.mass += part[b].mass;
        r.lam = fmax(r.lam, part[b].lam);
        r.nrupt += part[b].nrupt;
    }
    for (int off = 32; off > 0; off >>= 1) {
        r.mn = fmin(r.mn, __shfl_down(r.mn, off));
        r.mx = fmax(r.mx, __shfl_down(r.mx, off));
        r.mass += __shfl_down(r.mass, off);
        r.lam = fmax(r.lam, __shfl_down(r.lam, off));
        r.nrupt += __shfl_down(r.nrupt, off);
    }
    if (threadIdx.x == 0) *out = r;
}

__global__ void gather_kernel(const float* hi, const float* lo, const int64_t* idx, int n,
                              float2* out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) out[k] = make_float2(hi[idx[k]], lo[idx[k]]);
}

__global__ void rupture_mask_kernel(const float* hi, const float* lo, size_t n, float hc0,
                                    float hc1, float hc2, uint8_t* mask) {
    const size_t c = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
    if (c < n) mask[c] = ff_less(hi[c], lo[c], hc0, hc1, hc2) ? 1 : 0;
}

__global__ void tfe_mask_kernel(StepArgs a, uint8_t* mask) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y;
    if (j >= a.np) return;
    mask[static_cast<size_t>(i) * a.np + j] = in_fronts(a, i, j) ? 1 : 0;
}

// Reduces the last step's per-block FP64 partials into its history slot.
__global__ __launch_bounds__(256) void diag_finalize_kernel(const double2* part, int nblocks,
                                                            const float2* hpole, int ntiles,
                                                            int pole_sum, double pole_area_top,
                                                            double pole_area_bot,
                                                            double pole_mass_fixed, double2* hist,
                                                            int slot) {
    __shared__ double red_d[2][kBlockMax / 64];
    diag_reduce(part, nblocks, hpole, ntiles, pole_sum, pole_area_top, pole_area_bot,
                pole_mass_fixed, hist, slot, red_d);
}

__global__ void fill_kernel(float* p, size_t n, float v) {
    const size_t c = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
    if (c < n) p[c] = v;
}

// ---------------------------------------------------------------------------
// Host helpers.

constexpr double kPi = 3.14159265358979323846;
constexpr double kFracPi2 = 1.57079632679489661923;

static inline unsigned cdiv(size_t a, size_t b) { return static_cast<unsigned>((a + b - 1) / b); }

template <class T>
static T* dalloc(size_t count) {
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, count * sizeof(T) + 16);
    if (e != hipSuccess) {
        set_error("hipMalloc of %zu bytes failed: %s", count * sizeof(T), hipGetErrorString(e));
        throw Error{SF_ERR_NOMEM};
    }
    return static_cast<T*>(p);
}

static void dfree(void* p) {
    if (p) (void)hipFree(p);
}

// Step launchers; t0/t1 (may be null) are events the runtime records at the
// kernel's own start and end (hipExtLaunchKernelGGL), for kernel timing.
using StepFn = void (*)(dim3, dim3, hipStream_t, hipEvent_t, hipEvent_t, const StepArgs&);

template <bool RB, int GM, bool M, bool D>
static void launch_step(dim3 grid, dim3 block, hipStream_t s, hipEvent_t t0, hipEvent_t t1, const StepArgs& a) {
    hipExtLaunchKernelGGL((drainage_step_kernel<RB, GM, M, D>), grid, block, 0, s, t0, t1, 0, a);
}

template <bool RB, int GM, bool M, bool D, int ROWS>
static void launch_vec4(dim3 grid, dim3 block, hipStream_t s, hipEvent_t t0, hipEvent_t t1, const StepArgs& a) {
    hipExtLaunchKernelGGL((drainage_vec4_kernel<RB, GM, M, D, ROWS, kPrefetchRows>), grid, block, 0, s, t0, t1, 0, a);
}

template <bool RB, int GM, bool M, bool D>
static void launch_band(dim3 grid, dim3 block, hipStream_t s, hipEvent_t t0, hipEvent_t t1, const StepArgs& a) {
    hipExtLaunchKernelGGL((drainage_band_kernel<RB, GM, M, D>), grid, block, 0, s, t0, t1, 0, a);
}

// Kernel table indexed by [rebase][gmode][marangoni][diag].  The scalar
// kernel serves any n_phi; the float4 kernels need n_phi % 4 == 0: unrolled
// row bands of 4 and 8, or the band-streaming kernel for 16..64 rows.
static StepFn step_fn(bool rebase, int gmode, bool marangoni, bool diag, bool vec4, int rows) {
#define SF_E(RB, GM, M, D) &launch_step<RB, GM, M, D>
#define SF_B(RB, GM, M, D) &launch_band<RB, GM, M, D>
#define SF_V(RB, GM, M, D, R) &launch_vec4<RB, GM, M, D, R>
#define SF_ROW(E, RB, GM)                                                   \
    {{E(RB, GM, false, false), E(RB, GM, false, true)},                     \
     {E(RB, GM, true, false), E(RB, GM, true, true)}}
#define SF_TAB(E)                                                           \
    {{SF_ROW(E, false, 0), SF_ROW(E, false, 1), SF_ROW(E, false, 2)},       \
     {SF_ROW(E, true, 0), SF_ROW(E, true, 1), SF_ROW(E, true, 2)}}
#define SF_V4(RB, GM, M, D) SF_V(RB, GM, M, D, 4)
#define SF_V8(RB, GM, M, D) SF_V(RB, GM, M, D, 8)
    static const StepFn scalar[2][3][2][2] = SF_TAB(SF_E);
    static const StepFn v4[2][3][2][2] = SF_TAB(SF_V4);
    static const StepFn v8[2][3][2][2] = SF_TAB(SF_V8);
    static const StepFn band[2][3][2][2] = SF_TAB(SF_B);
#undef SF_E
#undef SF_B
#undef SF_V
#undef SF_ROW
#undef SF_TAB
#undef SF_V4
#undef SF_V8
    if (!vec4) return scalar[rebase][gmode][marangoni][diag];
    if (rows <= 4) return v4[rebase][gmode][marangoni][diag];
    if (rows <= 8) return v8[rebase][gmode][marangoni][diag];
    return band[rebase][gmode][marangoni][diag];
}

// Two-step launches: [rebase][gmode 0 or 2][OPT].
template <bool RB, int GM, int OPT>
static void launch_band2(dim3 grid, dim3 block, hipStream_t s, hipEvent_t t0, hipEvent_t t1, const StepArgs& a) {
    hipExtLaunchKernelGGL((drainage_band2_kernel<RB, GM, OPT>), grid, block, 0, s, t0, t1, 0, a);
}

static StepFn step2_fn(bool rebase, int gmode, int opt) {
#define SF_O(RB, GM) {&launch_band2<RB, GM, 0>, &launch_band2<RB, GM, 1>, &launch_band2<RB, GM, 2>, \
                      &launch_band2<RB, GM, 3>, &launch_band2<RB, GM, 4>, &launch_band2<RB, GM, 5>, \
                      &launch_band2<RB, GM, 6>, &launch_band2<RB, GM, 7>}
    static const StepFn tab[2][2][8] = {{SF_O(false, 0), SF_O(false, 2)}, {SF_O(true, 0), SF_O(true, 2)}};
#undef SF_O
    return tab[rebase][gmode == 2 ? 1 : 0][opt & 7];
}

}  // namespace sf

using namespace sf;

// ---------------------------------------------------------------------------
// Context.

Drainage::Drainage(const sf_params& prm, int n_theta, int n_phi) : p(prm), nt(n_theta), np(n_phi) {
    n = static_cast<size_t>(nt) * np;
    dtheta = kPi / static_cast<double>(nt - 1);
    dphi = 2.0 * kPi / static_cast<double>(np);
    SF_HIP(hipGetDevice(&device));
    SF_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    // Kernel choice.  Grids of up to kSmallGrid cells are latency-bound: the
    // scalar kernel with one row per block gives the most waves (6.5 us per
    // step at 128 x 256 against 30 us for 8-row float4 blocks).  Larger grids
    // take the float4 kernels: 16-row bands (the band-streaming kernel) when
    // that still gives two blocks per CU, else 4-row blocks.  SOAPFILM_STENCIL
    // (scalar | vec4) and SOAPFILM_ROWS_PER_BLOCK override the choice (tests).
    constexpr size_t kSmallGrid = 262144;
    const char* force = std::getenv("SOAPFILM_STENCIL");
    const bool small = n <= kSmallGrid;
    if (force && std::strcmp(force, "scalar") == 0)
        vec4 = false;
    else if (force && std::strcmp(force, "vec4") == 0)
        vec4 = (np % 4 == 0);
    else
        vec4 = (np % 4 == 0) && !small;
    const int cols_per_wave = vec4 ? 256 : 64;
    wx = static_cast<int>(std::min<size_t>(4, cdiv(np, cols_per_wave)));
    block = 64 * wx;
    ntiles = static_cast<int>(cdiv(np, static_cast<size_t>(cols_per_wave) * wx));
    rows_per_block = vec4 ? (static_cast<size_t>(ntiles) * cdiv(nt - 2, 16) >= 512 ? 16 : 4) : (small ? 1 : 16);
    if (const char* e = std::getenv("SOAPFILM_ROWS_PER_BLOCK")) rows_per_block = std::max(1, std::atoi(e));
    if (vec4) rows_per_block = rows_per_block <= 4 ? 4 : (rows_per_block <= 8 ? 8 : std::min(rows_per_block, kBandMax));
    // Two-step bands: the tallest (fewest recomputed halo rows) that still
    // gives 512 blocks, two per CU (measured at 1024 x 2048 .. 8192 x 16384:
    // 4, 16, 64, 64 rows).  Two steps per launch are the default where such
    // a band has at least four rows; on smaller grids the recomputed halo
    // rows outweigh the saved traffic.
    rows_per_block2 = 1;
    for (int rb = kBandMax; rb >= 1; rb /= 2) {
        if (static_cast<size_t>(ntiles) * cdiv(nt - 2, rb) >= 512) {
            rows_per_block2 = rb;
            break;
        }
    }
    steps_per_launch = rows_per_block2 >= 4 ? 2 : 1;
    if (const char* e = std::getenv("SOAPFILM_ROWS_PER_BLOCK2"))
        rows_per_block2 = std::min(std::max(1, std::atoi(e)), kBandMax);
    for (int k = 0; k < 2; ++k) {
        hi[k] = dalloc<float>(n);
        lo[k] = dalloc<float>(n);
        g[k] = dalloc<float>(n);
        hpole[k] = dalloc<float2>(2 * static_cast<size_t>(ntiles));
        gpole[k] = dalloc<float2>(2 * static_cast<size_t>(ntiles));
    }
    rows = dalloc<RowF>(nt);
    rowd = dalloc<RowD>(nt);
    nblocks_max = static_cast<size_t>(ntiles) * cdiv(nt, 1);
    for (int k = 0; k < kDiagBufs; ++k) diag_part[k] = dalloc<double2>(nblocks_max);
    hpole_mid[0] = dalloc<float2>(2 * static_cast<size_t>(ntiles));
    hpole_mid[1] = dalloc<float2>(2 * static_cast<size_t>(ntiles));
    diag_hist = dalloc<double2>(kHistCap);
    stats_part = dalloc<StatsPart>(kStatsBlocks);
    stats_out = dalloc<StatsPart>(1);
    tables_dirty = true;
}

Drainage::~Drainage() {
    if (stream) (void)hipStreamSynchronize(stream);
    for (int k = 0; k < 2; ++k) {
        dfree(hi[k]); dfree(lo[k]); dfree(g[k]); dfree(hpole[k]); dfree(gpole[k]);
    }
    dfree(rows); dfree(rowd); dfree(diag_hist);
    for (int k = 0; k < kDiagBufs; ++k) dfree(diag_part[k]);
    dfree(lo_mid); dfree(hpole_mid[0]); dfree(hpole_mid[1]);
    dfree(stats_part); dfree(stats_out); dfree(gather_idx); dfree(gather_out);
    dfree(rgb); dfree(cmf_dev); dfree(spec_rows); dfree(spec_cols); dfree(spec_view); dfree(caustic);
    step_timer.destroy();
    spec_timer.destroy();
    if (stream) (void)hipStreamDestroy(stream);
}

// Per-row coefficient tables, FP64 on the host then rounded (drainage.rs:315-324).
void Drainage::upload_tables() {
#pragma clang fp contract(off)
    std::vector<RowF> rf(nt);
    std::vector<RowD> rd(nt);
    double D = p.diffusion_coeff;
    if ((p.flags & SF_FLAG_SURFACTANT) && p.surfactant_diffusion > D) D = p.surfactant_diffusion;
    const double R2 = p.bubble_radius * p.bubble_radius;
    for (int i = 0; i < nt; ++i) {
        const double theta = static_cast<double>(i) * dtheta;
        const double s = std::sin(theta), c = std::cos(theta);
        const double ss = std::fabs(s) < p.sin_guard ? std::copysign(p.sin_guard, s) : s;
        rf[i].sin_t = static_cast<float>(s);
        rf[i].cot = static_cast<float>(c / ss);
        rf[i].inv_s2 = static_cast<float>(1.0 / (ss * ss));
        rf[i].inv_s = static_cast<float>(1.0 / ss);
        const double lo_b = i == 0 ? 0.0 : theta - 0.5 * dtheta;
        const double hi_b = i == nt - 1 ? kPi : theta + 0.5 * dtheta;
        rd[i].area = R2 * (std::cos(lo_b) - std::cos(hi_b)) * dphi;
        rd[i].lam_grid = 4.0 * D * (1.0 / (dtheta * dtheta) + 1.0 / (ss * ss * dphi * dphi)) / R2;
        rd[i].sin_abs = std::fabs(s);
        rd[i].pad = 0.0;
    }
    SF_HIP(hipMemcpyAsync(rows, rf.data(), nt * sizeof(RowF), hipMemcpyHostToDevice, stream));
    SF_HIP(hipMemcpyAsync(rowd, rd.data(), nt * sizeof(RowD), hipMemcpyHostToDevice, stream));
    SF_HIP(hipStreamSynchronize(stream));
    pole_area_top = rd[0].area;
    pole_area_bot = rd[nt - 1].area;
    tables_dirty = false;
}

// FP64 pole-row film volume, used while the h poles are copied through.
void Drainage::update_fixed_pole_mass() {
    std::vector<float> h0(2 * static_cast<size_t>(np)), l0(2 * static_cast<size_t>(np));
    SF_HIP(hipMemcpyAsync(h0.data(), hi[cur_hi], np * sizeof(float), hipMemcpyDeviceToHost, stream));
    SF_HIP(hipMemcpyAsync(l0.data(), lo[cur], np * sizeof(float), hipMemcpyDeviceToHost, stream));
    SF_HIP(hipMemcpyAsync(h0.data() + np, hi[cur_hi] + (nt - 1) * static_cast<size_t>(np),
                          np * sizeof(float), hipMemcpyDeviceToHost, stream));
    SF_HIP(hipMemcpyAsync(l0.data() + np, lo[cur] + (nt - 1) * static_cast<size_t>(np),
                          np * sizeof(float), hipMemcpyDeviceToHost, stream));
    SF_HIP(hipStreamSynchronize(stream));
    double st = 0.0, sb = 0.0;
    for (int j = 0; j < np; ++j) {
        st += static_cast<double>(h0[j]) + static_cast<double>(l0[j]);
        sb += static_cast<double>(h0[np + j]) + static_cast<double>(l0[np + j]);
    }
    pole_mass_fixed = pole_area_top * st + pole_area_bot * sb;
}

// Uploads hi/lo/Γ into both ping-pong halves so that rows a step never
// writes (the poles, and Γ when it is frozen) hold the same values in both.
void Drainage::upload(const float* h_hi, const float* h_lo, const float* gam) {
    if (h_hi) {
        for (int k = 0; k < 2; ++k) {
            SF_HIP(hipMemcpyAsync(hi[k], h_hi, n * sizeof(float), hipMemcpyHostToDevice, stream));
            SF_HIP(hipMemcpyAsync(lo[k], h_lo, n * sizeof(float), hipMemcpyHostToDevice, stream));
        }
    }
    if (gam) {
        for (int k = 0; k < 2; ++k)
            SF_HIP(hipMemcpyAsync(g[k], gam, n * sizeof(float), hipMemcpyHostToDevice, stream));
    }
    SF_HIP(hipStreamSynchronize(stream));
    cur = cur_hi = cur_g = 0;
    since_rebase = 0;
    poles_virtual = false;
    update_fixed_pole_mass();
}

void Drainage::fill(double h0, double gamma0) {
    const float h_hi = static_cast<float>(h0);
    const float h_lo = static_cast<float>(h0 - static_cast<double>(h_hi));
    for (int k = 0; k < 2; ++k) {
        hipLaunchKernelGGL(fill_kernel, dim3(cdiv(n, 256)), dim3(256), 0, stream, hi[k], n, h_hi);
        hipLaunchKernelGGL(fill_kernel, dim3(cdiv(n, 256)), dim3(256), 0, stream, lo[k], n, h_lo);
        hipLaunchKernelGGL(fill_kernel, dim3(cdiv(n, 256)), dim3(256), 0, stream, g[k], n,
                           static_cast<float>(gamma0));
    }
    SF_HIP(hipGetLastError());
    SF_HIP(hipStreamSynchronize(stream));
    cur = cur_hi = cur_g = 0;
    since_rebase = 0;
    poles_virtual = false;
    update_fixed_pole_mass();
}

// ---------------------------------------------------------------------------
// Marginal-regeneration schedule, FP64 on the host with the reference's
// accumulation order so that spawn times and front extents are bit-equal.

// Cyclic run of columns j with the phi part of TfeFront::contains true
// (drainage.rs:64-70).  The predicate is monotone in the cyclic distance to
// phi_center, so the set is one run.
void Drainage::front_columns(HostFront& f) const {
#pragma clang fp contract(off)
    std::vector<uint8_t> in(np);
    int count = 0;
    for (int j = 0; j < np; ++j) {
        const double phi = static_cast<double>(j) * dphi;
        double phi_diff = std::fabs(phi - f.phi_center);
        if (phi_diff > kPi) phi_diff = 2.0 * kPi - phi_diff;
        in[j] = phi_diff < f.width / 2.0;
        count += in[j];
    }
    f.j0 = 0;
    f.len = count;
    if (count == 0 || count == np) return;
    int start = -1;
    for (int j = 0; j < np; ++j)
        if (in[j] && !in[(j + np - 1) % np]) { start = j; break; }
    for (int k = 0; k < count; ++k)
        SF_REQUIRE(in[(start + k) % np], SF_ERR_INTERNAL, "TFE front columns are not one run");
    f.j0 = start;
}

void Drainage::spawn_front() {
#pragma clang fp contract(off)
    const double t = current_time;
    HostFront f;
    f.phi_center = std::fabs(std::sin(t * 7.89 + 1.23)) * 2.0 * kPi;      // drainage.rs:466
    f.width = 0.3 + std::fabs(std::sin(t * 3.21)) * 0.4;                // 467
    f.velocity = 0.05 + std::fabs(std::cos(t * 5.67)) * 0.05;           // 468
    f.extent = 0.0;
    front_columns(f);
    fronts.push_back(f);
    if (fronts.size() > static_cast<size_t>(kMaxFronts)) fronts.erase(fronts.begin());   // 474-476
}

// One step of update_marginal_regeneration's bookkeeping (drainage.rs:395-438);
// fills the kernel's front boxes for this step.
void Drainage::schedule_fronts(double dt, StepArgs& a) {
#pragma clang fp contract(off)
    a.nfronts = 0;
    if (!tfe_enabled) return;
    next_spawn -= dt;
    if (next_spawn <= 0.0) {
        spawn_front();
        const double random_factor = 0.5 + std::fabs(std::sin(current_time * 12.345));
        next_spawn = p.tfe_spawn_interval * random_factor;
    }
    for (auto& f : fronts) {
        const double e = f.extent + f.velocity * dt;
        f.extent = e < kFracPi2 ? e : kFracPi2;
    }
    fronts.erase(std::remove_if(fronts.begin(), fronts.end(),
                                [](const HostFront& f) { return !(f.extent < kFracPi2 * 0.95); }),
                 fronts.end());
    if (fronts.empty()) return;
    long min_all = nt, max_all = 0;
    for (const auto& f : fronts) {
        const double a0 = std::floor((kFracPi2 - f.extent) / dtheta);
        const double b0 = std::ceil((kFracPi2 + f.extent) / dtheta);
        long mi = a0 > 0.0 ? static_cast<long>(a0) : 0;
        long ma = b0 > 0.0 ? static_cast<long>(b0) : 0;
        if (ma > nt - 1) ma = nt - 1;
        min_all = std::min(min_all, mi);
        max_all = std::max(max_all, ma);
    }
    if (min_all > nt - 1) min_all = nt - 1;
    const double ratio = p.tfe_thickness_ratio;
    const double reduction_rate = (1.0 - ratio) * 0.5;
    const double factor = 1.0 - reduction_rate * dt;          // drainage.rs:454
    const double gm1 = factor - 1.0;                           // exact (Sterbenz)
    a.tfe_g_hi = static_cast<float>(gm1);
    a.tfe_g_lo = static_cast<float>(gm1 - static_cast<double>(a.tfe_g_hi));
    for (const auto& f : fronts) {
        int ilo = -1, ihi = -2;
        for (long i = min_all; i <= max_all; ++i) {
            const double theta = static_cast<double>(i) * dtheta;
            if (std::fabs(theta - kFracPi2) > f.extent) continue;   // drainage.rs:57-62
            if (ilo < 0) ilo = static_cast<int>(i);
            ihi = static_cast<int>(i);
        }
        if (ilo < 0 || f.len == 0) continue;
        ilo = std::max(ilo, 1);                     // poles are never inside a front
        ihi = std::min(ihi, nt - 2);
        if (ilo > ihi) continue;
        a.fr[a.nfronts++] = FrontBox{ilo, ihi, f.j0, f.len};
    }
}

void Drainage::fill_common_args(StepArgs& a, float dt_f) const {
#pragma clang fp contract(off)
    const double R = p.bubble_radius;
    a.rows = rows;
    a.rowd = rowd;
    a.nt = nt;
    a.np = np;
    a.ntiles = ntiles;
    a.rows_per_block = rows_per_block;
    a.dt = dt_f;
    a.inv_dth2 = static_cast<float>(1.0 / (dtheta * dtheta));
    a.inv_2dth = static_cast<float>(1.0 / (2.0 * dtheta));
    a.inv_dph2 = static_cast<float>(1.0 / (dphi * dphi));
    a.inv_2dph = static_cast<float>(1.0 / (2.0 * dphi));
    a.d_r2 = static_cast<float>(p.diffusion_coeff / (R * R));
    a.ds_r2 = static_cast<float>(p.surfactant_diffusion / (R * R));
    const double coef = p.density * p.gravity / (3.0 * p.viscosity);
    a.coef = static_cast<float>(coef);
    a.coef3 = 3.0 * coef;
    a.mgr = static_cast<float>(p.marangoni_coeff * p.gamma_reduction);
    a.inv_r = static_cast<float>(1.0 / R);
    const Float3 hc = split3(p.critical_thickness);
    a.hc0 = hc.c0;
    a.hc1 = hc.c1;
    a.hc2 = hc.c2;
    a.np_f = static_cast<float>(np);
    a.pole_area_top = pole_area_top;
    a.pole_area_bot = pole_area_bot;
    a.pole_mass_fixed = pole_mass_fixed;
    a.diag_hist = diag_hist;
}

void Drainage::materialize_poles() {
    if (!poles_virtual) return;
    const bool do_h = !(p.flags & SF_FLAG_POLES_COPY_H);
    const bool do_g = (p.flags & SF_FLAG_POLES_AVG_GAMMA) && (p.flags & SF_FLAG_SURFACTANT);
    hipLaunchKernelGGL(materialize_poles_kernel, dim3(cdiv(np, 256)), dim3(256), 0, stream,
                       hi[cur_hi], lo[cur], g[cur_g], hpole[cur_pole], gpole[cur_pole], nt, np,
                       ntiles, do_h ? 1 : 0, do_g ? 1 : 0);
    SF_HIP(hipGetLastError());
    poles_virtual = false;
}

// Reduces the per-block FP64 partials of the last diagnosed step, which no
// later launch has consumed yet.
void Drainage::finalize_diag() {
    for (int q = 0; q < npending; ++q) {
        const PendingDiag& e = pending[q];
        hipLaunchKernelGGL(diag_finalize_kernel, dim3(1), dim3(256), 0, stream, diag_part[e.buf], e.nblocks,
                           e.hpole, ntiles, e.pole_sum ? 1 : 0, pole_area_top, pole_area_bot, pole_mass_fixed,
                           diag_hist, e.slot);
        SF_HIP(hipGetLastError());
    }
    npending = 0;
}

// Partial buffers for a diagnosed launch of nsteps steps (ones no pending
// step holds) and the pending steps its block 0 reduces.
void Drainage::diag_args(StepArgs& a, int nsteps) {
    int k = 0;
    for (int b = 0; b < kDiagBufs && k < nsteps; ++b) {
        bool held = false;
        for (int q = 0; q < npending; ++q) held |= pending[q].buf == b;
        if (!held) new_buf[k++] = b;
    }
    a.diag_part = diag_part[new_buf[0]];
    a.diag_part2 = nsteps == 2 ? diag_part[new_buf[1]] : nullptr;
    a.nprev = npending;
    for (int q = 0; q < 2; ++q) {
        a.prev[q] = DiagPending{nullptr, nullptr, 0, 0, -1, 0};
        if (q < npending) {
            const PendingDiag& e = pending[q];
            a.prev[q] = DiagPending{diag_part[e.buf], e.hpole, e.nblocks, e.pole_sum ? 1 : 0, e.slot, 0};
        }
    }
}

// The launch's steps become the pending ones, in step order.
void Drainage::diag_commit(int nsteps, int nblocks, bool pole_sum, const float2* hpole_first,
                           const float2* hpole_last) {
    npending = nsteps;
    for (int q = 0; q < nsteps; ++q) {
        pending[q] = PendingDiag{new_buf[q], static_cast<int>(hist_count % kHistCap), nblocks, pole_sum,
                                 q == nsteps - 1 ? hpole_last : hpole_first};
        ++hist_count;
    }
}

void Drainage::materialize() {
    finalize_diag();
    materialize_poles();
}

double Drainage::resolve_dt(double dt) {
    if (dt > 0.0) return dt;
    const StatsPart s = compute_stats();
    const double cfl = s.lam > 0.0 ? 2.0 / s.lam : INFINITY;
    SF_REQUIRE(std::isfinite(cfl) && cfl > 0.0, SF_ERR_INVALID,
               "CFL time step is not finite (is diffusion zero?)");
    return p.cfl_safety * cfl;
}

bool Drainage::two_step_ok() const {
    const uint32_t f = p.flags;
    const bool g_avg = (f & SF_FLAG_SURFACTANT) && (f & SF_FLAG_POLES_AVG_GAMMA);
    // byte offsets of the two-step kernel are 32-bit: below 2^30 cells
    return vec4 && np % 256 == 0 && ntiles <= kMaxTiles2 && nt >= 4 && n < (size_t{1} << 30) &&
           !(f & SF_FLAG_MARANGONI) && !g_avg && p.rebase_interval != 1;
}

bool Drainage::two_step_next(int remaining) const {
    const int rebase_every = p.rebase_interval > 0 ? static_cast<int>(p.rebase_interval) : 1;
    return remaining >= 2 && steps_per_launch == 2 && two_step_ok() && since_rebase + 1 < rebase_every;
}

void Drainage::launch_one(double dt) {
    const bool marangoni = (p.flags & SF_FLAG_MARANGONI) != 0;
    const bool surfactant = (p.flags & SF_FLAG_SURFACTANT) != 0;
    const int gmode = surfactant ? 2 : (marangoni ? 1 : 0);
    const bool diag = (p.flags & SF_FLAG_DIAG) != 0;
    const bool h_avg = !(p.flags & SF_FLAG_POLES_COPY_H);
    const bool g_avg = surfactant && (p.flags & SF_FLAG_POLES_AVG_GAMMA);
    const int rebase_every = p.rebase_interval > 0 ? static_cast<int>(p.rebase_interval) : 1;
    const dim3 grid(ntiles, cdiv(nt - 2, rows_per_block));
    StepArgs a{};
    fill_common_args(a, static_cast<float>(dt));
    schedule_fronts(dt, a);
    last_boxes.assign(a.fr, a.fr + a.nfronts);
    const bool rebase = (since_rebase + 1 >= rebase_every);
    a.hi_in = hi[cur_hi];
    a.lo_in = lo[cur];
    a.g_in = g[cur_g];
    a.hi_out = hi[1 - cur_hi];
    a.lo_out = lo[1 - cur];
    a.g_out = g[1 - cur_g];
    a.h_pole_mean = (poles_virtual && h_avg) ? 1 : 0;
    a.g_pole_mean = (poles_virtual && g_avg) ? 1 : 0;
    a.h_pole_sum = h_avg ? 1 : 0;
    a.g_pole_sum = g_avg ? 1 : 0;
    a.hpole_in = hpole[cur_pole];
    a.gpole_in = gpole[cur_pole];
    a.hpole_out = hpole[1 - cur_pole];
    a.gpole_out = gpole[1 - cur_pole];
    if (diag)
        diag_args(a, 1);
    else
        finalize_diag();
    const StepFn fn = step_fn(rebase, gmode, marangoni, diag, vec4, rows_per_block);
    if (step_timer.enabled) {
        auto& ev = step_timer.next(stream);
        fn(grid, dim3(block), stream, ev.first, ev.second, a);
    } else {
        fn(grid, dim3(block), stream, nullptr, nullptr, a);
    }
    SF_HIP(hipGetLastError());
    cur = 1 - cur;
    if (surfactant) cur_g = 1 - cur_g;
    if (rebase) {
        cur_hi = 1 - cur_hi;
        since_rebase = 0;
    } else {
        ++since_rebase;
    }
    if (h_avg || g_avg) {
        cur_pole = 1 - cur_pole;
        poles_virtual = true;
    }
    if (diag) diag_commit(1, static_cast<int>(grid.x * grid.y), h_avg, a.hpole_out, a.hpole_out);
    current_time += dt;
    ++steps_taken;
}

// Two steps in one launch of drainage_band2_kernel (two_step_next holds).
void Drainage::launch_two(double dt, bool with_mid) {
    const bool surfactant = (p.flags & SF_FLAG_SURFACTANT) != 0;
    const int gmode = surfactant ? 2 : 0;
    const bool diag = (p.flags & SF_FLAG_DIAG) != 0;
    const bool h_avg = !(p.flags & SF_FLAG_POLES_COPY_H);
    const int rebase_every = p.rebase_interval > 0 ? static_cast<int>(p.rebase_interval) : 1;
    const dim3 grid2(ntiles, cdiv(nt - 2, rows_per_block2));
    StepArgs a{};
    fill_common_args(a, static_cast<float>(dt));
    a.rows_per_block = rows_per_block2;
    // Front schedules of both steps, each at its own start time, as two
    // one-step launches take them.
    schedule_fronts(dt, a);
    current_time += dt;
    StepArgs b = a;
    schedule_fronts(dt, b);
    current_time += dt;
    a.nfronts2 = b.nfronts;
    for (int k = 0; k < b.nfronts; ++k) a.fr2[k] = b.fr[k];
    if (b.nfronts > 0) {
        a.tfe_g_hi = b.tfe_g_hi;
        a.tfe_g_lo = b.tfe_g_lo;
    }
    last_boxes.assign(b.fr, b.fr + b.nfronts);
    const bool rebase = since_rebase + 2 >= rebase_every;
    a.hi_in = hi[cur_hi];
    a.lo_in = lo[cur];
    a.g_in = g[cur_g];
    a.hi_out = hi[1 - cur_hi];
    a.lo_out = lo[1 - cur];
    a.g_out = g[1 - cur_g];
    a.h_pole_mean = (poles_virtual && h_avg) ? 1 : 0;
    a.g_pole_mean = 0;
    a.h_pole_sum = h_avg ? 1 : 0;
    a.g_pole_sum = 0;
    a.hpole_in = hpole[cur_pole];
    a.hpole_out = hpole[1 - cur_pole];
    a.hpole_mid = hpole_mid[cur_mid];
    if (with_mid) {
        if (!lo_mid) lo_mid = dalloc<float>(n);
        a.lo_mid = lo_mid;
        mid_hi = hi[cur_hi];
        mid_hpole = hpole_mid[cur_mid];
    }
    if (diag)
        diag_args(a, 2);
    else
        finalize_diag();
    const int opt = (diag ? kOptDiag : 0) | (a.nfronts + a.nfronts2 > 0 ? kOptFronts : 0) | (with_mid ? kOptMid : 0);
    const StepFn fn = step2_fn(rebase, gmode, opt);
    if (step_timer.enabled) {
        auto& ev = step_timer.next(stream);
        fn(grid2, dim3(block), stream, ev.first, ev.second, a);
    } else {
        fn(grid2, dim3(block), stream, nullptr, nullptr, a);
    }
    SF_HIP(hipGetLastError());
    cur = 1 - cur;
    if (surfactant) cur_g = 1 - cur_g;
    if (rebase) {
        cur_hi = 1 - cur_hi;
        since_rebase = 0;
    } else {
        since_rebase += 2;
    }
    if (h_avg) {
        cur_pole = 1 - cur_pole;
        poles_virtual = true;
    }
    if (diag) diag_commit(2, static_cast<int>(grid2.x * grid2.y), h_avg, a.hpole_mid, a.hpole_out);
    cur_mid = 1 - cur_mid;
    steps_taken += 2;
}

void Drainage::step(int n_steps, double dt, bool flush) {
    if (tables_dirty) upload_tables();
    dt = resolve_dt(dt);
    for (int s = 0; s < n_steps;) {
        if (two_step_next(n_steps - s)) {
            launch_two(dt, false);
            s += 2;
        } else {
            launch_one(dt);
            s += 1;
        }
    }
    if (flush) materialize();
}

std::pair<hipEvent_t, hipEvent_t>& KernelTimer::next(hipStream_t s) {
    if (used == pool.size()) {
        if (pool.size() < kPool) {
            hipEvent_t a, b;
            SF_HIP(hipEventCreate(&a));
            SF_HIP(hipEventCreate(&b));
            pool.emplace_back(a, b);
        } else {
            drain(s);
        }
    }
    return pool[used++];
}

void KernelTimer::drain(hipStream_t s) {
    SF_HIP(hipStreamSynchronize(s));
    for (size_t k = 0; k < used; ++k) {
        float t = 0.f;
        SF_HIP(hipEventElapsedTime(&t, pool[k].first, pool[k].second));
        ms += t;
        ++launches;
    }
    used = 0;
}

void KernelTimer::destroy() {
    for (auto& e : pool) {
        (void)hipEventDestroy(e.first);
        (void)hipEventDestroy(e.second);
    }
    pool.clear();
    used = 0;
}

StatsPart Drainage::compute_stats() {
    if (tables_dirty) upload_tables();
    materialize();
    const size_t blocks = std::min<size_t>(kStatsBlocks, std::max<size_t>(1, cdiv(n, 256)));
    const Float3 hc = split3(p.critical_thickness);
    const double coef3 = 3.0 * (p.density * p.gravity / (3.0 * p.viscosity));
    hipLaunchKernelGGL(stats_partial_kernel, dim3(static_cast<unsigned>(blocks)), dim3(256), 0,
                       stream, hi[cur_hi], lo[cur], rowd, nt, np, hc.c0, hc.c1, hc.c2, coef3,
                       stats_part);
    hipLaunchKernelGGL(stats_final_kernel, dim3(1), dim3(64), 0, stream, stats_part,
                       static_cast<int>(blocks), stats_out);
    SF_HIP(hipGetLastError());
    StatsPart r;
    SF_HIP(hipMemcpyAsync(&r, stats_out, sizeof(StatsPart), hipMemcpyDeviceToHost, stream));
    SF_HIP(hipStreamSynchronize(stream));
    return r;
}

void Drainage::ensure_gather(size_t count) {
    if (count <= gather_cap) return;
    dfree(gather_idx);
    dfree(gather_out);
    gather_idx = dalloc<int64_t>(count);
    gather_out = dalloc<float2>(count);
    gather_cap = count;
}

// ---------------------------------------------------------------------------
// C ABI.

static Drainage* D(sf_drainage* c) { return reinterpret_cast<Drainage*>(c); }
static const Drainage* D(const sf_drainage* c) { return reinterpret_cast<const Drainage*>(c); }

static void check_params(const sf_params* p) {
    SF_REQUIRE(p != nullptr, SF_ERR_INVALID, "params is NULL");
    SF_REQUIRE(p->bubble_radius > 0.0 && std::isfinite(p->bubble_radius), SF_ERR_INVALID,
               "bubble_radius must be positive");
    SF_REQUIRE(p->viscosity > 0.0, SF_ERR_INVALID, "viscosity must be positive");
    SF_REQUIRE(p->sin_guard > 0.0, SF_ERR_INVALID, "sin_guard must be positive");
    SF_REQUIRE(std::isfinite(p->critical_thickness), SF_ERR_INVALID,
               "critical_thickness must be finite");
}

extern "C" int sf_params_cpu_default(sf_params* o) {
    return guard([&] {
        SF_REQUIRE(o != nullptr, SF_ERR_INVALID, "out is NULL");
        std::memset(o, 0, sizeof(*o));
        o->dt = 0.001;                          // config.rs:312
        o->gravity = 9.81;                      // config.rs:94
        o->viscosity = 0.001;                   // config.rs:61
        o->density = 1000.0;                    // config.rs:63
        o->diffusion_coeff = 1e-9;              // drainage.rs:277 <- config.rs:64
        o->bubble_radius = 0.05 / 2.0;          // config.rs:30, 358
        o->critical_thickness = 10.0 * 1e-9;    // drainage.rs:278 <- config.rs:31
        o->gamma_air = 0.072;                   // gpu_drainage.rs:60
        o->gamma_reduction = 0.045;             // gpu_drainage.rs:61
        o->surfactant_diffusion = 1e-9;         // gpu_drainage.rs:62
        o->marangoni_coeff = 0.01;              // gpu_drainage.rs:63
        o->sin_guard = 1e-10;                   // drainage.rs:320
        o->tfe_thickness_ratio = 0.85;          // drainage.rs:283
        o->tfe_first_spawn = 0.5;               // drainage.rs:285
        o->tfe_spawn_interval = 2.0;            // drainage.rs:286
        o->cfl_safety = 1.0;
        o->flags = SF_FLAG_TFE;                 // drainage.rs:284
        o->rebase_interval = 8;
    });
}

extern "C" int sf_params_gpu_default(sf_params* o) {
    return guard([&] {
        SF_REQUIRE(o != nullptr, SF_ERR_INVALID, "out is NULL");
        std::memset(o, 0, sizeof(*o));
        o->dt = 0.001;                          // gpu_drainage.rs:49
        o->gravity = 9.81;
        o->viscosity = 0.001;
        o->density = 1000.0;
        o->diffusion_coeff = 1e-9;              // gpu_drainage.rs:53
        o->bubble_radius = 0.025;               // gpu_drainage.rs:54
        o->critical_thickness = 30e-9;          // gpu_drainage.rs:55
        o->gamma_air = 0.072;
        o->gamma_reduction = 0.045;
        o->surfactant_diffusion = 1e-9;
        o->marangoni_coeff = 0.01;
        o->sin_guard = 1e-4;                    // drainage.wgsl:101
        o->tfe_thickness_ratio = 0.85;
        o->tfe_first_spawn = 0.5;
        o->tfe_spawn_interval = 2.0;
        o->cfl_safety = 1.0;
        o->flags = SF_FLAG_SURFACTANT | SF_FLAG_POLES_COPY_H;   // drainage.wgsl:85-89, 178-201
        o->rebase_interval = 8;
    });
}

extern "C" int sf_drainage_create(const sf_params* params, int32_t n_theta, int32_t n_phi,
                                  double h0, double gamma0, sf_drainage** out) {
    return guard([&] {
        SF_REQUIRE(out != nullptr, SF_ERR_INVALID, "out is NULL");
        *out = nullptr;
        check_params(params);
        SF_REQUIRE(n_theta >= 3 && n_phi >= 1, SF_ERR_INVALID,
                   "grid %dx%d: need n_theta >= 3 and n_phi >= 1", n_theta, n_phi);
        SF_REQUIRE(static_cast<int64_t>(n_theta) * n_phi < (int64_t(1) << 40), SF_ERR_INVALID,
                   "grid too large");
        require_device();
        Drainage* d = new Drainage(*params, n_theta, n_phi);
        try {
            d->current_time = 0.0;
            d->tfe_enabled = (params->flags & SF_FLAG_TFE) != 0;
            d->next_spawn = params->tfe_first_spawn;
            d->upload_tables();
            d->fill(h0, gamma0);
        } catch (...) {
            delete d;
            throw;
        }
        *out = reinterpret_cast<sf_drainage*>(d);
    });
}

extern "C" void sf_drainage_destroy(sf_drainage* ctx) { delete D(ctx); }

extern "C" int sf_drainage_set_params(sf_drainage* ctx, const sf_params* params) {
    return guard([&] {
        SF_REQUIRE(ctx != nullptr, SF_ERR_INVALID, "ctx is NULL");
        check_params(params);
        Drainage* d = D(ctx);
        d->materialize();
        const bool was_copy_h = d->p.flags & SF_FLAG_POLES_COPY_H;
        d->p = *params;
        d->tfe_enabled = (params->flags & SF_FLAG_TFE) != 0;
        if (!d->tfe_enabled) d->fronts.clear();
        d->upload_tables();
        if (was_copy_h != static_cast<bool>(params->flags & SF_FLAG_POLES_COPY_H))
            d->update_fixed_pole_mass();
    });
}

extern "C" int sf_drainage_get_params(const sf_drainage* ctx, sf_params* out) {
    return guard([&] {
        SF_REQUIRE(ctx != nullptr && out != nullptr, SF_ERR_INVALID, "NULL argument");
        *out = D(ctx)->p;
    });
}

extern "C" int sf_drainage_set_fields(sf_drainage* ctx, const float* h, const float* gamma) {
    return guard([&] {
        SF_REQUIRE(ctx != nullptr, SF_ERR_INVALID, "ctx is NULL");
        Drainage* d = D(ctx);
        d->materialize();
        std::vector<float> zeros;
        if (h) zeros.assign(d->n, 0.f);
        if (!h) {
            // keep h: copy the current one into both halves
            std::vector<float> hh(d->n), ll(d->n);
            SF_HIP(hipMemcpyAsync(hh.data(), d->hi[d->cur_hi], d->n * 4, hipMemcpyDeviceToHost, d->stream));
            SF_HIP(hipMemcpyAsync(ll.data(), d->lo[d->cur], d->n * 4, hipMemcpyDeviceToHost, d->stream));
            SF_HIP(hipStreamSynchronize(d->stream));
            if (!gamma) {
                std::vector<float> gg(d->n);
                SF_HIP(hipMemcpy(gg.data(), d->g[d->cur_g], d->n * 4, hipMemcpyDeviceToHost));
                d->upload(hh.data(), ll.data(), gg.data());
            } else {
                d->upload(hh.data(), ll.data(), gamma);
            }
            return;
        }
        if (!gamma) {
            std::vector<float> gg(d->n);
            SF_HIP(hipMemcpy(gg.data(), d->g[d->cur_g], d->n * 4, hipMemcpyDeviceToHost));
            d->upload(h, zeros.data(), gg.data());
        } else {
            d->upload(h, zeros.data(), gamma);
        }
    });
}

extern "C" int sf_drainage_set_fields_f64(sf_drainage* ctx, const double* h, const double* gamma) {
    return guard([&] {
        SF_REQUIRE(ctx != nullptr, SF_ERR_INVALID, "ctx is NULL");
        Drainage* d = D(ctx);
        d->materialize();
        std::vector<float> hh(d->n), ll(d->n), gg(d->n);
        if (h) {
            for (size_t k = 0; k < d->n; ++k) {
                hh[k] = static_cast<float>(h[k]);
                ll[k] = static_cast<float>(h[k] - static_cast<double>(hh[k]));
            }
        } else {
            SF_HIP(hipMemcpy(hh.data(), d->hi[d->cur_hi], d->n * 4, hipMemcpyDeviceToHost));
            SF_HIP(hipMemcpy(ll.data(), d->lo[d->cur], d->n * 4, hipMemcpyDeviceToHost));
        }
        if (gamma) {
            for (size_t k = 0; k < d->n; ++k) gg[k] = static_cast<float>(gamma[k]);
        } else {
            SF_HIP(hipMemcpy(gg.data(), d->g[d->cur_g], d->n * 4, hipMemcpyDeviceToHost));
        }
        d->upload(hh.data(), ll.data(), gg.data());
    });
}

extern "C" int sf_drainage_step(sf_drainage* ctx, int32_t n_steps, double dt) {
    return guard([&] {
        SF_REQUIRE(ctx != nullptr, SF_ERR_INVALID, "ctx is NULL");
        SF_REQUIRE(n_steps >= 0, SF_ERR_INVALID, "n_steps must be >= 0");
        SF_REQUIRE(!std::isnan(dt) && !std::isinf(dt), SF_ERR_INVALID, "dt must be finite");
        if (n_steps == 0) return;
        D(ctx)->step(n_steps, dt);
    });
}

extern "C" int sf_drainage_copy_field(sf_drainage* ctx, float* h, float* gamma) {
    return guard([&] {
        SF_REQUIRE(ctx != nullptr, SF_ERR_INVALID, "ctx is NULL");
        Drainage* d = D(ctx);
        d->materialize();
        if (h) {
            std::vector<float> ll(d->n);
            SF_HIP(hipMemcpyAsync(h, d->hi[d->cur_hi], d->n * 4, hipMemcpyDeviceToHost, d->stream));
            SF_HIP(hipMemcpyAsync(ll.data(), d->lo[d->cur], d->n * 4, hipMemcpyDeviceToHost, d->stream));
            SF_HIP(hipStreamSynchronize(d->stream));
            for (size_t k = 0; k < d->n; ++k)
                h[k] = static_cast<float>(static_cast<double>(h[k]) + static_cast<double>(ll[k]));
        }
        if (gamma) {
            SF_HIP(hipMemcpyAsync(gamma, d->g[d->cur_g], d->n * 4, hipMemcpyDeviceToHost, d->stream));
            SF_HIP(hipStreamSynchronize(d->stream));
        }
    });
}

extern "C" int sf_drainage_copy_field_f64(sf_drainage* ctx, double* h, double* gamma) {
    return guard([&] {
        SF_REQUIRE(ctx != nullptr, SF_ERR_INVALID, "ctx is NULL");
        Drainage* d = D(ctx);
        d->materialize();
        std::vector<float> hh(d->n), ll(d->n), gg(d->n);
        SF_HIP(hipMemcpyAsync(hh.data(), d->hi[d->cur_hi], d->n * 4, hipMemcpyDeviceToHost, d->stream));
        SF_HIP(hipMemcpyAsync(ll.data(), d->lo[d->cur], d->n * 4, hipMemcpyDeviceToHost, d->stream));
        SF_HIP(hipMemcpyAsync(gg.data(), d->g[d->cur_g], d->n * 4, hipMemcpyDeviceToHost, d->stream));
        SF_HIP(hipStreamSynchronize(d->stream));
        if (h)
            for (size_t k = 0; k < d->n; ++k)
                h[k] = static_cast<double>(hh[k]) + static_cast<double>(ll[k]);
        if (gamma)
            for (size_t k = 0; k < d->n; ++k) gamma[k] = static_cast<double>(gg[k]);
    });
}

// Index and weight arithmetic of DrainageSimulator::get_thickness
// (drainage.rs:516-552), FP64 on the host.
namespace {
struct QueryPlan {
    int64_t idx[4];
    double tf, pf;
};

size_t as_usize(double x) {
    if (!(x > 0.0)) return 0;
    if (x >= 18446744073709551615.0) return static_cast<size_t>(-1);
    return static_cast<size_t>(x);
}

QueryPlan plan_query(const Drainage& d, double theta, double phi) {
#pragma clang fp contract(off)
    const size_t nt = static_cast<size_t>(d.nt), np = static_cast<size_t>(d.np);
    const double theta_clamped = theta < 0.0 ? 0.0 : (theta > kPi ? kPi : theta);
    double r = std::fmod(phi, 2.0 * kPi);                 // f64::rem_euclid
    const double phi_normalized = r < 0.0 ? r + 2.0 * kPi : r;
    const double theta_continuous = theta_clamped / d.dtheta;
    const double phi_continuous = phi_normalized / d.dphi;
    size_t ti_lo = as_usize(std::floor(theta_continuous));
    if (ti_lo > nt - 2) ti_lo = nt - 2;
    const size_t pj_lo = as_usize(std::floor(phi_continuous)) % np;
    const size_t ti_hi = ti_lo + 1;
    const size_t pj_hi = (pj_lo + 1) % np;
    QueryPlan q;
    q.tf = theta_continuous - static_cast<double>(ti_lo);
    q.pf = phi_continuous - static_cast<double>(pj_lo);
    q.idx[0] = static_cast<int64_t>(ti_lo * np + pj_lo);
    q.idx[1] = static_cast<int64_t>(ti_lo * np + pj_hi);
    q.idx[2] = static_cast<int64_t>(ti_hi * np + pj_lo);
    q.idx[3] = static_cast<int64_t>(ti_hi * np + pj_hi);
    return q;
}

double interpolate(const QueryPlan& q, const float2* v) {
#pragma clang fp contract(off)
    double c[4];
    for (int k = 0; k < 4; ++k) c[k] = static_cast<double>(v[k].x) + static_cast<double>(v[k].y);
    const double v0 = c[0] * (1.0 - q.pf) + c[1] * q.pf;
    const double v1 = c[2] * (1.0 - q.pf) + c[3] * q.pf;
    return v0 * (1.0 - q.tf) + v1 * q.tf;
}
}  // namespace

extern "C" int sf_drainage_get_thickness_batch(sf_drainage* ctx, int32_t n, const double* theta,
                                               const double* phi, double* out) {
    return guard([&] {
        SF_REQUIRE(ctx != nullptr && n >= 0, SF_ERR_INVALID, "bad arguments");
        if (n == 0) return;
        SF_REQUIRE(theta && phi && out, SF_ERR_INVALID, "NULL array");
        Drainage* d = D(ctx);
        d->materialize();
        std::vector<QueryPlan> plans(n);
        std::vector<int64_t> idx(4 * static_cast<size_t>(n));
        for (int k = 0; k < n; ++k) {
            plans[k] = plan_query(*d, theta[k], phi[k]);
            for (int c = 0; c < 4; ++c) idx[4 * static_cast<size_t>(k) + c] = plans[k].idx[c];
        }
        const size_t m = idx.size();
        d->ensure_gather(m);
        SF_HIP(hipMemcpyAsync(d->gather_idx, idx.data(), m * sizeof(int64_t), hipMemcpyHostToDevice,
                              d->stream));
        hipLaunchKernelGGL(gather_kernel, dim3(cdiv(m, 256)), dim3(256), 0, d->stream,
                           d->hi[d->cur_hi], d->lo[d->cur], d->gather_idx, static_cast<int>(m),
                           d->gather_out);
        SF_HIP(hipGetLastError());
        std::vector<float2> v(m);
        SF_HIP(hipMemcpyAsync(v.data(), d->gather_out, m * sizeof(float2), hipMemcpyDeviceToHost,
                              d->stream));
        SF_HIP(hipStreamSynchronize(d->stream));
        for (int k = 0; k < n; ++k) out[k] = interpolate(plans[k], v.data() + 4 * static_cast<size_t>(k));
    });
}

extern "C" int sf_drainage_get_thickness(sf_drainage* ctx, double theta, double phi, double* out) {
    return sf_drainage_get_thickness_batch(ctx, 1, &theta, &phi, out);
}

extern "C" int sf_drainage_stats(sf_drainage* ctx, sf_stats* out) {
    return guard([&] {
        SF_REQUIRE(ctx != nullptr && out != nullptr, SF_ERR_INVALID, "NULL argument");
        Drainage* d = D(ctx);
        const StatsPart s = d->compute_stats();
        out->min_thickness = s.mn;
        out->max_thickness = s.mx;
        out->mass = s.mass;
        out->dt_cfl = s.lam > 0.0 ? 2.0 / s.lam : INFINITY;
        out->current_time = d->current_time;
        out->n_ruptured = static_cast<int64_t>(s.nrupt);
        out->n_tfe = static_cast<int32_t>(d->fronts.size());
        out->has_critical_region = s.mn < d->p.critical_thickness ? 1 : 0;
        out->steps = d->steps_taken;
    });
}

extern "C" int sf_drainage_reset(sf_drainage* ctx, double h0, double gamma0) {
    return guard([&] {
        SF_REQUIRE(ctx != nullptr, SF_ERR_INVALID, "ctx is NULL");
        Drainage* d = D(ctx);
        d->fill(h0, gamma0);
        d->current_time = 0.0;
        d->fronts.clear();
        d->next_spawn = d->p.tfe_first_spawn;
        d->steps_taken = 0;
        d->hist_count = 0;
    });
}

extern "C" int sf_drainage_set_marginal_regeneration(sf_drainage* ctx, int32_t enabled) {
    return guard([&] {
        SF_REQUIRE(ctx != nullptr, SF_ERR_INVALID, "ctx is NULL");
        Drainage* d = D(ctx);
        d->tfe_enabled = enabled != 0;
        if (enabled) d->p.flags |= SF_FLAG_TFE; else d->p.flags &= ~SF_FLAG_TFE;
        if (!enabled) d->fronts.clear();
    });
}

extern "C" int sf_drainage_set_tfe_thickness_ratio(sf_drainage* ctx, double ratio) {
    return guard([&] {
        SF_REQUIRE(ctx != nullptr, SF_ERR_INVALID, "ctx is NULL");
        D(ctx)->p.tfe_thickness_ratio = ratio < 0.5 ? 0.5 : (ratio > 0.95 ? 0.95 : ratio);
    });
}

extern "C" int sf_drainage_active_tfe_count(const sf_drainage* ctx, int32_t* out) {
    return guard([&] {
        SF_REQUIRE(ctx != nullptr && out != nullptr, SF_ERR_INVALID, "NULL argument");
        *out = static_cast<int32_t>(D(ctx)->fronts.size());
    });
}

extern "C" int sf_drainage_current_time(const sf_drainage* ctx, double* out) {
    return guard([&] {
        SF_REQUIRE(ctx != nullptr && out != nullptr, SF_ERR_INVALID, "NULL argument");
        *out = D(ctx)->current_time;
    });
}

extern "C" int sf_drainage_history(sf_drainage* ctx, int32_t n, double* mass, double* dt_cfl,
                                   int32_t* n_out) {
    return guard([&] {
        SF_REQUIRE(ctx != nullptr && n >= 0, SF_ERR_INVALID, "bad arguments");
        Drainage* d = D(ctx);
        d->finalize_diag();
        const int64_t avail = std::min<int64_t>(d->hist_count, Drainage::kHistCap);
        const int64_t m = std::min<int64_t>(avail, n);
        std::vector<double2> all(Drainage::kHistCap);
        SF_HIP(hipMemcpyAsync(all.data(), d->diag_hist, Drainage::kHistCap * sizeof(double2),
                              hipMemcpyDeviceToHost, d->stream));
        SF_HIP(hipStreamSynchronize(d->stream));
        for (int64_t k = 0; k < m; ++k) {
            const int64_t step = d->hist_count - m + k;
            const double2 e = all[static_cast<size_t>(step % Drainage::kHistCap)];
            if (mass) mass[k] = e.x;
            if (dt_cfl) dt_cfl[k] = e.y;
        }
        if (n_out) *n_out = static_cast<int32_t>(m);
    });
}

extern "C" int sf_drainage_tfe_mask(sf_drainage* ctx, uint8_t* mask) {
    return guard([&] {
        SF_REQUIRE(ctx != nullptr && mask != nullptr, SF_ERR_INVALID, "NULL argument");
        Drainage* d = D(ctx);
        StepArgs a{};
        d->fill_common_args(a, 0.f);
        a.nfronts = static_cast<int>(d->last_boxes.size());
        for (int k = 0; k < a.nfronts; ++k) a.fr[k] = d->last_boxes[k];
        if (d->fronts.empty()) a.nfronts = 0;
        uint8_t* dm = dalloc<uint8_t>(d->n);
        hipLaunchKernelGGL(tfe_mask_kernel, dim3(cdiv(d->np, 256), d->nt), dim3(256), 0, d->stream,
                           a, dm);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipMemcpyAsync(mask, dm, d->n, hipMemcpyDeviceToHost, d->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(d->stream);
        dfree(dm);
        SF_HIP(e);
    });
}

extern "C" int sf_drainage_rupture_mask(sf_drainage* ctx, uint8_t* mask) {
    return guard([&] {
        SF_REQUIRE(ctx != nullptr && mask != nullptr, SF_ERR_INVALID, "NULL argument");
        Drainage* d = D(ctx);
        d->materialize();
        const Float3 hc = split3(d->p.critical_thickness);
        uint8_t* dm = dalloc<uint8_t>(d->n);
        hipLaunchKernelGGL(rupture_mask_kernel, dim3(cdiv(d->n, 256)), dim3(256), 0, d->stream,
                           d->hi[d->cur_hi], d->lo[d->cur], d->n, hc.c0, hc.c1, hc.c2, dm);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipMemcpyAsync(mask, dm, d->n, hipMemcpyDeviceToHost, d->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(d->stream);
        dfree(dm);
        SF_HIP(e);
    });
}

extern "C" int sf_drainage_device_fields(sf_drainage* ctx, const float** h_hi, const float** h_lo,
                                         const float** gamma) {
    return guard([&] {
        SF_REQUIRE(ctx != nullptr, SF_ERR_INVALID, "ctx is NULL");
        Drainage* d = D(ctx);
        d->materialize();
        if (h_hi) *h_hi = d->hi[d->cur_hi];
        if (h_lo) *h_lo = d->lo[d->cur];
        if (gamma) *gamma = d->g[d->cur_g];
    });
}

extern "C" int sf_drainage_synchronize(sf_drainage* ctx) {
    return guard([&] {
        SF_REQUIRE(ctx != nullptr, SF_ERR_INVALID, "ctx is NULL");
        SF_HIP(hipStreamSynchronize(D(ctx)->stream));
    });
}

extern "C" int sf_drainage_stream(sf_drainage* ctx, void** stream) {
    return guard([&] {
        SF_REQUIRE(ctx != nullptr && stream != nullptr, SF_ERR_INVALID, "NULL argument");
        *stream = reinterpret_cast<void*>(D(ctx)->stream);
    });
}

static void timer_query(Drainage* d, KernelTimer& t, int32_t enable, double* avg_ms,
                        int64_t* launches) {
    t.drain(d->stream);
    if (avg_ms) *avg_ms = t.launches ? t.ms / static_cast<double>(t.launches) : 0.0;
    if (launches) *launches = t.launches;
    t.ms = 0.0;
    t.launches = 0;
    t.enabled = enable != 0;
}

extern "C" int sf_drainage_kernel_timing(sf_drainage* ctx, int32_t enable, double* avg_ms,
                                         int64_t* launches) {
    return guard([&] {
        SF_REQUIRE(ctx != nullptr, SF_ERR_INVALID, "ctx is NULL");
        Drainage* d = D(ctx);
        timer_query(d, d->step_timer, enable, avg_ms, launches);
    });
}

extern "C" int sf_spectral_kernel_timing(sf_drainage* ctx, int32_t enable, double* avg_ms,
                                         int64_t* launches) {
    return guard([&] {
        SF_REQUIRE(ctx != nullptr, SF_ERR_INVALID, "ctx is NULL");
        Drainage* d = D(ctx);
        timer_query(d, d->spec_timer, enable, avg_ms, launches);
    });
}

extern "C" int sf_drainage_set_steps_per_launch(sf_drainage* ctx, int32_t steps) {
    return guard([&] {
        SF_REQUIRE(ctx != nullptr, SF_ERR_INVALID, "ctx is NULL");
        SF_REQUIRE(steps == 1 || steps == 2, SF_ERR_INVALID, "steps per launch must be 1 or 2, got %d", steps);
        D(ctx)->steps_per_launch = steps;
    });
}

extern "C" int sf_drainage_get_steps_per_launch(sf_drainage* ctx, int32_t* steps) {
    return guard([&] {
        SF_REQUIRE(ctx != nullptr && steps != nullptr, SF_ERR_INVALID, "bad arguments");
        const Drainage* d = D(ctx);
        *steps = (d->steps_per_launch == 2 && d->two_step_ok()) ? 2 : 1;
    });
}

extern "C" int sf_spectral_set_operand(sf_drainage* ctx, int32_t kind) {
    return guard([&] {
        SF_REQUIRE(ctx != nullptr, SF_ERR_INVALID, "ctx is NULL");
        SF_REQUIRE(kind == SF_SPECTRAL_OPERAND_AUTO || kind == SF_SPECTRAL_OPERAND_BF16_PAIR ||
                       kind == SF_SPECTRAL_OPERAND_FP16,
                   SF_ERR_INVALID, "unknown spectral operand format %d", kind);
        D(ctx)->spec_operand = kind;
    });
}

extern "C" int sf_spectral_get_operand(sf_drainage* ctx, int32_t* kind) {
    return guard([&] {
        SF_REQUIRE(ctx != nullptr && kind != nullptr, SF_ERR_INVALID, "bad arguments");
        const Drainage* d = D(ctx);
        if (!d->cmf_dev)
            *kind = SF_SPECTRAL_OPERAND_AUTO;
        else if (d->spec_operand == SF_SPECTRAL_OPERAND_AUTO)
            *kind = (d->cmf_f16_range && d->cmf_f16_bound) ? SF_SPECTRAL_OPERAND_FP16 : SF_SPECTRAL_OPERAND_BF16_PAIR;
        else
            *kind = d->spec_operand;
    });
}

// ---------------------------------------------------------------------------
// Allocation-free field statistics (sf_field_stats_noalloc).
//
// The kernels of sf_drainage_stats, run on a caller-held device field with
// their per-row FP64 tables and per-block partials in `scratch`, a fixed
// array that every device the library runs on holds until the process ends.
// In host code `scratch` names a host-side stand-in, not device memory: the
// device address comes from hipGetSymbolAddress, once per device, and only
// that address is handed to kernels.

namespace {

constexpr int kNoallocMaxTheta = SF_NOALLOC_MAX_THETA;
constexpr size_t kNoallocRowBytes = static_cast<size_t>(kNoallocMaxTheta) * sizeof(RowD);
constexpr size_t kNoallocPartBytes = Drainage::kStatsBlocks * sizeof(StatsPart);
constexpr size_t kNoallocBytes = kNoallocRowBytes + kNoallocPartBytes + sizeof(StatsPart);
static_assert(kNoallocRowBytes % 256 == 0 && kNoallocPartBytes % 256 == 0, "scratch regions stay aligned");

static __device__ __align__(256) unsigned char scratch[kNoallocBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device (looked up on the first
// call there).
unsigned char* device_scratch() {
    int dev = 0;
    SF_HIP(hipGetDevice(&dev));
    SF_REQUIRE(dev >= 0 && dev < kMaxDevices, SF_ERR_DEVICE, "device index %d out of range", dev);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        SF_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        SF_REQUIRE(p != nullptr, SF_ERR_DEVICE, "no device address for the statistics scratch array");
        g_scratch_ptr[dev] = static_cast<unsigned char*>(p);
    }
    return g_scratch_ptr[dev];
}

// The FP64 row table of Drainage::upload_tables (cell area, CFL grid term,
// |sin theta|), one thread per row.
__global__ void noalloc_rows_kernel(RowD* rowd, int nt, double dtheta, double dphi, double R2, double D,
                                    double sin_guard) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nt) return;
    const double theta = static_cast<double>(i) * dtheta;
    const double s = sin(theta);
    const double ss = fabs(s) < sin_guard ? copysign(sin_guard, s) : s;
    const double lo_b = i == 0 ? 0.0 : theta - 0.5 * dtheta;
    const double hi_b = i == nt - 1 ? kPi : theta + 0.5 * dtheta;
    RowD r;
    r.area = R2 * (cos(lo_b) - cos(hi_b)) * dphi;
    r.lam_grid = 4.0 * D * (1.0 / (dtheta * dtheta) + 1.0 / (ss * ss * dphi * dphi)) / R2;
    r.sin_abs = fabs(s);
    r.pad = 0.0;
    rowd[i] = r;
}

// sf_stats of the folded partials, as sf_drainage_stats fills it for a field
// with no context (time, fronts and steps 0).
__global__ void noalloc_emit_kernel(const StatsPart* r, double h_crit, sf_stats* out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const StatsPart s = *r;
    sf_stats o;
    o.min_thickness = s.mn;
    o.max_thickness = s.mx;
    o.mass = s.mass;
    o.dt_cfl = s.lam > 0.0 ? 2.0 / s.lam : INFINITY;
    o.current_time = 0.0;
    o.n_ruptured = static_cast<int64_t>(s.nrupt);
    o.n_tfe = 0;
    o.has_critical_region = s.mn < h_crit ? 1 : 0;
    o.steps = 0;
    *out = o;
}

}  // namespace

extern "C" int sf_field_stats_noalloc(const sf_params* params, int32_t n_theta, int32_t n_phi, const float* h_hi,
                                      const float* h_lo, void* stream, sf_stats* out) {
    return guard([&] {
        check_params(params);
        SF_REQUIRE(h_hi != nullptr && out != nullptr, SF_ERR_INVALID, "NULL argument");
        SF_REQUIRE(n_theta >= 3 && n_phi >= 1, SF_ERR_INVALID, "grid %d x %d: need n_theta >= 3 and n_phi >= 1",
                   n_theta, n_phi);
        SF_REQUIRE(n_theta <= kNoallocMaxTheta, SF_ERR_INVALID,
                   "n_theta %d exceeds SF_NOALLOC_MAX_THETA (%d), the row capacity of the fixed statistics scratch",
                   n_theta, kNoallocMaxTheta);
        SF_REQUIRE(static_cast<int64_t>(n_theta) * n_phi < (int64_t(1) << 40), SF_ERR_INVALID, "grid too large");
        require_device();
        unsigned char* base = device_scratch();
        RowD* rowd = reinterpret_cast<RowD*>(base);
        StatsPart* part = reinterpret_cast<StatsPart*>(base + kNoallocRowBytes);
        StatsPart* folded = reinterpret_cast<StatsPart*>(base + kNoallocRowBytes + kNoallocPartBytes);
        hipStream_t s = static_cast<hipStream_t>(stream);
        const double dtheta = kPi / static_cast<double>(n_theta - 1);
        const double dphi = 2.0 * kPi / static_cast<double>(n_phi);
        double D = params->diffusion_coeff;
        if ((params->flags & SF_FLAG_SURFACTANT) && params->surfactant_diffusion > D) D = params->surfactant_diffusion;
        hipLaunchKernelGGL(noalloc_rows_kernel, dim3(cdiv(n_theta, 256)), dim3(256), 0, s, rowd, n_theta, dtheta,
                           dphi, params->bubble_radius * params->bubble_radius, D, params->sin_guard);
        const size_t n = static_cast<size_t>(n_theta) * n_phi;
        const size_t blocks = std::min<size_t>(Drainage::kStatsBlocks, std::max<size_t>(1, cdiv(n, 256)));
        const Float3 hc = split3(params->critical_thickness);
        const double coef3 = 3.0 * (params->density * params->gravity / (3.0 * params->viscosity));
        hipLaunchKernelGGL(stats_partial_kernel, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, s, h_hi, h_lo,
                           rowd, n_theta, n_phi, hc.c0, hc.c1, hc.c2, coef3, part);
        hipLaunchKernelGGL(stats_final_kernel, dim3(1), dim3(64), 0, s, part, static_cast<int>(blocks), folded);
        hipLaunchKernelGGL(noalloc_emit_kernel, dim3(1), dim3(64), 0, s, folded, params->critical_thickness, out);
        SF_HIP(hipGetLastError());
    });
}
