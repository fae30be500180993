// drainage_kernels.hpp -- device code of the drainage step (gfx950).
//
// State layout in HBM (row-major [theta][phi], index i * n_phi + j):
//   hi[2], lo[2]  film thickness h = hi + lo (m) as an FP32 float-float
//                 pair.  lo is rewritten every step; hi only on rebase steps,
//                 so a step streams 12 B in and 8 B out per cell (+4 B on
//                 rebase steps).
//   g[2]          surfactant concentration Γ in [0, 1], FP32.
// One explicit Euler step (drainage.rs:298-392 for h, drainage.wgsl:150-201
// for Γ and the Marangoni term) is one launch.  Each block owns a tile of
// 64*wx consecutive phi columns (one wave per 64 columns) and marches
// rows_per_block theta rows, carrying rows i-1, i, i+1 in registers; phi
// neighbours come from the adjacent lanes (cross-lane shuffles) and only the
// two wave-edge lanes load their wrapped neighbour column.
#pragma once

#include "sf_common.hpp"

namespace sf {

constexpr int kMaxFronts = 8;
constexpr int kBlockMax = 256;

// Rows [ilo, ihi] x cyclic columns [j0, j0 + len) of one marginal-regeneration
// front; the host derives them from the exact FP64 test of TfeFront::contains
// (drainage.rs:55-71), so membership matches the reference bit for bit.
struct FrontBox {
    int ilo, ihi, j0, len;
};

struct RowF {          // per theta row, rounded from FP64 on the host
    float sin_t;       // sin(theta_i)                        (drainage.rs:316)
    float cot;         // cos(theta_i) / sin_safe             (drainage.rs:367)
    float inv_s2;      // 1 / sin_safe^2                      (drainage.rs:368)
    float inv_s;       // 1 / sin_safe                        (drainage.wgsl:160)
};

struct RowD {          // per theta row, FP64 (diagnostics only)
    double area;       // cell area of the row, m^2
    double lam_grid;   // 4 D (1/dth^2 + 1/(sin_safe^2 dph^2)) / R^2
    double sin_abs;    // |sin(theta_i)|
    double pad;
};

struct StepArgs {
    const float* hi_in;
    const float* lo_in;
    const float* g_in;
    float* hi_out;     // written on rebase steps only
    float* lo_out;
    float* g_out;
    const RowF* rows;
    const RowD* rowd;
    const float2* hpole_in;   // ring partial sums of the previous step: [2][ntiles]
    float2* hpole_out;
    const float2* gpole_in;
    float2* gpole_out;
    double2* diag_part;       // per block {mass, max lambda}
    unsigned int* ticket;
    double2* diag_hist;       // ring buffer {mass, dt_cfl} per step
    int hist_slot;
    int nt, np, ntiles, rows_per_block;
    int h_pole_mean;          // ring-1 theta neighbour = mean from hpole_in
    int g_pole_mean;
    int h_pole_sum;           // produce ring partials for the next step
    int g_pole_sum;
    float dt;
    float inv_dth2, inv_2dth, inv_dph2, inv_2dph;
    float d_r2;               // D_h / R^2
    float ds_r2;              // D_s / R^2
    float coef;               // rho g / (3 eta)
    float mgr;                // marangoni_coeff * gamma_reduction
    float inv_r;
    float hc0, hc1, hc2;      // critical thickness as a three-float expansion
    float tfe_g_hi, tfe_g_lo; // (1 - rate dt) - 1 as a float pair
    float np_f;
    int nfronts;
    FrontBox fr[kMaxFronts];
    double coef3;             // 3 rho g / (3 eta), FP64
    double pole_area_top, pole_area_bot;
    double pole_mass_fixed;   // film volume of the pole rows when they are copied through
};

// Sum of the ring partials of one pole, divided by n_phi, as a float pair.
__device__ inline float2 pole_mean(const float2* part, int ntiles, float np_f) {
    float2 s = make_float2(0.f, 0.f);
    for (int t = 0; t < ntiles; ++t) s = ff_add(s, part[t]);
    float q = s.x / np_f;
    float r = fmaf(-q, np_f, s.x);
    float ql = (r + s.y) / np_f;
    float h, l;
    fast_two_sum(q, ql, h, l);
    return make_float2(h, l);
}

__device__ inline float2 wave_ff_sum(float2 v) {
    for (int off = 32; off > 0; off >>= 1) {
        float2 o;
        o.x = __shfl_down(v.x, off);
        o.y = __shfl_down(v.y, off);
        v = ff_add(v, o);
    }
    return v;
}

__device__ inline double wave_dsum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
    return v;
}

__device__ inline double wave_dmax(double v) {
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_down(v, off));
    return v;
}

__device__ inline bool in_fronts(const StepArgs& a, int i, int j) {
    bool hit = false;
    for (int k = 0; k < a.nfronts; ++k) {
        const FrontBox f = a.fr[k];
        if (i >= f.ilo && i <= f.ihi) {
            int d = j - f.j0;
            if (d < 0) d += a.np;
            hit |= d < f.len;
        }
    }
    return hit;
}

__device__ inline double atomic_ld(const double* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ inline float atomic_ldf(const float* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// GMODE: 0 = Γ unused, 1 = Γ read (Marangoni) but frozen, 2 = Γ evolves.
template <bool REBASE, int GMODE, bool MARANGONI, bool DIAG>
__global__ __launch_bounds__(kBlockMax) void drainage_step_kernel(const StepArgs a) {
    constexpr bool GREAD = GMODE >= 1;
    constexpr bool GWRITE = GMODE == 2;
    const int np = a.np, nt = a.nt;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int nwaves = blockDim.x >> 6;
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = j < np;
    const int jc = valid ? j : np - 1;
    const int jm = jc == 0 ? np - 1 : jc - 1;
    const int jp = jc + 1 == np ? 0 : jc + 1;
    const bool ledge = lane == 0;
    const bool redge = lane == 63 || j + 1 >= np;
    const int r0 = 1 + blockIdx.y * a.rows_per_block;
    const int r1 = min(r0 + a.rows_per_block, nt - 1);
    const bool touches_top = r0 == 1;
    const bool touches_bot = r1 == nt - 1;

    __shared__ float2 red_ff[4][kBlockMax / 64];
    __shared__ double red_d[2][kBlockMax / 64];
    __shared__ int s_last;

    float2 ptop = make_float2(0.f, 0.f), pbot = make_float2(0.f, 0.f);
    float gtop = 0.f, gbot = 0.f;
    if (a.h_pole_mean) {
        if (touches_top) ptop = pole_mean(a.hpole_in, a.ntiles, a.np_f);
        if (touches_bot) pbot = pole_mean(a.hpole_in + a.ntiles, a.ntiles, a.np_f);
    }
    if (GREAD && a.g_pole_mean) {
        if (touches_top) { float2 m = pole_mean(a.gpole_in, a.ntiles, a.np_f); gtop = m.x + m.y; }
        if (touches_bot) { float2 m = pole_mean(a.gpole_in + a.ntiles, a.ntiles, a.np_f); gbot = m.x + m.y; }
    }

    const float* __restrict__ hi = a.hi_in;
    const float* __restrict__ lo = a.lo_in;
    const float* __restrict__ g = a.g_in;
    auto at = [np](int i, int jj) { return static_cast<size_t>(i) * np + jj; };

    // Row r0 - 1 (the north neighbour) and row r0.
    float ph, pl, pg = 0.f;
    if (r0 - 1 == 0 && a.h_pole_mean) {
        ph = ptop.x;
        pl = ptop.y;
    } else {
        ph = hi[at(r0 - 1, jc)];
        pl = lo[at(r0 - 1, jc)];
    }
    if (GREAD) pg = (r0 - 1 == 0 && a.g_pole_mean) ? gtop : g[at(r0 - 1, jc)];
    float ch = hi[at(r0, jc)], cl = lo[at(r0, jc)], cg = GREAD ? g[at(r0, jc)] : 0.f;

    float2 sum_top = make_float2(0.f, 0.f), sum_bot = make_float2(0.f, 0.f);
    float2 gsum_top = make_float2(0.f, 0.f), gsum_bot = make_float2(0.f, 0.f);
    double mass = 0.0, lam = 0.0;

    for (int i = r0; i < r1; ++i) {
        // South neighbour row i + 1.
        float nh, nl, ng = 0.f;
        if (i + 1 == nt - 1 && a.h_pole_mean) {
            nh = pbot.x;
            nl = pbot.y;
        } else {
            nh = hi[at(i + 1, jc)];
            nl = lo[at(i + 1, jc)];
        }
        if (GREAD) ng = (i + 1 == nt - 1 && a.g_pole_mean) ? gbot : g[at(i + 1, jc)];

        // West / east neighbours of row i with the periodic phi wrap
        // (drainage.rs:340-343).
        float mh = __shfl_up(ch, 1), ml = __shfl_up(cl, 1);
        float qh = __shfl_down(ch, 1), ql = __shfl_down(cl, 1);
        float mg = 0.f, qg = 0.f;
        if (GREAD) {
            mg = __shfl_up(cg, 1);
            qg = __shfl_down(cg, 1);
        }
        if (ledge) {
            mh = hi[at(i, jm)];
            ml = lo[at(i, jm)];
            if (GREAD) mg = g[at(i, jm)];
        }
        if (redge) {
            qh = hi[at(i, jp)];
            ql = lo[at(i, jp)];
            if (GREAD) qg = g[at(i, jp)];
        }

        const RowF rc = a.rows[i];
        float lo_new = cl, g_new = cg;
        const float H = ch + cl;
        // Rupture mask on the input state (drainage.rs:331-333).
        const bool rupt = ff_less(ch, cl, a.hc0, a.hc1, a.hc2);
        if (!rupt) {
            // Differences of float-float values: hi parts first (exact when
            // the neighbours are within a factor two), then the lo parts.
            const float dtp = (nh - ch) + (nl - cl);
            const float dtm = (ch - ph) + (cl - pl);
            const float dpp = (qh - ch) + (ql - cl);
            const float dpm = (ch - mh) + (cl - ml);
            // Surface Laplacian times R^2 (drainage.rs:355-369).
            const float lap = (dtp - dtm) * a.inv_dth2 + rc.cot * (dtp + dtm) * a.inv_2dth +
                              (dpp - dpm) * a.inv_dph2 * rc.inv_s2;
            // Gravity sink (drainage.rs:345-348) plus diffusion.
            float rate = a.d_r2 * lap - a.coef * rc.sin_t * (H * H * H);
            if (MARANGONI) {                             // drainage.wgsl:150-171
                const float gth = (ng - pg) * a.inv_2dth * a.inv_r;
                const float gph = (qg - mg) * a.inv_2dph * a.inv_r * rc.inv_s;
                const float mag = sqrtf(gth * gth + gph * gph);
                rate += a.mgr * mag * (H * H);
            }
            lo_new = cl + a.dt * rate;
            if (lo_new < -ch) lo_new = -ch;              // h >= 0 (drainage.rs:378)
            if (GWRITE) {                                // drainage.wgsl:180-201
                const float gtp = ng - cg, gtm = cg - pg, gpp = qg - cg, gpm = cg - mg;
                const float glap = (gtp - gtm) * a.inv_dth2 + rc.cot * (gtp + gtm) * a.inv_2dth +
                                   (gpp - gpm) * a.inv_dph2 * rc.inv_s2;
                float grate = a.ds_r2 * glap;
                if (MARANGONI) grate += 0.001f * lap * cg;
                g_new = fminf(fmaxf(cg + a.dt * grate, 0.f), 1.f);
            }
        }

        // Ring partials for the pole means (before the fronts act, as
        // update_poles runs before update_marginal_regeneration).
        if (a.h_pole_sum && valid && (i == 1 || i == nt - 2)) {
            float s, e;
            two_sum(ch, lo_new, s, e);
            if (i == 1) sum_top = ff_add(sum_top, make_float2(s, e));
            if (i == nt - 2) sum_bot = ff_add(sum_bot, make_float2(s, e));
        }
        if (GWRITE && a.g_pole_sum && valid && (i == 1 || i == nt - 2)) {
            if (i == 1) gsum_top = ff_add(gsum_top, make_float2(g_new, 0.f));
            if (i == nt - 2) gsum_bot = ff_add(gsum_bot, make_float2(g_new, 0.f));
        }

        // Marginal regeneration: h *= 1 - rate dt on front cells, rupture
        // ignored (drainage.rs:440-460).
        if (a.nfronts > 0 && in_fronts(a, i, jc)) {
            const float p = ch * a.tfe_g_hi;
            const float pe = fmaf(ch, a.tfe_g_hi, -p);
            const float corr = pe + (ch * a.tfe_g_lo + lo_new * a.tfe_g_hi);
            lo_new = lo_new + (p + corr);
        }

        if (valid) {
            const size_t c = at(i, j);
            if (REBASE) {
                float s, e;
                two_sum(ch, lo_new, s, e);
                a.hi_out[c] = s;
                a.lo_out[c] = e;
            } else {
                a.lo_out[c] = lo_new;
            }
            if (GWRITE) a.g_out[c] = g_new;
            if (DIAG) {
                const RowD rd = a.rowd[i];
                const double hv = static_cast<double>(ch) + static_cast<double>(lo_new);
                mass += hv * rd.area;
                lam = fmax(lam, rd.lam_grid + a.coef3 * rd.sin_abs * hv * hv);
            }
        }

        ph = ch; pl = cl; pg = cg;
        ch = nh; cl = nl; cg = ng;
    }

    // Block reduction of the ring partials.
    if (touches_top || touches_bot) {
        if (a.h_pole_sum) {
            sum_top = wave_ff_sum(sum_top);
            sum_bot = wave_ff_sum(sum_bot);
        }
        if (GWRITE && a.g_pole_sum) {
            gsum_top = wave_ff_sum(gsum_top);
            gsum_bot = wave_ff_sum(gsum_bot);
        }
        if (lane == 0) {
            red_ff[0][wave] = sum_top;
            red_ff[1][wave] = sum_bot;
            red_ff[2][wave] = gsum_top;
            red_ff[3][wave] = gsum_bot;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            float2 t[4];
            for (int q = 0; q < 4; ++q) {
                t[q] = red_ff[q][0];
                for (int w = 1; w < nwaves; ++w) t[q] = ff_add(t[q], red_ff[q][w]);
            }
            if (a.h_pole_sum) {
                if (touches_top) a.hpole_out[blockIdx.x] = t[0];
                if (touches_bot) a.hpole_out[a.ntiles + blockIdx.x] = t[1];
            }
            if (GWRITE && a.g_pole_sum) {
                if (touches_top) a.gpole_out[blockIdx.x] = t[2];
                if (touches_bot) a.gpole_out[a.ntiles + blockIdx.x] = t[3];
            }
        }
    }

    if (DIAG) {
        // Per-block FP64 partials, then the last block to finish reduces
        // them in block order (deterministic) into the history slot.
        mass = wave_dsum(mass);
        lam = wave_dmax(lam);
        if (lane == 0) {
            red_d[0][wave] = mass;
            red_d[1][wave] = lam;
        }
        __syncthreads();
        const unsigned nblocks = gridDim.x * gridDim.y;
        if (threadIdx.x == 0) {
            double m = red_d[0][0], l = red_d[1][0];
            for (int w = 1; w < nwaves; ++w) {
                m += red_d[0][w];
                l = fmax(l, red_d[1][w]);
            }
            a.diag_part[blockIdx.y * gridDim.x + blockIdx.x] = make_double2(m, l);
            __threadfence();
            const unsigned t = atomicAdd(a.ticket, 1u);
            s_last = (t == nblocks - 1);
        }
        __syncthreads();
        if (s_last) {
            __threadfence();
            double m = 0.0, l = 0.0;
            for (unsigned b = threadIdx.x; b < nblocks; b += blockDim.x) {
                m += atomic_ld(&a.diag_part[b].x);
                l = fmax(l, atomic_ld(&a.diag_part[b].y));
            }
            m = wave_dsum(m);
            l = wave_dmax(l);
            if (lane == 0) {
                red_d[0][wave] = m;
                red_d[1][wave] = l;
            }
            __syncthreads();
            if (threadIdx.x == 0) {
                double mt = red_d[0][0], lt = red_d[1][0];
                for (int w = 1; w < nwaves; ++w) {
                    mt += red_d[0][w];
                    lt = fmax(lt, red_d[1][w]);
                }
                if (a.h_pole_sum) {
                    double st = 0.0, sb = 0.0;
                    for (int t = 0; t < a.ntiles; ++t) {
                        st += static_cast<double>(atomic_ldf(&a.hpole_out[t].x)) +
                              static_cast<double>(atomic_ldf(&a.hpole_out[t].y));
                        sb += static_cast<double>(atomic_ldf(&a.hpole_out[a.ntiles + t].x)) +
                              static_cast<double>(atomic_ldf(&a.hpole_out[a.ntiles + t].y));
                    }
                    mt += a.pole_area_top * st + a.pole_area_bot * sb;
                } else {
                    mt += a.pole_mass_fixed;
                }
                a.diag_hist[a.hist_slot] = make_double2(mt, lt > 0.0 ? 2.0 / lt : INFINITY);
                __hip_atomic_store(a.ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
    }
}

}  // namespace sf
