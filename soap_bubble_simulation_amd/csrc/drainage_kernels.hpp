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
constexpr int kPrefetchRows = 2;
constexpr unsigned kXcds = 8;      // MI355X: 8 accelerator complex dies

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

// A diagnosed step whose per-block partials the next launch reduces: the
// partials, the step's ring partial sums (its pole-row film volume), and its
// history slot.
struct DiagPending {
    const double2* part;
    const float2* hpole;
    int nblocks;
    int pole_sum;
    int slot;
    int pad;
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
    double2* diag_part;       // this step's per-block {mass, max lambda}
    double2* diag_part2;      // two-step launches: the second step's partials
    DiagPending prev[2];      // steps of the previous launch not yet reduced
    int nprev;
    double2* diag_hist;       // ring buffer {mass, dt_cfl} per step
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
    // two-step launches: the second step's fronts, and the state between the
    // two steps (its low word and ring partial sums; null: not written)
    int nfronts2;
    FrontBox fr2[kMaxFronts];
    float* lo_mid;
    float2* hpole_mid;
    double coef3;             // 3 rho g / (3 eta), FP64
    double pole_area_top, pole_area_bot;
    double pole_mass_fixed;   // film volume of the pole rows when they are copied through
};

// A pole's summed ring partials divided by n_phi, as a float pair.
__device__ inline float2 pole_mean_of_sum(float2 s, float np_f) {
    float q = s.x / np_f;
    float r = fmaf(-q, np_f, s.x);
    float ql = (r + s.y) / np_f;
    float h, l;
    fast_two_sum(q, ql, h, l);
    return make_float2(h, l);
}

// Sum of the ring partials of one pole, divided by n_phi, as a float pair.
__device__ inline float2 pole_mean(const float2* part, int ntiles, float np_f) {
    float2 s = make_float2(0.f, 0.f);
    for (int t = 0; t < ntiles; ++t) s = ff_add(s, part[t]);
    return pole_mean_of_sum(s, np_f);
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

__device__ inline bool in_front_boxes(const FrontBox* fr, int nfronts, int np, int i, int j) {
    bool hit = false;
    for (int k = 0; k < nfronts; ++k) {
        const FrontBox f = fr[k];
        if (i >= f.ilo && i <= f.ihi) {
            int d = j - f.j0;
            if (d < 0) d += np;
            hit |= d < f.len;
        }
    }
    return hit;
}

__device__ inline bool in_fronts(const StepArgs& a, int i, int j) {
    return in_front_boxes(a.fr, a.nfronts, a.np, i, j);
}

// Thinning inside a front (drainage.rs:441-459): h <- h * factor with
// factor - 1 = tfe_g_hi + tfe_g_lo, added to the low word as a float-float
// product of (ch, lo) and the factor's offset.
__device__ __forceinline__ float front_thin(const StepArgs& a, float ch, float lo) {
    const float p = ch * a.tfe_g_hi;
    const float pe = fmaf(ch, a.tfe_g_hi, -p);
    const float corr = pe + (ch * a.tfe_g_lo + lo * a.tfe_g_hi);
    return lo + (p + corr);
}

__device__ inline double atomic_ld(const double* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ inline float atomic_ldf(const float* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Reduces one step's per-block FP64 partials in block order (deterministic)
// plus the pole rows, and stores {film volume, dt_cfl} in its history slot.
// Run by one block of the following launch, or by diag_finalize_kernel.
__device__ inline void diag_reduce(const double2* part, int nblocks, const float2* hpole, int ntiles,
                                   int pole_sum, double pole_area_top, double pole_area_bot,
                                   double pole_mass_fixed, double2* hist, int slot,
                                   double (&red_d)[2][kBlockMax / 64]) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwaves = blockDim.x >> 6;
    double m = 0.0, l = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += blockDim.x) {
        m += part[b].x;
        l = fmax(l, part[b].y);
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
        if (pole_sum) {
            double st = 0.0, sb = 0.0;
            for (int t = 0; t < ntiles; ++t) {
                st += static_cast<double>(hpole[t].x) + static_cast<double>(hpole[t].y);
                sb += static_cast<double>(hpole[ntiles + t].x) + static_cast<double>(hpole[ntiles + t].y);
            }
            mt += pole_area_top * st + pole_area_bot * sb;
        } else {
            mt += pole_mass_fixed;
        }
        hist[slot] = make_double2(mt, lt > 0.0 ? 2.0 / lt : INFINITY);
    }
    __syncthreads();
}

// Block 0 of a diagnosed launch reduces the previous launch's steps (one or
// two) into their history slots.
__device__ inline void diag_reduce_pending(const StepArgs& a, double (&red_d)[2][kBlockMax / 64]) {
    __syncthreads();
    diag_reduce(a.prev[0].part, a.prev[0].nblocks, a.prev[0].hpole, a.ntiles, a.prev[0].pole_sum,
                a.pole_area_top, a.pole_area_bot, a.pole_mass_fixed, a.diag_hist, a.prev[0].slot, red_d);
    if (a.nprev > 1)
        diag_reduce(a.prev[1].part, a.prev[1].nblocks, a.prev[1].hpole, a.ntiles, a.prev[1].pole_sum,
                    a.pole_area_top, a.pole_area_bot, a.pole_mass_fixed, a.diag_hist, a.prev[1].slot, red_d);
}

// Per-block FP64 partials of film volume and CFL lambda for this step (no
// fence and no atomics: they are consumed by the next launch); block 0 also
// reduces the previous step's partials into the history.
__device__ inline void diag_block_partial(double mass, double lam, double2* part, int lane, int wave,
                                          int nwaves, double (&red_d)[2][kBlockMax / 64]) {
    // the sum and the maximum in one pass, so the shuffles of both overlap
    for (int off = 32; off > 0; off >>= 1) {
        const double m = __shfl_down(mass, off);
        const double l = __shfl_down(lam, off);
        mass += m;
        lam = fmax(lam, l);
    }
    __syncthreads();
    if (lane == 0) {
        red_d[0][wave] = mass;
        red_d[1][wave] = lam;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = red_d[0][0], l = red_d[1][0];
        for (int w = 1; w < nwaves; ++w) {
            m += red_d[0][w];
            l = fmax(l, red_d[1][w]);
        }
        part[blockIdx.y * gridDim.x + blockIdx.x] = make_double2(m, l);
    }
}

__device__ inline void diag_epilogue(const StepArgs& a, double mass, double lam, int lane, int wave,
                                     int nwaves, double (&red_d)[2][kBlockMax / 64]) {
    diag_block_partial(mass, lam, a.diag_part, lane, wave, nwaves, red_d);
    if (a.nprev > 0 && blockIdx.x == 0 && blockIdx.y == 0) diag_reduce_pending(a, red_d);
}

// A lane's four cells of one row into its FP64 film volume and CFL lambda:
// the cells share the row's area and CFL terms, so their FP64 sum is
// multiplied once per row.
__device__ __forceinline__ void diag_row(const RowD& rd, double coef3, const float (&h)[4], const float (&l)[4],
                                         double& mass, double& lam) {
    double hsum = 0.0;
#pragma unroll
    for (int v = 0; v < 4; ++v) hsum += static_cast<double>(h[v]) + static_cast<double>(l[v]);
    const float hmax = fmaxf(fmaxf(h[0] + l[0], h[1] + l[1]), fmaxf(h[2] + l[2], h[3] + l[3]));
    mass = fma(hsum, rd.area, mass);
    const double hm = static_cast<double>(hmax);
    lam = fmax(lam, rd.lam_grid + coef3 * rd.sin_abs * hm * hm);
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
        if (a.nfronts > 0 && in_fronts(a, i, jc)) lo_new = front_thin(a, ch, lo_new);

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

    if (DIAG) diag_epilogue(a, mass, lam, lane, wave, nwaves, red_d);
}

// ---------------------------------------------------------------------------
// Wide variant for n_phi % 4 == 0: each lane owns four consecutive columns
// (one float4 per field and row), a wave spans 256 columns and a block of
// four waves 1024 columns x RB rows.  Rows are streamed with a prefetch
// depth of PF rows ahead of the row being updated.  Loads are issued without
// any branch (row index clamped into the grid, every lane also loading its
// wrapped west/east neighbour column), so the compiler never has to wait for
// them before their first use; the pole-row substitution is applied where a
// row is consumed.

struct Row4 {
    float4 h, l, g;      // the lane's four cells
    float hw, lw, gw;    // west neighbour of column 0
    float he, le, ge;    // east neighbour of column 3
};

template <int GMODE>
__device__ inline void load_row4(const StepArgs& a, int r, int cc, int jw, int je, Row4& R) {
    const size_t base = static_cast<size_t>(min(r, a.nt - 1)) * a.np;
    R.h = *reinterpret_cast<const float4*>(a.hi_in + base + cc);
    R.l = *reinterpret_cast<const float4*>(a.lo_in + base + cc);
    R.hw = a.hi_in[base + jw];
    R.lw = a.lo_in[base + jw];
    R.he = a.hi_in[base + je];
    R.le = a.lo_in[base + je];
    if (GMODE >= 1) {
        R.g = *reinterpret_cast<const float4*>(a.g_in + base + cc);
        R.gw = a.g_in[base + jw];
        R.ge = a.g_in[base + je];
    }
}

// Replaces a loaded row by the pole mean (all cells equal).
__device__ inline void pole_row(Row4& R, float2 P, float G, bool h, bool g) {
    if (h) {
        R.h = make_float4(P.x, P.x, P.x, P.x);
        R.l = make_float4(P.y, P.y, P.y, P.y);
        R.hw = R.he = P.x;
        R.lw = R.le = P.y;
    }
    if (g) {
        R.g = make_float4(G, G, G, G);
        R.gw = R.ge = G;
    }
}

template <bool REBASE, int GMODE, bool MARANGONI, bool DIAG, int RB, int PF>
__global__ __launch_bounds__(kBlockMax) void drainage_vec4_kernel(const StepArgs a) {
    constexpr bool GREAD = GMODE >= 1;
    constexpr bool GWRITE = GMODE == 2;
    const int np = a.np, nt = a.nt;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int nwaves = blockDim.x >> 6;
    // Blocks are dealt to the 8 XCDs round-robin; renumber them so that each
    // XCD walks a contiguous band of rows and the theta-neighbour rows it
    // re-reads stay in its own L2.
    const unsigned nblk = gridDim.x * gridDim.y;
    const unsigned b = blockIdx.y * gridDim.x + blockIdx.x;
    const unsigned per_xcd = (nblk + kXcds - 1) / kXcds;
    unsigned v = (b % kXcds) * per_xcd + b / kXcds;
    if (nblk % kXcds != 0) v = b;                 // uneven split: keep the launch order
    const unsigned bx = v % gridDim.x, by = v / gridDim.x;
    const int c0 = (static_cast<int>(bx) * nwaves + wave) * 256 + 4 * lane;
    const bool valid = c0 < np;
    const int cc = valid ? c0 : np - 4;
    const int jw = cc == 0 ? np - 1 : cc - 1;
    const int je = cc + 4 == np ? 0 : cc + 4;
    const int r0 = 1 + static_cast<int>(by) * RB;
    const int r1 = min(r0 + RB, nt - 1);          // rows [r0, r1) are updated
    const bool touches_top = r0 == 1;
    const bool touches_bot = r1 == nt - 1;

    __shared__ float2 red_ff[4][kBlockMax / 64];
    __shared__ double red_d[2][kBlockMax / 64];

    float2 ptop = make_float2(0.f, 0.f), pbot = make_float2(0.f, 0.f);
    float gtop = 0.f, gbot = 0.f;
    if (a.h_pole_mean) {
        if (touches_top) ptop = pole_mean(a.hpole_in, a.ntiles, a.np_f);
        if (touches_bot) pbot = pole_mean(a.hpole_in + a.ntiles, a.ntiles, a.np_f);
    }
    if (GREAD && a.g_pole_mean) {
        if (touches_top) { const float2 m = pole_mean(a.gpole_in, a.ntiles, a.np_f); gtop = m.x + m.y; }
        if (touches_bot) { const float2 m = pole_mean(a.gpole_in + a.ntiles, a.ntiles, a.np_f); gbot = m.x + m.y; }
    }

    // The band's per-row coefficients go to LDS before any row is streamed,
    // so that reading them inside the row loop never waits on the vector
    // memory counter that the prefetched rows are queued on.
    __shared__ RowF s_rowf[RB];
    __shared__ RowD s_rowd[DIAG ? RB : 1];
    if (threadIdx.x < RB) {
        const int ri = min(r0 + static_cast<int>(threadIdx.x), nt - 1);
        s_rowf[threadIdx.x] = a.rows[ri];
        if (DIAG) s_rowd[threadIdx.x] = a.rowd[ri];
    }

    // Slot k holds row r0 - 1 + k; slots 0 .. RB+1.
    Row4 S[RB + 2];
#pragma unroll
    for (int k = 0; k < PF + 2 && k < RB + 2; ++k) load_row4<GMODE>(a, r0 - 1 + k, cc, jw, je, S[k]);
    __syncthreads();

    const float G0 = a.inv_dth2;
    float top_h[4] = {0.f, 0.f, 0.f, 0.f}, top_l[4] = {0.f, 0.f, 0.f, 0.f}, top_g[4] = {0.f, 0.f, 0.f, 0.f};
    float bot_h[4] = {0.f, 0.f, 0.f, 0.f}, bot_l[4] = {0.f, 0.f, 0.f, 0.f}, bot_g[4] = {0.f, 0.f, 0.f, 0.f};
    double mass = 0.0, lam = 0.0;

#pragma unroll
    for (int t = 0; t < RB; ++t) {
        if (t + PF + 2 < RB + 2) load_row4<GMODE>(a, r0 + t + PF + 1, cc, jw, je, S[t + PF + 2]);
        const int i = r0 + t;
        if (i < r1) {
        Row4 N = S[t];               // row i - 1
        const Row4 C = S[t + 1];     // row i
        Row4 Sx = S[t + 2];          // row i + 1
        if (i == 1) pole_row(N, ptop, gtop, a.h_pole_mean, GREAD && a.g_pole_mean);
        if (i == nt - 2) pole_row(Sx, pbot, gbot, a.h_pole_mean, GREAD && a.g_pole_mean);
        const RowF rc = s_rowf[t];
        const float G1 = rc.cot * a.inv_2dth;
        const float G2 = rc.inv_s2 * a.inv_dph2;
        const float dtD = a.dt * a.d_r2;
        const float cS = a.dt * a.coef * rc.sin_t;

        // West / east neighbours of the lane's first and last column.
        const float hw = C.hw, lw = C.lw, he = C.he, le = C.le;
        const float gw = GREAD ? C.gw : 0.f, ge = GREAD ? C.ge : 0.f;

        const float ch[4] = {C.h.x, C.h.y, C.h.z, C.h.w};
        const float cl[4] = {C.l.x, C.l.y, C.l.z, C.l.w};
        const float nh[4] = {N.h.x, N.h.y, N.h.z, N.h.w};
        const float nl[4] = {N.l.x, N.l.y, N.l.z, N.l.w};
        const float sh[4] = {Sx.h.x, Sx.h.y, Sx.h.z, Sx.h.w};
        const float sl[4] = {Sx.l.x, Sx.l.y, Sx.l.z, Sx.l.w};
        const float wh[4] = {hw, C.h.x, C.h.y, C.h.z};
        const float wl[4] = {lw, C.l.x, C.l.y, C.l.z};
        const float eh[4] = {C.h.y, C.h.z, C.h.w, he};
        const float el[4] = {C.l.y, C.l.z, C.l.w, le};
        float cg[4] = {0.f, 0.f, 0.f, 0.f}, ng[4] = {0.f, 0.f, 0.f, 0.f}, sg[4] = {0.f, 0.f, 0.f, 0.f};
        float wg[4] = {0.f, 0.f, 0.f, 0.f}, eg[4] = {0.f, 0.f, 0.f, 0.f};
        if (GREAD) {
            cg[0] = C.g.x; cg[1] = C.g.y; cg[2] = C.g.z; cg[3] = C.g.w;
            ng[0] = N.g.x; ng[1] = N.g.y; ng[2] = N.g.z; ng[3] = N.g.w;
            sg[0] = Sx.g.x; sg[1] = Sx.g.y; sg[2] = Sx.g.z; sg[3] = Sx.g.w;
            wg[0] = gw; wg[1] = C.g.x; wg[2] = C.g.y; wg[3] = C.g.z;
            eg[0] = C.g.y; eg[1] = C.g.z; eg[2] = C.g.w; eg[3] = ge;
        }

        float lo_new[4], g_new[4];
#pragma unroll
        for (int v = 0; v < 4; ++v)
            cell_step<GMODE, MARANGONI>(a, rc, G0, G1, G2, dtD, cS, ch[v], cl[v], nh[v], nl[v], sh[v], sl[v], wh[v],
                                        wl[v], eh[v], el[v], cg[v], ng[v], sg[v], wg[v], eg[v], lo_new[v], g_new[v]);

        // Keep the ring rows' new values (before the fronts act, as
        // update_poles runs before update_marginal_regeneration) for the
        // pole means; they are reduced once after the row loop.
        if (i == 1) {
#pragma unroll
            for (int v = 0; v < 4; ++v) { top_h[v] = ch[v]; top_l[v] = lo_new[v]; top_g[v] = g_new[v]; }
        }
        if (i == nt - 2) {
#pragma unroll
            for (int v = 0; v < 4; ++v) { bot_h[v] = ch[v]; bot_l[v] = lo_new[v]; bot_g[v] = g_new[v]; }
        }

        if (a.nfronts > 0) {
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                if (in_fronts(a, i, cc + v)) lo_new[v] = front_thin(a, ch[v], lo_new[v]);
            }
        }

        if (valid) {
            const size_t off = static_cast<size_t>(i) * np + cc;
            if (REBASE) {
                float4 hh, ll;
                two_sum(ch[0], lo_new[0], hh.x, ll.x);
                two_sum(ch[1], lo_new[1], hh.y, ll.y);
                two_sum(ch[2], lo_new[2], hh.z, ll.z);
                two_sum(ch[3], lo_new[3], hh.w, ll.w);
                *reinterpret_cast<float4*>(a.hi_out + off) = hh;
                *reinterpret_cast<float4*>(a.lo_out + off) = ll;
            } else {
                *reinterpret_cast<float4*>(a.lo_out + off) =
                    make_float4(lo_new[0], lo_new[1], lo_new[2], lo_new[3]);
            }
            if (GWRITE)
                *reinterpret_cast<float4*>(a.g_out + off) = make_float4(g_new[0], g_new[1], g_new[2], g_new[3]);
            if (DIAG) diag_row(s_rowd[DIAG ? t : 0], a.coef3, ch, lo_new, mass, lam);
        }
        }
    }

    if (touches_top || touches_bot) {
        float2 sum_top = make_float2(0.f, 0.f), sum_bot = make_float2(0.f, 0.f);
        float2 gsum_top = make_float2(0.f, 0.f), gsum_bot = make_float2(0.f, 0.f);
        if (valid) {
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                float s_, e_;
                two_sum(top_h[v], top_l[v], s_, e_);
                sum_top = ff_add(sum_top, make_float2(s_, e_));
                two_sum(bot_h[v], bot_l[v], s_, e_);
                sum_bot = ff_add(sum_bot, make_float2(s_, e_));
                gsum_top = ff_add(gsum_top, make_float2(top_g[v], 0.f));
                gsum_bot = ff_add(gsum_bot, make_float2(bot_g[v], 0.f));
            }
        }
        sum_top = wave_ff_sum(sum_top);
        sum_bot = wave_ff_sum(sum_bot);
        if (GWRITE) {
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
                if (touches_top) a.hpole_out[bx] = t[0];
                if (touches_bot) a.hpole_out[a.ntiles + bx] = t[1];
            }
            if (GWRITE && a.g_pole_sum) {
                if (touches_top) a.gpole_out[bx] = t[2];
                if (touches_bot) a.gpole_out[a.ntiles + bx] = t[3];
            }
        }
    }

    if (DIAG) diag_epilogue(a, mass, lam, lane, wave, nwaves, red_d);
}
// One cell's step from its centre (c), north (n, theta - 1), south (s), west
// (w, phi - 1) and east (e) values, h as float-float pairs (h, l) and Γ (g):
// the low word's update lo_new and Γ's (drainage.rs:331-378,
// drainage.wgsl:150-201).  Rupture mask on the input state
// (drainage.rs:331-333): frozen cells keep h and Γ; the update is computed
// for every cell and selected, which keeps the wave free of divergent
// branches.  Shared by the one- and two-step band kernels so both do the same
// arithmetic.
template <int GMODE, bool MARANGONI>
__device__ __forceinline__ void cell_step(const StepArgs& a, const RowF& rc, float G0, float G1, float G2, float dtD,
                                          float cS, float ch, float cl, float nh, float nl, float sh, float sl, float wh,
                                          float wl, float eh, float el, float cg, float ng, float sg, float wg, float eg,
                                          float& lo_new, float& g_new) {
    constexpr bool GWRITE = GMODE == 2;
    const bool rupt = ff_less(ch, cl, a.hc0, a.hc1, a.hc2);
    const float dtp = (sh - ch) + (sl - cl);   // toward theta + 1
    const float dtm = (ch - nh) + (cl - nl);
    const float dpp = (eh - ch) + (el - cl);
    const float dpm = (ch - wh) + (cl - wl);
    const float lap = (dtp - dtm) * G0 + (dtp + dtm) * G1 + (dpp - dpm) * G2;
    const float H = ch + cl;
    float rate_dt = fmaf(dtD, lap, -cS * (H * H * H));
    if (MARANGONI) {
        const float gth = (sg - ng) * a.inv_2dth * a.inv_r;
        const float gph = (eg - wg) * a.inv_2dph * a.inv_r * rc.inv_s;
        const float mag = sqrtf(gth * gth + gph * gph);
        rate_dt += a.dt * a.mgr * mag * (H * H);
    }
    const float ln = fmaxf(cl + rate_dt, -ch);          // h >= 0 (drainage.rs:378)
    lo_new = rupt ? cl : ln;
    g_new = cg;
    if (GWRITE) {
        const float gtp = sg - cg, gtm = cg - ng;
        const float gpp = eg - cg, gpm = cg - wg;
        const float glap = (gtp - gtm) * G0 + (gtp + gtm) * G1 + (gpp - gpm) * G2;
        float grate = a.ds_r2 * glap;
        if (MARANGONI) grate += 0.001f * lap * cg;
        const float gn = fminf(fmaxf(fmaf(a.dt, grate, cg), 0.f), 1.f);
        g_new = rupt ? cg : gn;
    }
}

// Band-streaming variant of drainage_vec4_kernel: the same row update on a
// band of a.rows_per_block rows (up to kBandMax) streamed through a window of
// five row buffers, so tall bands cost no extra registers and re-read only
// two halo rows per band.
constexpr int kBandMax = 64;

template <bool REBASE, int GMODE, bool MARANGONI, bool DIAG>
__global__ __launch_bounds__(kBlockMax) void drainage_band_kernel(const StepArgs a) {
    constexpr bool GREAD = GMODE >= 1;
    constexpr bool GWRITE = GMODE == 2;
    const int np = a.np, nt = a.nt;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int nwaves = blockDim.x >> 6;
    // Blocks are dealt to the 8 XCDs round-robin; renumber them so that each
    // XCD walks a contiguous band of rows and the theta-neighbour rows it
    // re-reads stay in its own L2.
    const unsigned nblk = gridDim.x * gridDim.y;
    const unsigned b = blockIdx.y * gridDim.x + blockIdx.x;
    const unsigned per_xcd = (nblk + kXcds - 1) / kXcds;
    unsigned v = (b % kXcds) * per_xcd + b / kXcds;
    if (nblk % kXcds != 0) v = b;                 // uneven split: keep the launch order
    const unsigned bx = v % gridDim.x, by = v / gridDim.x;
    const int c0 = (static_cast<int>(bx) * nwaves + wave) * 256 + 4 * lane;
    const bool valid = c0 < np;
    const int cc = valid ? c0 : np - 4;
    const int jw = cc == 0 ? np - 1 : cc - 1;
    const int je = cc + 4 == np ? 0 : cc + 4;
    const int RB = a.rows_per_block;              // band height, <= kBandMax
    const int r0 = 1 + static_cast<int>(by) * RB;
    const int r1 = min(r0 + RB, nt - 1);          // rows [r0, r1) are updated
    const bool touches_top = r0 == 1;
    const bool touches_bot = r1 == nt - 1;

    __shared__ float2 red_ff[4][kBlockMax / 64];
    __shared__ double red_d[2][kBlockMax / 64];

    float2 ptop = make_float2(0.f, 0.f), pbot = make_float2(0.f, 0.f);
    float gtop = 0.f, gbot = 0.f;
    if (a.h_pole_mean) {
        if (touches_top) ptop = pole_mean(a.hpole_in, a.ntiles, a.np_f);
        if (touches_bot) pbot = pole_mean(a.hpole_in + a.ntiles, a.ntiles, a.np_f);
    }
    if (GREAD && a.g_pole_mean) {
        if (touches_top) { const float2 m = pole_mean(a.gpole_in, a.ntiles, a.np_f); gtop = m.x + m.y; }
        if (touches_bot) { const float2 m = pole_mean(a.gpole_in + a.ntiles, a.ntiles, a.np_f); gbot = m.x + m.y; }
    }

    // The band's per-row coefficients go to LDS before any row is streamed,
    // so that reading them inside the row loop never waits on the vector
    // memory counter that the prefetched rows are queued on.
    __shared__ RowF s_rowf[kBandMax];
    __shared__ RowD s_rowd[DIAG ? kBandMax : 1];
    if (threadIdx.x < RB) {
        const int ri = min(r0 + static_cast<int>(threadIdx.x), nt - 1);
        s_rowf[threadIdx.x] = a.rows[ri];
        if (DIAG) s_rowd[threadIdx.x] = a.rowd[ri];
    }

    // Five row buffers: north, centre, south and two rows in flight.  The
    // loop is unrolled over the five rotations, so no buffer is ever copied
    // while its load is outstanding.
    Row4 P0, P1, P2, P3, P4;
    load_row4<GMODE>(a, r0 - 1, cc, jw, je, P0);
    load_row4<GMODE>(a, r0, cc, jw, je, P1);
    load_row4<GMODE>(a, r0 + 1, cc, jw, je, P2);
    load_row4<GMODE>(a, r0 + 2, cc, jw, je, P3);
    load_row4<GMODE>(a, r0 + 3, cc, jw, je, P4);
    __syncthreads();

    const float G0 = a.inv_dth2;
    float top_h[4] = {0.f, 0.f, 0.f, 0.f}, top_l[4] = {0.f, 0.f, 0.f, 0.f}, top_g[4] = {0.f, 0.f, 0.f, 0.f};
    float bot_h[4] = {0.f, 0.f, 0.f, 0.f}, bot_l[4] = {0.f, 0.f, 0.f, 0.f}, bot_g[4] = {0.f, 0.f, 0.f, 0.f};
    double mass = 0.0, lam = 0.0;

    // Updates row i from (N, C, Sx) = rows (i - 1, i, i + 1).
    auto update_row = [&](int i, Row4& N, const Row4& C, Row4& Sx) {
        // pole rows are substituted in place: row 0 leaves the window after
        // row 1, and row nt - 2 is the last row any band updates
        if (i == 1) pole_row(N, ptop, gtop, a.h_pole_mean, GREAD && a.g_pole_mean);
        if (i == nt - 2) pole_row(Sx, pbot, gbot, a.h_pole_mean, GREAD && a.g_pole_mean);
        const int t = i - r0;
        const RowF rc = s_rowf[t];
        const float G1 = rc.cot * a.inv_2dth;
        const float G2 = rc.inv_s2 * a.inv_dph2;
        const float dtD = a.dt * a.d_r2;
        const float cS = a.dt * a.coef * rc.sin_t;

        // West / east neighbours of the lane's first and last column.
        const float hw = C.hw, lw = C.lw, he = C.he, le = C.le;
        const float gw = GREAD ? C.gw : 0.f, ge = GREAD ? C.ge : 0.f;

        const float ch[4] = {C.h.x, C.h.y, C.h.z, C.h.w};
        const float cl[4] = {C.l.x, C.l.y, C.l.z, C.l.w};
        const float nh[4] = {N.h.x, N.h.y, N.h.z, N.h.w};
        const float nl[4] = {N.l.x, N.l.y, N.l.z, N.l.w};
        const float sh[4] = {Sx.h.x, Sx.h.y, Sx.h.z, Sx.h.w};
        const float sl[4] = {Sx.l.x, Sx.l.y, Sx.l.z, Sx.l.w};
        const float wh[4] = {hw, C.h.x, C.h.y, C.h.z};
        const float wl[4] = {lw, C.l.x, C.l.y, C.l.z};
        const float eh[4] = {C.h.y, C.h.z, C.h.w, he};
        const float el[4] = {C.l.y, C.l.z, C.l.w, le};
        float cg[4] = {0.f, 0.f, 0.f, 0.f}, ng[4] = {0.f, 0.f, 0.f, 0.f}, sg[4] = {0.f, 0.f, 0.f, 0.f};
        float wg[4] = {0.f, 0.f, 0.f, 0.f}, eg[4] = {0.f, 0.f, 0.f, 0.f};
        if (GREAD) {
            cg[0] = C.g.x; cg[1] = C.g.y; cg[2] = C.g.z; cg[3] = C.g.w;
            ng[0] = N.g.x; ng[1] = N.g.y; ng[2] = N.g.z; ng[3] = N.g.w;
            sg[0] = Sx.g.x; sg[1] = Sx.g.y; sg[2] = Sx.g.z; sg[3] = Sx.g.w;
            wg[0] = gw; wg[1] = C.g.x; wg[2] = C.g.y; wg[3] = C.g.z;
            eg[0] = C.g.y; eg[1] = C.g.z; eg[2] = C.g.w; eg[3] = ge;
        }

        float lo_new[4], g_new[4];
#pragma unroll
        for (int v = 0; v < 4; ++v)
            cell_step<GMODE, MARANGONI>(a, rc, G0, G1, G2, dtD, cS, ch[v], cl[v], nh[v], nl[v], sh[v], sl[v], wh[v],
                                        wl[v], eh[v], el[v], cg[v], ng[v], sg[v], wg[v], eg[v], lo_new[v], g_new[v]);

        // Keep the ring rows' new values (before the fronts act, as
        // update_poles runs before update_marginal_regeneration) for the
        // pole means; they are reduced once after the row loop.
        if (i == 1) {
#pragma unroll
            for (int v = 0; v < 4; ++v) { top_h[v] = ch[v]; top_l[v] = lo_new[v]; top_g[v] = g_new[v]; }
        }
        if (i == nt - 2) {
#pragma unroll
            for (int v = 0; v < 4; ++v) { bot_h[v] = ch[v]; bot_l[v] = lo_new[v]; bot_g[v] = g_new[v]; }
        }

        if (a.nfronts > 0) {
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                if (in_fronts(a, i, cc + v)) lo_new[v] = front_thin(a, ch[v], lo_new[v]);
            }
        }

        if (valid) {
            const size_t off = static_cast<size_t>(i) * np + cc;
            if (REBASE) {
                float4 hh, ll;
                two_sum(ch[0], lo_new[0], hh.x, ll.x);
                two_sum(ch[1], lo_new[1], hh.y, ll.y);
                two_sum(ch[2], lo_new[2], hh.z, ll.z);
                two_sum(ch[3], lo_new[3], hh.w, ll.w);
                *reinterpret_cast<float4*>(a.hi_out + off) = hh;
                *reinterpret_cast<float4*>(a.lo_out + off) = ll;
            } else {
                *reinterpret_cast<float4*>(a.lo_out + off) =
                    make_float4(lo_new[0], lo_new[1], lo_new[2], lo_new[3]);
            }
            if (GWRITE)
                *reinterpret_cast<float4*>(a.g_out + off) = make_float4(g_new[0], g_new[1], g_new[2], g_new[3]);
            if (DIAG) diag_row(s_rowd[DIAG ? t : 0], a.coef3, ch, lo_new, mass, lam);
        }
    };
#define SF_BAND_ROW(Nb, Cb, Sb)                                                  \
    {                                                                            \
        if (i >= r1) break;                                                      \
        update_row(i, Nb, Cb, Sb);                                               \
        if (i + 4 <= r1) load_row4<GMODE>(a, i + 4, cc, jw, je, Nb);             \
        ++i;                                                                     \
    }
    for (int i = r0;;) {
        SF_BAND_ROW(P0, P1, P2)
        SF_BAND_ROW(P1, P2, P3)
        SF_BAND_ROW(P2, P3, P4)
        SF_BAND_ROW(P3, P4, P0)
        SF_BAND_ROW(P4, P0, P1)
    }
#undef SF_BAND_ROW

    if (touches_top || touches_bot) {
        float2 sum_top = make_float2(0.f, 0.f), sum_bot = make_float2(0.f, 0.f);
        float2 gsum_top = make_float2(0.f, 0.f), gsum_bot = make_float2(0.f, 0.f);
        if (valid) {
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                float s_, e_;
                two_sum(top_h[v], top_l[v], s_, e_);
                sum_top = ff_add(sum_top, make_float2(s_, e_));
                two_sum(bot_h[v], bot_l[v], s_, e_);
                sum_bot = ff_add(sum_bot, make_float2(s_, e_));
                gsum_top = ff_add(gsum_top, make_float2(top_g[v], 0.f));
                gsum_bot = ff_add(gsum_bot, make_float2(bot_g[v], 0.f));
            }
        }
        sum_top = wave_ff_sum(sum_top);
        sum_bot = wave_ff_sum(sum_bot);
        if (GWRITE) {
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
                if (touches_top) a.hpole_out[bx] = t[0];
                if (touches_bot) a.hpole_out[a.ntiles + bx] = t[1];
            }
            if (GWRITE && a.g_pole_sum) {
                if (touches_top) a.gpole_out[bx] = t[2];
                if (touches_bot) a.gpole_out[a.ntiles + bx] = t[3];
            }
        }
    }

    if (DIAG) diag_epilogue(a, mass, lam, lane, wave, nwaves, red_d);
}

// ---- two steps per launch (temporal blocking) -------------------------------
//
// drainage_band2_kernel advances the state by two steps in one pass over the
// band: each lane streams the step-s rows of its four columns through a
// register window (with the west and east neighbour columns and one column
// beyond, for the wave's edge cells), forms the step-(s+1) rows and keeps them
// in a three-row ring in the wave's LDS, and forms the step-(s+2) rows from
// that ring.  A band thus reads h, its low word and Γ once and writes them once
// per two steps.  Every cell update is cell_step, the arithmetic of the
// one-step band kernel, and the pole means between the two steps are formed
// from the same partial sums in the same order — a block whose band touches a
// pole recomputes that pole's ring row of step s+1 over every column tile —
// so a launch equals two launches of drainage_band_kernel bit for bit.
// Marginal-regeneration fronts act on the step-(s+1) rows in the ring (halo
// rows and the edge columns included) and on the step-(s+2) rows, each step
// with its own front boxes.  With DIAG the launch leaves per-block FP64
// partials for both steps.  With a.lo_mid set it also writes the state
// between the two steps (its low word; h's high word is unchanged, as a
// rebase only follows the second step) and that state's ring partial sums,
// for consumers of every step such as the spectral pass.
// Scope: no Marangoni term, Γ poles fixed, n_phi a multiple of 256 (whole
// waves), a rebase only after the second step.
constexpr int kRing2W = 264;     // LDS row of a wave: edge at 3, columns 4 .. 259, edge at 260
constexpr int kMaxTiles2 = 64;   // column tiles (n_phi <= 65536 with four-wave blocks)

// Element at a 32-bit byte offset from a uniform base: the global access
// then takes the base in scalar registers and one 32-bit offset register.
// The two-step kernel runs on grids below 2^30 cells, whose byte offsets
// fit 32 bits.
template <class T>
__device__ __forceinline__ const T* at_bytes(const float* base, uint32_t off) {
    return reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + off);
}
template <class T>
__device__ __forceinline__ T* at_bytes(float* base, uint32_t off) {
    return reinterpret_cast<T*>(reinterpret_cast<char*>(base) + off);
}

struct Row5 {
    float4 h, l, g;              // the lane's four columns cc .. cc + 3
    float hx, lx, gx;            // lane 0: column cc - 1; other lanes: cc + 4
};

struct Far5 {                    // lane 0: column cc - 2; other lanes: cc + 5
    float h, l, g;
};

// The far column of row r, which only the wave's edge cells use (as the
// centre row of a first step): loaded one row ahead of its use, not with the
// row window.
template <int GMODE>
__device__ inline void load_far5(const StepArgs& a, int r, int jf, Far5& F) {
    const uint32_t off = static_cast<uint32_t>(max(0, min(r, a.nt - 1)) * a.np + jf) * 4u;
    F.h = *at_bytes<float>(a.hi_in, off);
    F.l = *at_bytes<float>(a.lo_in, off);
    F.g = GMODE >= 1 ? *at_bytes<float>(a.g_in, off) : 0.f;
}

// Loads step-s row r (clamped into the grid).  With pm set, the pole rows
// (0 for r <= 0, nt - 1 for r >= nt - 1) hold the pole means P0 of h and its
// low word instead of the stored values, which is what every use of them
// needs: they are the theta neighbours of rows 1 and nt - 2, and rows 0 and
// nt - 1 themselves are replaced when stepped.
template <int GMODE>
__device__ inline void load_row5(const StepArgs& a, int r, int cc, int jx, bool pm, float2 ptop, float2 pbot,
                                 Row5& R) {
    const int rr = max(0, min(r, a.nt - 1));
    const uint32_t oc = static_cast<uint32_t>(rr * a.np + cc) * 4u;
    const uint32_t ox = static_cast<uint32_t>(rr * a.np + jx) * 4u;
    const bool pole = pm && (rr == 0 || rr == a.nt - 1);
    if (pole) {
        const float2 P = rr == 0 ? ptop : pbot;
        R.h = make_float4(P.x, P.x, P.x, P.x);
        R.l = make_float4(P.y, P.y, P.y, P.y);
        R.hx = P.x;
        R.lx = P.y;
    } else {
        R.h = *at_bytes<float4>(a.hi_in, oc);
        R.l = *at_bytes<float4>(a.lo_in, oc);
        R.hx = *at_bytes<float>(a.hi_in, ox);
        R.lx = *at_bytes<float>(a.lo_in, ox);
    }
    if (GMODE >= 1) {
        R.g = *at_bytes<float4>(a.g_in, oc);
        R.gx = *at_bytes<float>(a.g_in, ox);
    } else {
        R.g = make_float4(0.f, 0.f, 0.f, 0.f);
        R.gx = 0.f;
    }
}

// The west neighbour of the lane's first column: lane l - 1's last column,
// or, in lane 0, x (the column before the wave's first).  DPP wave_shr:1.
__device__ __forceinline__ float from_west(float last, float x) {
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(x), __float_as_int(last), 0x138, 0xf, 0xf, false));
}

// The east neighbour of the lane's last column: lane l + 1's first column,
// or, in lane 63, x (the column after the wave's last).  DPP wave_shl:1.
__device__ __forceinline__ float from_east(float first, float x) {
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(x), __float_as_int(first), 0x130, 0xf, 0xf, false));
}

// m ? a : b with m all ones or all zeros, as a bit blend of register values.
__device__ __forceinline__ float blend(unsigned m, float a, float b) {
    return __uint_as_float((__float_as_uint(a) & m) | (__float_as_uint(b) & ~m));
}

// Ring-1 (or ring nt-2) partial sums of step s+1 for every column tile, as
// drainage_band_kernel forms them (same lanes, waves and order), and the pole
// mean they give.  ring = 1 or nt - 2; Pin: the pole of step s (applied when
// a.h_pole_mean).  All threads of the block call it; returns the mean.
template <int GMODE>
__device__ float2 pole_after_first_step(const StepArgs& a, int ring, float2 Pin, float2* part, float2* red) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwaves = blockDim.x >> 6;
    const int np = a.np;
    const RowF rc = a.rows[ring];
    const float G0 = a.inv_dth2, G1 = rc.cot * a.inv_2dth, G2 = rc.inv_s2 * a.inv_dph2;
    const float dtD = a.dt * a.d_r2, cS = a.dt * a.coef * rc.sin_t;
    const bool top = ring == 1;
    for (int t = 0; t < a.ntiles; ++t) {
        const int c0 = (t * nwaves + wave) * 256 + 4 * lane;
        const bool valid = c0 < np;
        const int cc = valid ? c0 : np - 4;
        const int jw = cc == 0 ? np - 1 : cc - 1;
        const int je = cc + 4 == np ? 0 : cc + 4;
        Row4 N, C, S;
        load_row4<GMODE>(a, ring - 1, cc, jw, je, N);
        load_row4<GMODE>(a, ring, cc, jw, je, C);
        load_row4<GMODE>(a, ring + 1, cc, jw, je, S);
        if (top)
            pole_row(N, Pin, 0.f, a.h_pole_mean, false);
        else
            pole_row(S, Pin, 0.f, a.h_pole_mean, false);
        const float ch[4] = {C.h.x, C.h.y, C.h.z, C.h.w}, cl[4] = {C.l.x, C.l.y, C.l.z, C.l.w};
        const float nh[4] = {N.h.x, N.h.y, N.h.z, N.h.w}, nl[4] = {N.l.x, N.l.y, N.l.z, N.l.w};
        const float sh[4] = {S.h.x, S.h.y, S.h.z, S.h.w}, sl[4] = {S.l.x, S.l.y, S.l.z, S.l.w};
        const float wh[4] = {C.hw, C.h.x, C.h.y, C.h.z}, wl[4] = {C.lw, C.l.x, C.l.y, C.l.z};
        const float eh[4] = {C.h.y, C.h.z, C.h.w, C.he}, el[4] = {C.l.y, C.l.z, C.l.w, C.le};
        float cg[4] = {0.f, 0.f, 0.f, 0.f}, ng[4] = {0.f, 0.f, 0.f, 0.f}, sg[4] = {0.f, 0.f, 0.f, 0.f};
        float wg[4] = {0.f, 0.f, 0.f, 0.f}, eg[4] = {0.f, 0.f, 0.f, 0.f};
        if (GMODE >= 1) {
            cg[0] = C.g.x; cg[1] = C.g.y; cg[2] = C.g.z; cg[3] = C.g.w;
            ng[0] = N.g.x; ng[1] = N.g.y; ng[2] = N.g.z; ng[3] = N.g.w;
            sg[0] = S.g.x; sg[1] = S.g.y; sg[2] = S.g.z; sg[3] = S.g.w;
            wg[0] = C.gw; wg[1] = C.g.x; wg[2] = C.g.y; wg[3] = C.g.z;
            eg[0] = C.g.y; eg[1] = C.g.z; eg[2] = C.g.w; eg[3] = C.ge;
        }
        float2 sum = make_float2(0.f, 0.f);
        if (valid) {
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                float lo_new, g_new;
                cell_step<GMODE, false>(a, rc, G0, G1, G2, dtD, cS, ch[v], cl[v], nh[v], nl[v], sh[v], sl[v], wh[v],
                                        wl[v], eh[v], el[v], cg[v], ng[v], sg[v], wg[v], eg[v], lo_new, g_new);
                float s_, e_;
                two_sum(ch[v], lo_new, s_, e_);
                sum = ff_add(sum, make_float2(s_, e_));
            }
        }
        sum = wave_ff_sum(sum);
        if (lane == 0) red[wave] = sum;
        __syncthreads();
        if (threadIdx.x == 0) {
            float2 tt = red[0];
            for (int w = 1; w < nwaves; ++w) tt = ff_add(tt, red[w]);
            part[t] = tt;
        }
        __syncthreads();
    }
    return pole_mean(part, a.ntiles, a.np_f);
}

// Three waves per SIMD (168 VGPRs with Γ transport): one step-s row in
// flight per wave measured faster than three rows at two waves per SIMD.
// OPT: kOptDiag (fused FP64 diagnostics), kOptFronts (marginal-regeneration
// fronts in either step), kOptMid (write the state between the steps); each
// its own instantiation, so that launches without them keep their registers.
constexpr int kOptDiag = 1, kOptFronts = 2, kOptMid = 4;

template <bool REBASE, int GMODE, int OPT>
__global__ __launch_bounds__(kBlockMax) __attribute__((amdgpu_waves_per_eu(3, 3)))
void drainage_band2_kernel(const StepArgs a) {
    constexpr bool DIAG = (OPT & kOptDiag) != 0;
    constexpr bool FRONTS = (OPT & kOptFronts) != 0;
    constexpr bool MID = (OPT & kOptMid) != 0;
    constexpr bool GREAD = GMODE >= 1;
    constexpr bool GWRITE = GMODE == 2;
    const int np = a.np, nt = a.nt;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int nwaves = blockDim.x >> 6;
    // The blocks of the first and last band (the pole rings, the longest
    // blocks) go first; the rest keep the one-step kernel's XCD-contiguous
    // order.
    const unsigned nblk = gridDim.x * gridDim.y;
    const unsigned b = blockIdx.y * gridDim.x + blockIdx.x;
    unsigned bx, by;
    if (b < 2 * gridDim.x && gridDim.y > 1) {
        bx = b % gridDim.x;
        by = b < gridDim.x ? 0 : gridDim.y - 1;
    } else {
        const unsigned rest = gridDim.y > 1 ? b - 2 * gridDim.x : b;      // index among the inner bands
        const unsigned ninner = gridDim.y > 1 ? nblk - 2 * gridDim.x : nblk;
        const unsigned per_xcd = (ninner + kXcds - 1) / kXcds;
        unsigned v = (rest % kXcds) * per_xcd + rest / kXcds;
        if (ninner % kXcds != 0) v = rest;
        bx = v % gridDim.x;
        by = v / gridDim.x + (gridDim.y > 1 ? 1 : 0);
    }
    const int c0 = (static_cast<int>(bx) * nwaves + wave) * 256 + 4 * lane;
    const bool valid = c0 < np;                   // whole waves: n_phi % 256 == 0
    const int cc = valid ? c0 : np - 4;
    const int jw = cc == 0 ? np - 1 : cc - 1;
    const int je = cc + 4 == np ? 0 : cc + 4;
    const int jx = lane == 0 ? jw : je;
    const int jf = lane == 0 ? (cc < 2 ? cc - 2 + np : cc - 2) : (cc + 5 >= np ? cc + 5 - np : cc + 5);
    const int RB = a.rows_per_block;
    const int r0 = 1 + static_cast<int>(by) * RB;
    const int r1 = min(r0 + RB, nt - 1);          // rows [r0, r1) of step s+2 are written
    const bool touches_top = r0 == 1;
    const bool touches_bot = r1 == nt - 1;

    __shared__ float2 red_ff[4][kBlockMax / 64];
    __shared__ float2 s_part[kMaxTiles2];
    __shared__ RowF s_rowf[kBandMax + 2];
    __shared__ RowD s_rowd[DIAG ? kBandMax : 1];
    __shared__ double red_d[2][kBlockMax / 64];
    // Each lane's FP64 film volume and CFL lambda of steps s+1 and s+2,
    // accumulated in LDS (the registers are taken by the row window).
    __shared__ double s_diag[DIAG ? 4 : 1][DIAG ? kBlockMax : 1];
    if constexpr (DIAG) {
        for (int q = 0; q < 4; ++q) s_diag[q][threadIdx.x] = 0.0;
    }
    __shared__ __align__(16) float ring[kBlockMax / 64][3][3][kRing2W];   // [wave][slot][h, l, g][column]

    // Row coefficients of rows r0 - 1 .. r1 (step s+1 rows r0 - 1 .. r1 - 1 are
    // formed with those of their own rows; step s+2 rows r0 .. r1 - 1), and
    // the FP64 terms of the band's own rows r0 .. r1 - 1.
    for (int t = threadIdx.x; t < RB + 2; t += blockDim.x) {
        const int ri = min(max(r0 - 1 + t, 0), nt - 1);
        s_rowf[t] = a.rows[ri];
        if (DIAG && t < RB) s_rowd[t] = a.rowd[min(r0 + t, nt - 1)];
    }
    // Poles of step s (P0) and of step s+1 (P1).
    float2 p0top = make_float2(0.f, 0.f), p0bot = make_float2(0.f, 0.f);
    if (a.h_pole_mean) {
        if (touches_top || r0 <= 2) p0top = pole_mean(a.hpole_in, a.ntiles, a.np_f);
        if (touches_bot || r1 >= nt - 2) p0bot = pole_mean(a.hpole_in + a.ntiles, a.ntiles, a.np_f);
    }
    float2 p1top = make_float2(0.f, 0.f), p1bot = make_float2(0.f, 0.f);
    const bool need_top = a.h_pole_sum && touches_top;           // step s+1 row 0 is used
    const bool need_bot = a.h_pole_sum && touches_bot;           // step s+1 row nt - 1 is used
    // The partials of step s+1 go out with the state between the steps, and
    // for that step's pole-row film volume (written by one block per pole).
    const bool mid_poles = (MID || DIAG) && bx == 0;
    if (need_top) {
        p1top = pole_after_first_step<GMODE>(a, 1, p0top, s_part, red_ff[0]);
        if (mid_poles && threadIdx.x < a.ntiles) a.hpole_mid[threadIdx.x] = s_part[threadIdx.x];
    }
    if (need_bot) {
        p1bot = pole_after_first_step<GMODE>(a, nt - 2, p0bot, s_part, red_ff[1]);
        if (mid_poles && threadIdx.x < a.ntiles) a.hpole_mid[a.ntiles + threadIdx.x] = s_part[threadIdx.x];
    }
    // Every lane holds the same pole means: keep them in scalar registers.
    auto uniform = [](float2 v) {
        return make_float2(__int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v.x))),
                           __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v.y))));
    };
    p0top = uniform(p0top);
    p0bot = uniform(p0bot);
    p1top = uniform(p1top);
    p1bot = uniform(p1bot);
    __syncthreads();

    float* rw = &ring[wave][0][0][0];
    auto slot = [&](int k, int arr) { return rw + (((k % 3) + 3) % 3) * 3 * kRing2W + arr * kRing2W; };
    const int lc = 4 + 4 * lane;                  // the lane's first column in an LDS row

    // Step s+1 row k from step-s rows (N, C, S) into the ring.
    auto first_step = [&](int k, const Row5& N, const Row5& C, const Row5& S, const Far5& F) {
        const RowF rc = s_rowf[k - (r0 - 1)];
        const float G0 = a.inv_dth2, G1 = rc.cot * a.inv_2dth, G2 = rc.inv_s2 * a.inv_dph2;
        const float dtD = a.dt * a.d_r2, cS = a.dt * a.coef * rc.sin_t;
        const float ch[4] = {C.h.x, C.h.y, C.h.z, C.h.w}, cl[4] = {C.l.x, C.l.y, C.l.z, C.l.w};
        const float nh[4] = {N.h.x, N.h.y, N.h.z, N.h.w}, nl[4] = {N.l.x, N.l.y, N.l.z, N.l.w};
        const float sh[4] = {S.h.x, S.h.y, S.h.z, S.h.w}, sl[4] = {S.l.x, S.l.y, S.l.z, S.l.w};
        const float hw = from_west(C.h.w, C.hx), lw = from_west(C.l.w, C.lx);
        const float he = from_east(C.h.x, C.hx), le = from_east(C.l.x, C.lx);
        const float gw = GREAD ? from_west(C.g.w, C.gx) : 0.f, ge = GREAD ? from_east(C.g.x, C.gx) : 0.f;
        const float wh[4] = {hw, C.h.x, C.h.y, C.h.z}, wl[4] = {lw, C.l.x, C.l.y, C.l.z};
        const float eh[4] = {C.h.y, C.h.z, C.h.w, he}, el[4] = {C.l.y, C.l.z, C.l.w, le};
        const float cg[4] = {C.g.x, C.g.y, C.g.z, C.g.w}, ng[4] = {N.g.x, N.g.y, N.g.z, N.g.w};
        const float sg[4] = {S.g.x, S.g.y, S.g.z, S.g.w};
        const float wg[4] = {gw, C.g.x, C.g.y, C.g.z}, eg[4] = {C.g.y, C.g.z, C.g.w, ge};
        float lo1[4], g1[4];
#pragma unroll
        for (int v = 0; v < 4; ++v)
            cell_step<GMODE, false>(a, rc, G0, G1, G2, dtD, cS, ch[v], cl[v], nh[v], nl[v], sh[v], sl[v], wh[v], wl[v],
                                    eh[v], el[v], cg[v], ng[v], sg[v], wg[v], eg[v], lo1[v], g1[v]);
        // the wave's edge cell: lane 0 its west column cc - 1, lane 63 its
        // east column cc + 4 (the other lanes form the east one and drop it)
        const unsigned wm = lane == 0 ? ~0u : 0u;
        const float xch = C.hx, xcl = C.lx, xcg = C.gx;
        const float xnh = N.hx, xnl = N.lx, xng = N.gx;
        const float xsh = S.hx, xsl = S.lx, xsg = S.gx;
        const float xwh = blend(wm, F.h, C.h.w), xwl = blend(wm, F.l, C.l.w), xwg = blend(wm, F.g, C.g.w);
        const float xeh = blend(wm, C.h.x, F.h), xel = blend(wm, C.l.x, F.l), xeg = blend(wm, C.g.x, F.g);
        float xlo, xg;
        cell_step<GMODE, false>(a, rc, G0, G1, G2, dtD, cS, xch, xcl, xnh, xnl, xsh, xsl, xwh, xwl, xeh, xel, xcg, xng,
                                xsg, xwg, xeg, xlo, xg);
        // Fronts of step s+1 (never on a pole row), the edge column with its
        // own index: lane 0 column cc - 1, lane 63 column cc + 4.
        if (FRONTS && a.nfronts > 0) {
#pragma unroll
            for (int v = 0; v < 4; ++v)
                if (in_front_boxes(a.fr, a.nfronts, np, k, cc + v)) lo1[v] = front_thin(a, ch[v], lo1[v]);
            if (in_front_boxes(a.fr, a.nfronts, np, k, jx)) xlo = front_thin(a, xch, xlo);
        }
        // A pole row of step s+1 is not stepped: h and its low word are the
        // pole mean P1 (or, with copied poles, the stored row) and Γ is
        // copied through.  Its cells were formed above from clamped rows and
        // are replaced here by value, so the row loop has no branches.
        const bool pole = k == 0 || k == nt - 1;
        const bool use_p = pole && a.h_pole_sum;
        const float2 P = k == 0 ? p1top : p1bot;
        float oh[4], ol[4], og[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            oh[v] = use_p ? P.x : ch[v];
            ol[v] = use_p ? P.y : (pole ? cl[v] : lo1[v]);
            og[v] = pole ? cg[v] : g1[v];
        }
        const float xoh = use_p ? P.x : xch;
        const float xol = use_p ? P.y : (pole ? xcl : xlo);
        const float xog = pole ? xcg : xg;
        float* rh = slot(k, 0);
        float* rl = slot(k, 1);
        float* rg = slot(k, 2);
        *reinterpret_cast<float4*>(rh + lc) = make_float4(oh[0], oh[1], oh[2], oh[3]);
        *reinterpret_cast<float4*>(rl + lc) = make_float4(ol[0], ol[1], ol[2], ol[3]);
        if (GREAD) *reinterpret_cast<float4*>(rg + lc) = make_float4(og[0], og[1], og[2], og[3]);
        if (lane == 0) { rh[3] = xoh; rl[3] = xol; if (GREAD) rg[3] = xog; }
        if (lane == 63) { rh[260] = xoh; rl[260] = xol; if (GREAD) rg[260] = xog; }
        const bool own = k >= r0 && k < r1;
        if (MID && valid && (own || pole))
            *at_bytes<float4>(a.lo_mid, static_cast<uint32_t>(k * np + cc) * 4u) =
                make_float4(ol[0], ol[1], ol[2], ol[3]);
        if constexpr (DIAG) {
            if (valid && own) diag_row(s_rowd[k - r0], a.coef3, oh, ol, s_diag[0][threadIdx.x], s_diag[1][threadIdx.x]);
        }
    };

    // The lane's float-float sums of its ring-row cells of step s+2 (rows 1
    // and nt - 2), in the order of the one-step kernel.
    float2 sum_top = make_float2(0.f, 0.f), sum_bot = make_float2(0.f, 0.f);
    auto lane_ring_sum = [&](const float (&h)[4], const float (&l)[4]) {
        float2 t = make_float2(0.f, 0.f);
        if (valid) {
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                float s_, e_;
                two_sum(h[v], l[v], s_, e_);
                t = ff_add(t, make_float2(s_, e_));
            }
        }
        return t;
    };

    // Step s+2 row i from ring rows i - 1, i, i + 1.
    auto second_step = [&](int i) {
        const float* nh_ = slot(i - 1, 0); const float* nl_ = slot(i - 1, 1); const float* ng_ = slot(i - 1, 2);
        const float* ch_ = slot(i, 0);     const float* cl_ = slot(i, 1);     const float* cg_ = slot(i, 2);
        const float* sh_ = slot(i + 1, 0); const float* sl_ = slot(i + 1, 1); const float* sg_ = slot(i + 1, 2);
        const float4 C4 = *reinterpret_cast<const float4*>(ch_ + lc), L4 = *reinterpret_cast<const float4*>(cl_ + lc);
        const float4 N4 = *reinterpret_cast<const float4*>(nh_ + lc), NL4 = *reinterpret_cast<const float4*>(nl_ + lc);
        const float4 S4 = *reinterpret_cast<const float4*>(sh_ + lc), SL4 = *reinterpret_cast<const float4*>(sl_ + lc);
        const float ch[4] = {C4.x, C4.y, C4.z, C4.w}, cl[4] = {L4.x, L4.y, L4.z, L4.w};
        const float nh[4] = {N4.x, N4.y, N4.z, N4.w}, nl[4] = {NL4.x, NL4.y, NL4.z, NL4.w};
        const float sh[4] = {S4.x, S4.y, S4.z, S4.w}, sl[4] = {SL4.x, SL4.y, SL4.z, SL4.w};
        const float wh[4] = {ch_[lc - 1], C4.x, C4.y, C4.z}, wl[4] = {cl_[lc - 1], L4.x, L4.y, L4.z};
        const float eh[4] = {C4.y, C4.z, C4.w, ch_[lc + 4]}, el[4] = {L4.y, L4.z, L4.w, cl_[lc + 4]};
        float cg[4] = {0.f, 0.f, 0.f, 0.f}, ng[4] = {0.f, 0.f, 0.f, 0.f}, sg[4] = {0.f, 0.f, 0.f, 0.f};
        float wg[4] = {0.f, 0.f, 0.f, 0.f}, eg[4] = {0.f, 0.f, 0.f, 0.f};
        if (GREAD) {
            const float4 G4 = *reinterpret_cast<const float4*>(cg_ + lc);
            const float4 NG4 = *reinterpret_cast<const float4*>(ng_ + lc);
            const float4 SG4 = *reinterpret_cast<const float4*>(sg_ + lc);
            cg[0] = G4.x; cg[1] = G4.y; cg[2] = G4.z; cg[3] = G4.w;
            ng[0] = NG4.x; ng[1] = NG4.y; ng[2] = NG4.z; ng[3] = NG4.w;
            sg[0] = SG4.x; sg[1] = SG4.y; sg[2] = SG4.z; sg[3] = SG4.w;
            wg[0] = cg_[lc - 1]; wg[1] = G4.x; wg[2] = G4.y; wg[3] = G4.z;
            eg[0] = G4.y; eg[1] = G4.z; eg[2] = G4.w; eg[3] = cg_[lc + 4];
        }
        const RowF rc = s_rowf[i - (r0 - 1)];
        const float G0 = a.inv_dth2, G1 = rc.cot * a.inv_2dth, G2 = rc.inv_s2 * a.inv_dph2;
        const float dtD = a.dt * a.d_r2, cS = a.dt * a.coef * rc.sin_t;
        float lo2[4], g2[4];
#pragma unroll
        for (int v = 0; v < 4; ++v)
            cell_step<GMODE, false>(a, rc, G0, G1, G2, dtD, cS, ch[v], cl[v], nh[v], nl[v], sh[v], sl[v], wh[v], wl[v],
                                    eh[v], el[v], cg[v], ng[v], sg[v], wg[v], eg[v], lo2[v], g2[v]);
        if (i == 1) sum_top = lane_ring_sum(ch, lo2);
        if (i == nt - 2) sum_bot = lane_ring_sum(ch, lo2);
        if (FRONTS && a.nfronts2 > 0) {
#pragma unroll
            for (int v = 0; v < 4; ++v)
                if (in_front_boxes(a.fr2, a.nfronts2, np, i, cc + v)) lo2[v] = front_thin(a, ch[v], lo2[v]);
        }
        if constexpr (DIAG) {
            if (valid) diag_row(s_rowd[i - r0], a.coef3, ch, lo2, s_diag[2][threadIdx.x], s_diag[3][threadIdx.x]);
        }
        if (valid) {
            const uint32_t off = static_cast<uint32_t>(i * np + cc) * 4u;
            if (REBASE) {
                float4 hh, ll;
                two_sum(ch[0], lo2[0], hh.x, ll.x);
                two_sum(ch[1], lo2[1], hh.y, ll.y);
                two_sum(ch[2], lo2[2], hh.z, ll.z);
                two_sum(ch[3], lo2[3], hh.w, ll.w);
                *at_bytes<float4>(a.hi_out, off) = hh;
                *at_bytes<float4>(a.lo_out, off) = ll;
            } else {
                *at_bytes<float4>(a.lo_out, off) = make_float4(lo2[0], lo2[1], lo2[2], lo2[3]);
            }
            if (GWRITE) *at_bytes<float4>(a.g_out, off) = make_float4(g2[0], g2[1], g2[2], g2[3]);
        }
    };

    // Step-s rows k - 1, k, k + 1 form step s+1 row k (k = r0 - 1 .. r1);
    // after row k is in the ring, step s+2 row k - 1 follows.  Four buffers,
    // one row in flight, unrolled over the rotations (no buffer is copied
    // while its load is outstanding).
    // Step-s rows k - 1, k, k + 1 form step s+1 row k (k = r0 - 1 .. r1);
    // after row k is in the ring, step s+2 row k - 1 follows.  Four row
    // buffers (one row in flight) unrolled over their rotations, so no
    // buffer is copied while its load is outstanding; the far columns of the
    // edge cells come one row ahead in two more.
    Row5 B0, B1, B2, B3;
    Far5 F0, F1;
    const int ks = r0 - 1;
    const bool pm = a.h_pole_mean != 0;
    load_row5<GMODE>(a, ks - 1, cc, jx, pm, p0top, p0bot, B0);
    load_row5<GMODE>(a, ks, cc, jx, pm, p0top, p0bot, B1);
    load_row5<GMODE>(a, ks + 1, cc, jx, pm, p0top, p0bot, B2);
    load_row5<GMODE>(a, ks + 2, cc, jx, pm, p0top, p0bot, B3);
    load_far5<GMODE>(a, ks, jf, F0);
    auto row_k = [&](int k, const Row5& N, const Row5& C, const Row5& S, const Far5& F) {
        first_step(k, N, C, S, F);
        __builtin_amdgcn_wave_barrier();
        if (k - 1 >= r0) second_step(k - 1);
        __builtin_amdgcn_wave_barrier();
    };
#define SF_BAND2_ROW(Nb, Cb, Sb, Fc, Fn)                                         \
    {                                                                            \
        if (k > r1) break;                                                       \
        if (k + 1 <= r1) load_far5<GMODE>(a, k + 1, jf, Fn);                     \
        row_k(k, Nb, Cb, Sb, Fc);                                                \
        if (k + 3 <= r1 + 1) load_row5<GMODE>(a, k + 3, cc, jx, pm, p0top, p0bot, Nb); \
        ++k;                                                                     \
    }
    for (int k = ks;;) {
        SF_BAND2_ROW(B0, B1, B2, F0, F1)
        SF_BAND2_ROW(B1, B2, B3, F1, F0)
        SF_BAND2_ROW(B2, B3, B0, F0, F1)
        SF_BAND2_ROW(B3, B0, B1, F1, F0)
    }
#undef SF_BAND2_ROW

    // Ring partial sums of step s+2 for the next launch's pole means.
    if (touches_top || touches_bot) {
        sum_top = wave_ff_sum(sum_top);
        sum_bot = wave_ff_sum(sum_bot);
        __syncthreads();
        if (lane == 0) {
            red_ff[0][wave] = sum_top;
            red_ff[1][wave] = sum_bot;
        }
        __syncthreads();
        if (threadIdx.x == 0 && a.h_pole_sum) {
            float2 t0 = red_ff[0][0], t1 = red_ff[1][0];
            for (int w = 1; w < nwaves; ++w) {
                t0 = ff_add(t0, red_ff[0][w]);
                t1 = ff_add(t1, red_ff[1][w]);
            }
            if (touches_top) a.hpole_out[bx] = t0;
            if (touches_bot) a.hpole_out[a.ntiles + bx] = t1;
        }
    }
    if constexpr (DIAG) {
        // Block partials of the four accumulators (volume and lambda of both
        // steps), one wave each: a lane folds the entries lane, lane + 64, ...
        // in order, then the wave reduces (the reductions of the four waves
        // run side by side).
        __syncthreads();
        const unsigned slot = b;      // partials are indexed by hardware block
        for (int q = wave; q < 4; q += nwaves) {
            const bool is_max = (q & 1) != 0;
            double v = 0.0;
            for (int t = lane; t < static_cast<int>(blockDim.x); t += 64)
                v = is_max ? fmax(v, s_diag[q][t]) : v + s_diag[q][t];
            v = is_max ? wave_dmax(v) : wave_dsum(v);
            if (lane == 0) {
                double2* part = q < 2 ? a.diag_part : a.diag_part2;
                if (is_max)
                    part[slot].y = v;
                else
                    part[slot].x = v;
            }
        }
        // The previous launch's steps are reduced by the first inner-band
        // blocks, one step each: the pole-band blocks, dispatched first, are
        // the longest, and a reduction appended to one of them would end
        // the launch late.
        const unsigned first_inner = gridDim.y > 2 ? 2 * gridDim.x : 0;
        for (int q = 0; q < a.nprev; ++q) {
            if (b != (first_inner + static_cast<unsigned>(q)) % nblk) continue;
            const DiagPending e = q == 0 ? a.prev[0] : a.prev[1];
            __syncthreads();
            diag_reduce(e.part, e.nblocks, e.hpole, a.ntiles, e.pole_sum, a.pole_area_top, a.pole_area_bot,
                        a.pole_mass_fixed, a.diag_hist, e.slot, red_d);
        }
    }
}
}  // namespace sf
