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
    const double2* diag_prev; // the previous step's per-block partials
    const float2* hpole_prev; // the previous step's ring partials (its pole mass)
    double2* diag_hist;       // ring buffer {mass, dt_cfl} per step
    int prev_slot;            // history slot of the previous step, -1 if none
    int prev_nblocks;
    int prev_pole_sum;
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

// Per-block FP64 partials of film volume and CFL lambda for this step (no
// fence and no atomics: they are consumed by the next launch); block 0 also
// reduces the previous step's partials into the history.
__device__ inline void diag_epilogue(const StepArgs& a, double mass, double lam, int lane, int wave,
                                     int nwaves, double (&red_d)[2][kBlockMax / 64]) {
    mass = wave_dsum(mass);
    lam = wave_dmax(lam);
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
        a.diag_part[blockIdx.y * gridDim.x + blockIdx.x] = make_double2(m, l);
    }
    if (a.prev_slot >= 0 && blockIdx.x == 0 && blockIdx.y == 0) {
        __syncthreads();
        diag_reduce(a.diag_prev, a.prev_nblocks, a.hpole_prev, a.ntiles, a.prev_pole_sum,
                    a.pole_area_top, a.pole_area_bot, a.pole_mass_fixed, a.diag_hist, a.prev_slot, red_d);
    }
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
        for (int v = 0; v < 4; ++v) {
            // Rupture mask on the input state (drainage.rs:331-333): frozen
            // cells keep h and Γ; the update is computed for every cell and
            // selected, which keeps the wave free of divergent branches.
            const bool rupt = ff_less(ch[v], cl[v], a.hc0, a.hc1, a.hc2);
            const float dtp = (sh[v] - ch[v]) + (sl[v] - cl[v]);   // toward theta + 1
            const float dtm = (ch[v] - nh[v]) + (cl[v] - nl[v]);
            const float dpp = (eh[v] - ch[v]) + (el[v] - cl[v]);
            const float dpm = (ch[v] - wh[v]) + (cl[v] - wl[v]);
            const float lap = (dtp - dtm) * G0 + (dtp + dtm) * G1 + (dpp - dpm) * G2;
            const float H = ch[v] + cl[v];
            float rate_dt = fmaf(dtD, lap, -cS * (H * H * H));
            if (MARANGONI) {
                const float gth = (sg[v] - ng[v]) * a.inv_2dth * a.inv_r;
                const float gph = (eg[v] - wg[v]) * a.inv_2dph * a.inv_r * rc.inv_s;
                const float mag = sqrtf(gth * gth + gph * gph);
                rate_dt += a.dt * a.mgr * mag * (H * H);
            }
            const float ln = fmaxf(cl[v] + rate_dt, -ch[v]);          // h >= 0 (drainage.rs:378)
            lo_new[v] = rupt ? cl[v] : ln;
            g_new[v] = cg[v];
            if (GWRITE) {
                const float gtp = sg[v] - cg[v], gtm = cg[v] - ng[v];
                const float gpp = eg[v] - cg[v], gpm = cg[v] - wg[v];
                const float glap = (gtp - gtm) * G0 + (gtp + gtm) * G1 + (gpp - gpm) * G2;
                float grate = a.ds_r2 * glap;
                if (MARANGONI) grate += 0.001f * lap * cg[v];
                const float gn = fminf(fmaxf(fmaf(a.dt, grate, cg[v]), 0.f), 1.f);
                g_new[v] = rupt ? cg[v] : gn;
            }
        }

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
                if (in_fronts(a, i, cc + v)) {
                    const float p = ch[v] * a.tfe_g_hi;
                    const float pe = fmaf(ch[v], a.tfe_g_hi, -p);
                    const float corr = pe + (ch[v] * a.tfe_g_lo + lo_new[v] * a.tfe_g_hi);
                    lo_new[v] = lo_new[v] + (p + corr);
                }
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
            if (DIAG) {
                // The lane's four cells share the row's area and CFL terms:
                // their FP64 sum is multiplied once per row.
                double hsum = 0.0;
#pragma unroll
                for (int v = 0; v < 4; ++v)
                    hsum += static_cast<double>(ch[v]) + static_cast<double>(lo_new[v]);
                const float hmax = fmaxf(fmaxf(ch[0] + lo_new[0], ch[1] + lo_new[1]),
                                         fmaxf(ch[2] + lo_new[2], ch[3] + lo_new[3]));
                const RowD rd = s_rowd[DIAG ? t : 0];
                mass = fma(hsum, rd.area, mass);
                const double hm = static_cast<double>(hmax);
                lam = fmax(lam, rd.lam_grid + a.coef3 * rd.sin_abs * hm * hm);
            }
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
        for (int v = 0; v < 4; ++v) {
            // Rupture mask on the input state (drainage.rs:331-333): frozen
            // cells keep h and Γ; the update is computed for every cell and
            // selected, which keeps the wave free of divergent branches.
            const bool rupt = ff_less(ch[v], cl[v], a.hc0, a.hc1, a.hc2);
            const float dtp = (sh[v] - ch[v]) + (sl[v] - cl[v]);   // toward theta + 1
            const float dtm = (ch[v] - nh[v]) + (cl[v] - nl[v]);
            const float dpp = (eh[v] - ch[v]) + (el[v] - cl[v]);
            const float dpm = (ch[v] - wh[v]) + (cl[v] - wl[v]);
            const float lap = (dtp - dtm) * G0 + (dtp + dtm) * G1 + (dpp - dpm) * G2;
            const float H = ch[v] + cl[v];
            float rate_dt = fmaf(dtD, lap, -cS * (H * H * H));
            if (MARANGONI) {
                const float gth = (sg[v] - ng[v]) * a.inv_2dth * a.inv_r;
                const float gph = (eg[v] - wg[v]) * a.inv_2dph * a.inv_r * rc.inv_s;
                const float mag = sqrtf(gth * gth + gph * gph);
                rate_dt += a.dt * a.mgr * mag * (H * H);
            }
            const float ln = fmaxf(cl[v] + rate_dt, -ch[v]);          // h >= 0 (drainage.rs:378)
            lo_new[v] = rupt ? cl[v] : ln;
            g_new[v] = cg[v];
            if (GWRITE) {
                const float gtp = sg[v] - cg[v], gtm = cg[v] - ng[v];
                const float gpp = eg[v] - cg[v], gpm = cg[v] - wg[v];
                const float glap = (gtp - gtm) * G0 + (gtp + gtm) * G1 + (gpp - gpm) * G2;
                float grate = a.ds_r2 * glap;
                if (MARANGONI) grate += 0.001f * lap * cg[v];
                const float gn = fminf(fmaxf(fmaf(a.dt, grate, cg[v]), 0.f), 1.f);
                g_new[v] = rupt ? cg[v] : gn;
            }
        }

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
                if (in_fronts(a, i, cc + v)) {
                    const float p = ch[v] * a.tfe_g_hi;
                    const float pe = fmaf(ch[v], a.tfe_g_hi, -p);
                    const float corr = pe + (ch[v] * a.tfe_g_lo + lo_new[v] * a.tfe_g_hi);
                    lo_new[v] = lo_new[v] + (p + corr);
                }
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
            if (DIAG) {
                // The lane's four cells share the row's area and CFL terms:
                // their FP64 sum is multiplied once per row.
                double hsum = 0.0;
#pragma unroll
                for (int v = 0; v < 4; ++v)
                    hsum += static_cast<double>(ch[v]) + static_cast<double>(lo_new[v]);
                const float hmax = fmaxf(fmaxf(ch[0] + lo_new[0], ch[1] + lo_new[1]),
                                         fmaxf(ch[2] + lo_new[2], ch[3] + lo_new[3]));
                const RowD rd = s_rowd[DIAG ? t : 0];
                mass = fma(hsum, rd.area, mass);
                const double hm = static_cast<double>(hmax);
                lam = fmax(lam, rd.lam_grid + a.coef3 * rd.sin_abs * hm * hm);
            }
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
}  // namespace sf
