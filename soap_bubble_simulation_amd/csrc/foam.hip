// foam.hip -- batched drainage of a foam of bubbles with uneven grids.
//
// Every bubble is its own latitude-longitude grid (n_theta x n_phi, any
// even n_phi) with its own radius, viscosity and time step; all of them
// advance together, one step of every bubble per launch sequence.  The
// state of all bubbles lives in shared concatenated arrays (the same
// float-float h + FP32 Γ layout as a single bubble, see
// drainage_kernels.hpp), so a step streams the whole foam once.
//
// Work is cut into wave items: one item is a band of up to 16-64 rows of a
// chunk of 64 x CPL columns of one bubble, streamed top to bottom by one
// wave through a five-row register window (north, centre, south, two rows in
// flight).  The tile shape (CPL in {1, 2, 4} columns per lane, band height)
// is chosen per bubble-size bucket by the autotuner below.  Items are
// grouped by code path (lane width, vector access) and each path is its own
// kernel, so that each runs at the occupancy of its own register count;
// within a path items keep their natural order (bubble, band, chunk) and
// blocks are renumbered per XCD, so that neighbouring items share halo rows
// and edge columns in one L2.  A run of 16 steps (one period of the
// ping-pong and rebase buffers) is captured in a HIP graph and replayed.
//
// Physics per cell is the drainage step of drainage.rs:298-392 with the
// surfactant transport of drainage.wgsl:178-201: h poles are ring means, Γ
// poles are copied through, no marginal-regeneration fronts (the reference
// keeps a scalar thickness per foam bubble, foam.rs:32-49, so there is no
// per-bubble front schedule to follow).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include <hip/hip_ext.h>

#include "drainage.hpp"

namespace sf {

struct FoamBubble {
    int64_t off;          // first cell of the bubble in the concatenated arrays
    int nt, np;
    int row_off;          // first row in the row-coefficient table
    int pole_off;         // first ring partial: [nchunks top][nchunks bottom]
    int nchunks;          // column chunks of the bubble's tile shape
    int variant;          // tile shape index (kShapeCpl, kShapeBand)
    int path;             // 0: 1 column per lane, 1: 2, 2: 4 (float4), 3: 4 (two float2)
    int active;           // stepped (1) or held (0), see Foam::set_active
    int lane0;            // lane of a chunk's first column: 1 where lanes 0 and 63 are halo lanes (two-step layout)
    int cw;               // columns per chunk: (64 - 2 lane0) x columns per lane
    float dt;             // s
    float dtD;            // dt D_h / R^2
    float dtDs;           // dt D_s / R^2
    float dtc;            // dt rho g / (3 eta)
    float G0;             // 1 / dtheta^2
    float np_f;
    float hc0, hc1, hc2;  // critical thickness, three-float expansion
};

// One wave's work: rows [r0, r1) of the column chunk starting at c0.
struct FoamItem {
    int bubble, r0, r1, c0, chunk, pad_[3];
};

// Per-row coefficients of one bubble: sin(theta), cot/(2 dtheta),
// 1/(sin^2 dphi^2).
struct FoamRow {
    float sin_t, G1, G2, pad;
};

struct FoamArgs {
    const float* hi_in;
    const float* lo_in;
    const float* g_in;
    float* hi_out;
    float* lo_out;
    float* g_out;
    const FoamRow* rows;
    const FoamBubble* bub;
    const float2* hpole_in;
    float2* hpole_out;
    int pole_mean;        // ring neighbours of the poles from hpole_in
};

// Tile shapes: columns per lane x rows per item (a band of rows is streamed
// top to bottom by one wave, so taller bands re-read fewer halo rows; the
// short bands of shapes 8 and 9 give small foams enough waves).
constexpr int kShapes = 10;
constexpr int kShapeCpl[kShapes] = {1, 1, 2, 2, 2, 4, 4, 4, 1, 1};
constexpr int kShapeBand[kShapes] = {16, 32, 16, 32, 64, 16, 32, 64, 4, 2};
// Foams above this many cells are bandwidth-bound: the autotuner leaves out
// the short-band shapes there, whose extra halo reads a bucket timed on its
// own (latency-bound) does not show.
constexpr int64_t kShortBandMaxCells = int64_t{1} << 23;
constexpr int kPaths = 4;     // 1, 2, 4 (float4) and 4 (two float2) columns per lane

// One row of a lane's columns plus its west and east neighbours.
template <int CPL>
struct FRow {
    float h[CPL], l[CPL], g[CPL];
    float hw, lw, gw, he, le, ge;
};

// Lane columns: CPL values from column cc.  With four columns per lane and
// n_phi = 2 mod 4 (V4 false) the two halves are separate float2 accesses and
// the second half starts at column cc2, which for the row's last lane wraps
// to column 0: its columns 2 and 3 then hold the east neighbour of column 1
// and are never stored.
template <int CPL, bool V4>
__device__ inline void load_vec(const float* row, int cc, int cc2, float (&v)[CPL]) {
    if constexpr (CPL == 4 && V4) {
        const float4 t = *reinterpret_cast<const float4*>(row + cc);
        v[0] = t.x; v[1] = t.y; v[2] = t.z; v[3] = t.w;
    } else if constexpr (CPL == 4) {
        const float2 t0 = *reinterpret_cast<const float2*>(row + cc);
        const float2 t1 = *reinterpret_cast<const float2*>(row + cc2);
        v[0] = t0.x; v[1] = t0.y; v[2] = t1.x; v[3] = t1.y;
    } else if constexpr (CPL == 2) {
        const float2 t = *reinterpret_cast<const float2*>(row + cc);
        v[0] = t.x; v[1] = t.y;
    } else {
        v[0] = row[cc];
    }
}

template <int CPL, bool V4>
__device__ inline void store_vec(float* row, int cc, bool full, const float (&v)[CPL]) {
    if constexpr (CPL == 4 && V4) {
        *reinterpret_cast<float4*>(row + cc) = make_float4(v[0], v[1], v[2], v[3]);
    } else if constexpr (CPL == 4) {
        *reinterpret_cast<float2*>(row + cc) = make_float2(v[0], v[1]);
        if (full) *reinterpret_cast<float2*>(row + cc + 2) = make_float2(v[2], v[3]);
    } else if constexpr (CPL == 2) {
        *reinterpret_cast<float2*>(row + cc) = make_float2(v[0], v[1]);
    } else {
        row[cc] = v[0];
    }
}

// Loads row r (clamped to the last row) of the lane's columns.  Every lane
// loads, valid or not, so the wave never diverges around memory accesses.
template <int CPL, bool V4>
__device__ inline void foam_load(const FoamArgs& a, int64_t off, int np, int nt, int r, int cc, int cc2,
                                 int jw, int je, FRow<CPL>& R) {
    const int64_t base = off + static_cast<int64_t>(min(r, nt - 1)) * np;
    load_vec<CPL, V4>(a.hi_in + base, cc, cc2, R.h);
    load_vec<CPL, V4>(a.lo_in + base, cc, cc2, R.l);
    load_vec<CPL, V4>(a.g_in + base, cc, cc2, R.g);
    R.hw = a.hi_in[base + jw];
    R.lw = a.lo_in[base + jw];
    R.gw = a.g_in[base + jw];
    R.he = a.hi_in[base + je];
    R.le = a.lo_in[base + je];
    R.ge = a.g_in[base + je];
}

template <int CPL>
__device__ inline void foam_pole_row(FRow<CPL>& R, float2 P) {
#pragma unroll
    for (int v = 0; v < CPL; ++v) {
        R.h[v] = P.x;
        R.l[v] = P.y;
    }
    R.hw = R.he = P.x;
    R.lw = R.le = P.y;
}

// Wave-wide float-float sum of the lane's new ring values; lane 0 stores it.
template <int CPL>
__device__ inline void ring_partial(const float (&h)[CPL], const float (&l)[CPL], int nvalid, int lane,
                                    float2* out) {
    float2 s = make_float2(0.f, 0.f);
#pragma unroll
    for (int v = 0; v < CPL; ++v) {
        if (v < nvalid) {
            float a_, e_;
            two_sum(h[v], l[v], a_, e_);
            s = ff_add(s, make_float2(a_, e_));
        }
    }
    s = wave_ff_sum(s);
    if (lane == 0) *out = s;
}

// Per-lane constants of one work item.
struct BandLane {
    int cc, cc2, jw, je;      // first column, second-half column, west and east neighbours
    int nvalid;               // columns this lane stores (0 for idle lanes)
    bool valid, full;
    float2 ptop, pbot;        // pole means (virtual poles)
    FoamRow rl;               // coefficients of row r0 + lane (read back with readlane)
};

__device__ inline float lane_value(float v, int t) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), t));
}

// One cell of a drainage step: the new low word of h and the new Γ from the
// cell (c) and its north, south, west and east neighbours (drainage.rs:298-392
// with the Γ transport of drainage.wgsl:178-201).  G1, G2 and cS are the
// row's coefficients (cot / 2 dtheta, 1 / (sin^2 dphi^2), dt rho g sin / 3 eta).
__device__ __forceinline__ void foam_cell(const FoamBubble& B, float G1, float G2, float cS, float ch, float cl,
                                          float cg, float nh, float nl, float ng, float sh, float sl, float sg,
                                          float wh, float wl, float wg, float eh, float el, float eg, float& lo_new,
                                          float& g_new) {
    // Every fused multiply-add is written out and no other is formed, so the
    // one- and two-step kernels, which inline this in different contexts,
    // round the same operations.
#pragma clang fp contract(off)
    // Rupture mask on the input state (drainage.rs:331-333).
    const bool rupt = ff_less(ch, cl, B.hc0, B.hc1, B.hc2);
    const float dtp = (sh - ch) + (sl - cl);
    const float dtm = (ch - nh) + (cl - nl);
    const float dpp = (eh - ch) + (el - cl);
    const float dpm = (ch - wh) + (cl - wl);
    const float lap = fmaf(dpp - dpm, G2, fmaf(dtp + dtm, G1, (dtp - dtm) * B.G0));
    const float H = ch + cl;
    const float ln = fmaxf(cl + fmaf(B.dtD, lap, -cS * (H * H * H)), -ch);   // h >= 0
    lo_new = rupt ? cl : ln;
    const float gtp = sg - cg, gtm = cg - ng;
    const float gpp = eg - cg, gpm = cg - wg;
    const float glap = fmaf(gpp - gpm, G2, fmaf(gtp + gtm, G1, (gtp - gtm) * B.G0));
    const float gn = fminf(fmaxf(fmaf(B.dtDs, glap, cg), 0.f), 1.f);  // drainage.wgsl:198-201
    g_new = rupt ? cg : gn;
}

// Computes and stores row i from its north, centre and south rows.  The
// pole rows are replaced in place: row 0 is reloaded after row 1, and row
// nt - 2 is the last row any band computes.
template <int CPL, bool V4, bool REBASE>
__device__ inline void band_row(const FoamArgs& a, const FoamItem& it, const FoamBubble& B, const BandLane& L,
                                int lane, int i, FRow<CPL>& N, const FRow<CPL>& C, FRow<CPL>& Sx) {
    const int np = B.np, nt = B.nt;
    if (a.pole_mean && i == 1) foam_pole_row<CPL>(N, L.ptop);
    if (a.pole_mean && i == nt - 2) foam_pole_row<CPL>(Sx, L.pbot);
    const int t = i - it.r0;                   // wave-uniform, < 64
    const float G1 = lane_value(L.rl.G1, t), G2 = lane_value(L.rl.G2, t);
    const float cS = B.dtc * lane_value(L.rl.sin_t, t);

    float wh[CPL], wl[CPL], wg[CPL], eh[CPL], el[CPL], eg[CPL];
    wh[0] = C.hw; wl[0] = C.lw; wg[0] = C.gw;
    eh[CPL - 1] = C.he; el[CPL - 1] = C.le; eg[CPL - 1] = C.ge;
#pragma unroll
    for (int v = 1; v < CPL; ++v) {
        wh[v] = C.h[v - 1]; wl[v] = C.l[v - 1]; wg[v] = C.g[v - 1];
        eh[v - 1] = C.h[v]; el[v - 1] = C.l[v]; eg[v - 1] = C.g[v];
    }
    float lo_new[CPL], g_new[CPL];
#pragma unroll
    for (int v = 0; v < CPL; ++v)
        foam_cell(B, G1, G2, cS, C.h[v], C.l[v], C.g[v], N.h[v], N.l[v], N.g[v], Sx.h[v], Sx.l[v], Sx.g[v], wh[v],
                  wl[v], wg[v], eh[v], el[v], eg[v], lo_new[v], g_new[v]);
    // Ring partials of this chunk for the pole means of the next step
    // (drainage.rs:480-503), from the new ring values.
    if (i == 1) ring_partial<CPL>(C.h, lo_new, L.nvalid, lane, a.hpole_out + B.pole_off + it.chunk);
    if (i == nt - 2) ring_partial<CPL>(C.h, lo_new, L.nvalid, lane, a.hpole_out + B.pole_off + B.nchunks + it.chunk);

    if (L.valid) {
        const int64_t o = B.off + static_cast<int64_t>(i) * np;
        if (REBASE) {
            float hh[CPL], ll[CPL];
#pragma unroll
            for (int v = 0; v < CPL; ++v) two_sum(C.h[v], lo_new[v], hh[v], ll[v]);
            store_vec<CPL, V4>(a.hi_out + o, L.cc, L.full, hh);
            store_vec<CPL, V4>(a.lo_out + o, L.cc, L.full, ll);
        } else {
            store_vec<CPL, V4>(a.lo_out + o, L.cc, L.full, lo_new);
        }
        store_vec<CPL, V4>(a.g_out + o, L.cc, L.full, g_new);
    }
}

// Streams rows [r0, r1) of one column chunk of one bubble through a window
// of five row buffers per lane: north, centre, south and two rows in
// flight.  The loop is unrolled over the five buffer rotations so that no
// buffer is ever copied (a copy would wait for the loads in flight).
template <int CPL, bool V4, bool REBASE>
__device__ void foam_band(const FoamArgs& a, const FoamItem& it, const FoamBubble& B, int lane) {
    const int np = B.np, nt = B.nt;
    BandLane L;
    const int c = it.c0 + CPL * (lane - B.lane0);      // lanes lane0 .. 63 - lane0 hold the chunk
    L.valid = lane >= B.lane0 && lane < 64 - B.lane0 && c < np;
    L.cc = L.valid ? c : 0;                            // idle lanes shadow column 0
    L.jw = L.cc == 0 ? np - 1 : L.cc - 1;
    L.je = (L.cc + CPL) % np;
    L.full = L.cc + CPL <= np;                         // false only for a 2-column tail lane
    L.cc2 = L.full ? L.cc + 2 : L.cc + 2 - np;
    L.nvalid = L.valid ? (L.full ? CPL : 2) : 0;
    const int r0 = it.r0, r1 = it.r1;
    L.rl = a.rows[B.row_off + min(r0 + lane, nt - 1)];   // bands hold at most 64 rows
    L.ptop = L.pbot = make_float2(0.f, 0.f);
    if (a.pole_mean) {
        if (r0 == 1) L.ptop = pole_mean(a.hpole_in + B.pole_off, B.nchunks, B.np_f);
        if (r1 == nt - 1) L.pbot = pole_mean(a.hpole_in + B.pole_off + B.nchunks, B.nchunks, B.np_f);
    }

    FRow<CPL> P0, P1, P2, P3, P4;
    foam_load<CPL, V4>(a, B.off, np, nt, r0 - 1, L.cc, L.cc2, L.jw, L.je, P0);
    foam_load<CPL, V4>(a, B.off, np, nt, r0, L.cc, L.cc2, L.jw, L.je, P1);
    foam_load<CPL, V4>(a, B.off, np, nt, r0 + 1, L.cc, L.cc2, L.jw, L.je, P2);
    foam_load<CPL, V4>(a, B.off, np, nt, r0 + 2, L.cc, L.cc2, L.jw, L.je, P3);
    foam_load<CPL, V4>(a, B.off, np, nt, r0 + 3, L.cc, L.cc2, L.jw, L.je, P4);

    // Row i uses (N, C, S); the N buffer then receives row i + 4.
#define SF_FOAM_ROW(Nb, Cb, Sb)                                                                    \
    {                                                                                              \
        if (i >= r1) break;                                                                        \
        band_row<CPL, V4, REBASE>(a, it, B, L, lane, i, Nb, Cb, Sb);                               \
        if (i + 4 <= r1) foam_load<CPL, V4>(a, B.off, np, nt, i + 4, L.cc, L.cc2, L.jw, L.je, Nb); \
        ++i;                                                                                       \
    }
    for (int i = r0;;) {
        SF_FOAM_ROW(P0, P1, P2)
        SF_FOAM_ROW(P1, P2, P3)
        SF_FOAM_ROW(P2, P3, P4)
        SF_FOAM_ROW(P3, P4, P0)
        SF_FOAM_ROW(P4, P0, P1)
    }
#undef SF_FOAM_ROW
}

// One work item per wave.  Items are grouped by code path (lane width and
// vector access) and every path is its own kernel, so that each runs at the
// occupancy its own register count allows.
template <int PATH, bool REBASE>
__global__ __launch_bounds__(256) void foam_step_kernel(const FoamArgs a, const FoamItem* items, int n) {
    const int lane = threadIdx.x & 63;
    // The item index is wave-uniform; readfirstlane lets the compiler keep it
    // (and the item and bubble records) in scalar registers.
    // Blocks are dealt to the 8 XCDs round-robin; renumbering them gives each
    // XCD a contiguous run of items, so neighbouring bands and chunks (their
    // halo rows and edge columns) meet in the same L2.
    const int nb = static_cast<int>(gridDim.x), b = static_cast<int>(blockIdx.x);
    const int v = (nb % kXcds == 0) ? (b % kXcds) * (nb / kXcds) + b / kXcds : b;
    const int k = __builtin_amdgcn_readfirstlane(v * 4 + static_cast<int>(threadIdx.x >> 6));
    if (k >= n) return;
    const FoamItem it = items[k];
    const FoamBubble B = a.bub[it.bubble];
    if constexpr (PATH == 0) foam_band<1, false, REBASE>(a, it, B, lane);
    else if constexpr (PATH == 1) foam_band<2, false, REBASE>(a, it, B, lane);
    else if constexpr (PATH == 2) foam_band<4, true, REBASE>(a, it, B, lane);
    else foam_band<4, false, REBASE>(a, it, B, lane);
}

// ---- two steps per launch (temporal blocking) -------------------------------
//
// foam_step2_kernel advances every bubble by two steps in one pass: a wave
// streams the step-s rows r0 - 2 .. r1 + 1 of its item's columns, forms the
// step-(s+1) rows r0 - 1 .. r1 into a three-row ring in its LDS, and forms the
// step-(s+2) rows r0 .. r1 - 1 from that ring, so the foam's state is read
// and written once per two steps.  Every cell update is foam_cell, the
// arithmetic of the one-step kernel, and the ring partial sums and pole means
// are formed in the same order, so a launch equals two launches of
// foam_step_kernel bit for bit:
//  * halo columns: with the two-step layout (FoamBubble::lane0 = 1) a chunk's
//    columns sit in lanes 1 .. nl, so lane 0 holds the CPL columns before the
//    chunk and lane nl + 1 those after it (both wrap in phi; a bubble that is
//    a single chunk thus re-forms its own edge columns).  Every lane forms
//    step s+1 of its columns, so the halo columns cost two lanes of 64, not an
//    extra cell per lane.  With a two-column tail lane (4 columns per lane,
//    n_phi = 2 mod 4) the east halo is that lane's third column;
//  * poles: a band touching a pole first re-forms that pole's ring row of step
//    s+1 over every chunk of its bubble and takes the step-(s+1) pole mean
//    from those partial sums, in chunk order.
// Scope: the pole rows are the ring means of the previous step (virtual
// poles), and a rebase only follows the second step.
template <int CPL>
constexpr int foam_ring_width() { return 64 * CPL + 4; }   // lane L's columns at CPL L .. CPL L + CPL - 1

template <int CPL>
struct FRow2 {
    float h[CPL], l[CPL], g[CPL];
    float hx, lx, gx;     // column jx: lane 0 the one west of its columns, lane nl + 1 the one east
};

// Loads step-s row r (clamped into the grid) of the lane's columns and its
// column jx.  Pole rows hold the ring mean of step s (virtual poles: ptop,
// pbot); Γ poles are stored.
template <int CPL, bool V4>
__device__ inline void foam_load2(const FoamArgs& a, const FoamBubble& B, int r, int cc, int cc2, int jx,
                                  float2 ptop, float2 pbot, FRow2<CPL>& R) {
    const int rr = max(0, min(r, B.nt - 1));
    const int64_t base = B.off + static_cast<int64_t>(rr) * B.np;
    if (rr == 0 || rr == B.nt - 1) {
        const float px = rr == 0 ? ptop.x : pbot.x;
        const float py = rr == 0 ? ptop.y : pbot.y;
#pragma unroll
        for (int v = 0; v < CPL; ++v) {
            R.h[v] = px;
            R.l[v] = py;
        }
        R.hx = px;
        R.lx = py;
    } else {
        load_vec<CPL, V4>(a.hi_in + base, cc, cc2, R.h);
        load_vec<CPL, V4>(a.lo_in + base, cc, cc2, R.l);
        R.hx = a.hi_in[base + jx];
        R.lx = a.lo_in[base + jx];
    }
    load_vec<CPL, V4>(a.g_in + base, cc, cc2, R.g);
    R.gx = a.g_in[base + jx];
}

__device__ __forceinline__ float lane_value_f(float v, int t) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), t));
}

// The step-(s+1) mean of pole ring row `ring` (1 or nt - 2) of bubble B: the
// ring row is formed over every chunk as foam_step_kernel forms it, each
// chunk's float-float sum reduced over the wave as ring_partial does, and
// the chunk sums added in chunk order as pole_mean adds the partials.
template <int CPL, bool V4>
__device__ __forceinline__ float2 foam_pole_after_step(const FoamArgs& a, const FoamBubble& B, int ring, float2 P0,
                                                       int lane) {
    const int np = B.np, nt = B.nt;
    const FoamRow rw = a.rows[B.row_off + ring];
    const float cS = B.dtc * rw.sin_t;
    float2 total = make_float2(0.f, 0.f);
    for (int ch = 0; ch < B.nchunks; ++ch) {
        const int c = ch * B.cw + CPL * (lane - B.lane0);
        const bool valid = lane >= B.lane0 && lane < 64 - B.lane0 && c < np;
        const int cc = valid ? c : 0;
        const int jw = cc == 0 ? np - 1 : cc - 1;
        const int je = (cc + CPL) % np;
        const bool full = cc + CPL <= np;
        const int cc2 = full ? cc + 2 : cc + 2 - np;
        const int nvalid = valid ? (full ? CPL : 2) : 0;
        FRow<CPL> N, C, S;
        foam_load<CPL, V4>(a, B.off, np, nt, ring - 1, cc, cc2, jw, je, N);
        foam_load<CPL, V4>(a, B.off, np, nt, ring, cc, cc2, jw, je, C);
        foam_load<CPL, V4>(a, B.off, np, nt, ring + 1, cc, cc2, jw, je, S);
        if (ring == 1) foam_pole_row<CPL>(N, P0);
        if (ring == nt - 2) foam_pole_row<CPL>(S, P0);
        float wh[CPL], wl[CPL], wg[CPL], eh[CPL], el[CPL], eg[CPL];
        wh[0] = C.hw; wl[0] = C.lw; wg[0] = C.gw;
        eh[CPL - 1] = C.he; el[CPL - 1] = C.le; eg[CPL - 1] = C.ge;
#pragma unroll
        for (int v = 1; v < CPL; ++v) {
            wh[v] = C.h[v - 1]; wl[v] = C.l[v - 1]; wg[v] = C.g[v - 1];
            eh[v - 1] = C.h[v]; el[v - 1] = C.l[v]; eg[v - 1] = C.g[v];
        }
        float lo1[CPL], g1[CPL];
#pragma unroll
        for (int v = 0; v < CPL; ++v)
            foam_cell(B, rw.G1, rw.G2, cS, C.h[v], C.l[v], C.g[v], N.h[v], N.l[v], N.g[v], S.h[v], S.l[v], S.g[v],
                      wh[v], wl[v], wg[v], eh[v], el[v], eg[v], lo1[v], g1[v]);
        float2 sum = make_float2(0.f, 0.f);
#pragma unroll
        for (int v = 0; v < CPL; ++v) {
            if (v < nvalid) {
                float a_, e_;
                two_sum(C.h[v], lo1[v], a_, e_);
                sum = ff_add(sum, make_float2(a_, e_));
            }
        }
        sum = wave_ff_sum(sum);
        total = ff_add(total, make_float2(lane_value_f(sum.x, 0), lane_value_f(sum.y, 0)));
    }
    return pole_mean_of_sum(total, B.np_f);
}

template <int CPL, bool V4, bool REBASE>
__device__ __forceinline__ void foam_band2(const FoamArgs& a, const FoamItem& it, const FoamBubble& B, int lane,
                                           float* ring) {
    constexpr int W = foam_ring_width<CPL>();
    const int np = B.np, nt = B.nt;
    const int r0 = it.r0, r1 = it.r1;
    // Lanes 1 .. nl hold the chunk's columns, lane 0 the CPL columns before it
    // and lane nl + 1 those after it; lanes past nl + 1 are idle.
    const int ncols = min(B.cw, np - it.c0);           // columns of the chunk inside the bubble
    const int nl = (ncols + CPL - 1) / CPL;            // at most 62
    const bool valid = lane >= 1 && lane <= nl;
    const bool active = lane <= nl + 1;
    const int craw = it.c0 + CPL * (lane - 1);         // below 0 for lane 0 of chunk 0, past n_phi for a halo
    const int cc = active ? (craw < 0 ? craw + np : (craw >= np ? craw - np : craw)) : 0;
    const bool full = cc + CPL <= np;                  // false only where a 4-column lane wraps (n_phi = 2 mod 4)
    const int cc2 = full ? cc + 2 : cc + 2 - np;
    const int nvalid = valid ? (full ? CPL : 2) : 0;
    const bool east = lane == nl + 1;                  // the east halo lane: its east neighbour is jx
    const int jx = lane == 0 ? (cc == 0 ? np - 1 : cc - 1) : (full ? (cc + CPL) % np : (cc2 + 2) % np);

    // Pole means of step s (p0) and of step s+1 (p1).
    float2 p0top = make_float2(0.f, 0.f), p0bot = make_float2(0.f, 0.f);
    float2 p1top = make_float2(0.f, 0.f), p1bot = make_float2(0.f, 0.f);
    if (r0 <= 2) p0top = pole_mean(a.hpole_in + B.pole_off, B.nchunks, B.np_f);
    if (r1 >= nt - 2) p0bot = pole_mean(a.hpole_in + B.pole_off + B.nchunks, B.nchunks, B.np_f);
    if (r0 == 1) p1top = foam_pole_after_step<CPL, V4>(a, B, 1, p0top, lane);
    if (r1 == nt - 1) p1bot = foam_pole_after_step<CPL, V4>(a, B, nt - 2, p0bot, lane);

    // Row coefficients of rows r0 - 1 .. r1 (at most 66), read with readlane.
    const int rb = r0 - 1;
    const FoamRow rl0 = a.rows[B.row_off + min(rb + lane, nt - 1)];
    const FoamRow rl1 = a.rows[B.row_off + min(rb + 64 + lane, nt - 1)];
    auto row_coef = [&](int k, float& G1, float& G2, float& sn) __attribute__((always_inline)) {
        const int t = k - rb;                          // wave-uniform
        if (t < 64) {
            G1 = lane_value_f(rl0.G1, t); G2 = lane_value_f(rl0.G2, t); sn = lane_value_f(rl0.sin_t, t);
        } else {
            G1 = lane_value_f(rl1.G1, t - 64); G2 = lane_value_f(rl1.G2, t - 64); sn = lane_value_f(rl1.sin_t, t - 64);
        }
    };
    auto slot = [&](int k, int arr) __attribute__((always_inline)) { return ring + ((k % 3) * 3 + arr) * W; };
    const int lc = CPL * lane;                         // the lane's first column in a ring row

    // Step s+1 row k from step-s rows (N, C, S) into the ring.
    auto first_step = [&](int k, const FRow2<CPL>& N, const FRow2<CPL>& C,
                          const FRow2<CPL>& S) __attribute__((always_inline)) {
        float G1, G2, sn;
        row_coef(k, G1, G2, sn);
        const float cS = B.dtc * sn;
        float wh[CPL], wl[CPL], wg[CPL], eh[CPL], el[CPL], eg[CPL];
        // West of the first column: lane - 1's last (lane 0: its column jx);
        // east of the last: lane + 1's first (the east halo lane: its jx).
        wh[0] = from_west(C.h[CPL - 1], C.hx);
        wl[0] = from_west(C.l[CPL - 1], C.lx);
        wg[0] = from_west(C.g[CPL - 1], C.gx);
        const float xeh = from_east(C.h[0], C.hx), xel = from_east(C.l[0], C.lx), xeg = from_east(C.g[0], C.gx);
        eh[CPL - 1] = east ? C.hx : xeh;
        el[CPL - 1] = east ? C.lx : xel;
        eg[CPL - 1] = east ? C.gx : xeg;
#pragma unroll
        for (int v = 1; v < CPL; ++v) {
            wh[v] = C.h[v - 1]; wl[v] = C.l[v - 1]; wg[v] = C.g[v - 1];
            eh[v - 1] = C.h[v]; el[v - 1] = C.l[v]; eg[v - 1] = C.g[v];
        }
        float lo1[CPL], g1[CPL];
#pragma unroll
        for (int v = 0; v < CPL; ++v)
            foam_cell(B, G1, G2, cS, C.h[v], C.l[v], C.g[v], N.h[v], N.l[v], N.g[v], S.h[v], S.l[v], S.g[v], wh[v],
                      wl[v], wg[v], eh[v], el[v], eg[v], lo1[v], g1[v]);
        // A pole row of step s+1 is not stepped: h and its low word are the
        // ring mean of step s+1, Γ is the stored pole value.
        const bool pole = k == 0 || k == nt - 1;
        const float px = k == 0 ? p1top.x : p1bot.x, py = k == 0 ? p1top.y : p1bot.y;
        float* rh = slot(k, 0);
        float* rl = slot(k, 1);
        float* rg = slot(k, 2);
        if (active) {
#pragma unroll
            for (int v = 0; v < CPL; ++v) {
                rh[lc + v] = pole ? px : C.h[v];
                rl[lc + v] = pole ? py : lo1[v];
                rg[lc + v] = pole ? C.g[v] : g1[v];
            }
        }
    };

    // Step s+2 row i from ring rows i - 1, i, i + 1.
    auto second_step = [&](int i) __attribute__((always_inline)) {
        const float* nh_ = slot(i - 1, 0); const float* nl_ = slot(i - 1, 1); const float* ng_ = slot(i - 1, 2);
        const float* ch_ = slot(i, 0);     const float* cl_ = slot(i, 1);     const float* cg_ = slot(i, 2);
        const float* sh_ = slot(i + 1, 0); const float* sl_ = slot(i + 1, 1); const float* sg_ = slot(i + 1, 2);
        float G1, G2, sn;
        row_coef(i, G1, G2, sn);
        const float cS = B.dtc * sn;
        float lo2[CPL], g2[CPL], chh[CPL];
#pragma unroll
        for (int v = 0; v < CPL; ++v) {
            const int p = lc + v;
            chh[v] = ch_[p];
            foam_cell(B, G1, G2, cS, ch_[p], cl_[p], cg_[p], nh_[p], nl_[p], ng_[p], sh_[p], sl_[p], sg_[p],
                      ch_[p - 1], cl_[p - 1], cg_[p - 1], ch_[p + 1], cl_[p + 1], cg_[p + 1], lo2[v], g2[v]);
        }
        // Ring partials of step s+2 for the next launch's pole means.
        if (i == 1) ring_partial<CPL>(chh, lo2, nvalid, lane, a.hpole_out + B.pole_off + it.chunk);
        if (i == nt - 2) ring_partial<CPL>(chh, lo2, nvalid, lane, a.hpole_out + B.pole_off + B.nchunks + it.chunk);
        if (valid) {
            const int64_t o = B.off + static_cast<int64_t>(i) * np;
            if (REBASE) {
                float hh[CPL], ll[CPL];
#pragma unroll
                for (int v = 0; v < CPL; ++v) two_sum(chh[v], lo2[v], hh[v], ll[v]);
                store_vec<CPL, V4>(a.hi_out + o, cc, full, hh);
                store_vec<CPL, V4>(a.lo_out + o, cc, full, ll);
            } else {
                store_vec<CPL, V4>(a.lo_out + o, cc, full, lo2);
            }
            store_vec<CPL, V4>(a.g_out + o, cc, full, g2);
        }
    };

    // Step-s rows k - 1, k, k + 1 form step s+1 row k (k = r0 - 1 .. r1);
    // after row k is in the ring, step s+2 row k - 1 follows.  Four row
    // buffers (one row in flight) unrolled over their rotations, so no buffer
    // is copied while its load is outstanding.
    FRow2<CPL> R0, R1, R2, R3;
    const int ks = r0 - 1;
    foam_load2<CPL, V4>(a, B, ks - 1, cc, cc2, jx, p0top, p0bot, R0);
    foam_load2<CPL, V4>(a, B, ks, cc, cc2, jx, p0top, p0bot, R1);
    foam_load2<CPL, V4>(a, B, ks + 1, cc, cc2, jx, p0top, p0bot, R2);
    foam_load2<CPL, V4>(a, B, ks + 2, cc, cc2, jx, p0top, p0bot, R3);
#define SF_FOAM2_ROW(Nb, Cb, Sb)                                                         \
    {                                                                                    \
        if (k > r1) break;                                                               \
        first_step(k, Nb, Cb, Sb);                                                       \
        __builtin_amdgcn_wave_barrier();                                                 \
        if (k - 1 >= r0) second_step(k - 1);                                             \
        __builtin_amdgcn_wave_barrier();                                                 \
        if (k + 3 <= r1 + 1) foam_load2<CPL, V4>(a, B, k + 3, cc, cc2, jx, p0top, p0bot, Nb); \
        ++k;                                                                             \
    }
    for (int k = ks;;) {
        SF_FOAM2_ROW(R0, R1, R2)
        SF_FOAM2_ROW(R1, R2, R3)
        SF_FOAM2_ROW(R2, R3, R0)
        SF_FOAM2_ROW(R3, R0, R1)
    }
#undef SF_FOAM2_ROW
}

// One two-step work item per wave; the ring of each wave is its own slice of
// the block's LDS.
template <int PATH, bool REBASE>
__global__ __launch_bounds__(256) void foam_step2_kernel(const FoamArgs a, const FoamItem* items, int n) {
    constexpr int CPL = PATH == 0 ? 1 : (PATH == 1 ? 2 : 4);
    constexpr int W = foam_ring_width<CPL>();
    __shared__ __align__(16) float ring[4][3 * 3 * W];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int nb = static_cast<int>(gridDim.x), b = static_cast<int>(blockIdx.x);
    const int v = (nb % kXcds == 0) ? (b % kXcds) * (nb / kXcds) + b / kXcds : b;
    const int k = __builtin_amdgcn_readfirstlane(v * 4 + wave);
    if (k >= n) return;
    const FoamItem it = items[k];
    const FoamBubble B = a.bub[it.bubble];
    if constexpr (PATH == 0) foam_band2<1, false, REBASE>(a, it, B, lane, ring[wave]);
    else if constexpr (PATH == 1) foam_band2<2, false, REBASE>(a, it, B, lane, ring[wave]);
    else if constexpr (PATH == 2) foam_band2<4, true, REBASE>(a, it, B, lane, ring[wave]);
    else foam_band2<4, false, REBASE>(a, it, B, lane, ring[wave]);
}

// Writes every bubble's pole rows from the ring partials (one block per bubble).
__global__ void foam_poles_kernel(float* hi, float* lo, const float2* hpole, const FoamBubble* bub) {
    const FoamBubble B = bub[blockIdx.x];
    if (!B.active) return;      // held bubbles have no fresh ring partials
    const float2 t = pole_mean(hpole + B.pole_off, B.nchunks, B.np_f);
    const float2 b = pole_mean(hpole + B.pole_off + B.nchunks, B.nchunks, B.np_f);
    for (int j = threadIdx.x; j < B.np; j += blockDim.x) {
        hi[B.off + j] = t.x;
        lo[B.off + j] = t.y;
        const int64_t o = B.off + static_cast<int64_t>(B.nt - 1) * B.np + j;
        hi[o] = b.x;
        lo[o] = b.y;
    }
}

// Per-bubble FP64 statistics: film volume, CFL lambda, min and max
// (one block per bubble).
struct FoamStat {
    double mass, lam, mn, mx;
};

__global__ __launch_bounds__(256) void foam_stats_kernel(const float* hi, const float* lo,
                                                          const FoamBubble* bub, const double* area,
                                                          const double* lamg, const double* sinab,
                                                          const double* coef3, FoamStat* out) {
    const FoamBubble B = bub[blockIdx.x];
    const int n = B.nt * B.np;
    double m = 0.0, l = 0.0, mn = INFINITY, mx = -INFINITY;
    for (int c = threadIdx.x; c < n; c += blockDim.x) {
        const int i = c / B.np;
        const double v = static_cast<double>(hi[B.off + c]) + static_cast<double>(lo[B.off + c]);
        m += v * area[B.row_off + i];
        mn = fmin(mn, v);
        mx = fmax(mx, v);
        if (i > 0 && i < B.nt - 1) l = fmax(l, lamg[B.row_off + i] + coef3[blockIdx.x] * sinab[B.row_off + i] * v * v);
    }
    __shared__ double s[4][4];
    m = wave_dsum(m);
    l = wave_dmax(l);
    for (int off = 32; off > 0; off >>= 1) {
        mn = fmin(mn, __shfl_down(mn, off));
        mx = fmax(mx, __shfl_down(mx, off));
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) { s[0][wave] = m; s[1][wave] = l; s[2][wave] = mn; s[3][wave] = mx; }
    __syncthreads();
    if (threadIdx.x == 0) {
        FoamStat r{s[0][0], s[1][0], s[2][0], s[3][0]};
        for (int w = 1; w < static_cast<int>(blockDim.x >> 6); ++w) {
            r.mass += s[0][w];
            r.lam = fmax(r.lam, s[1][w]);
            r.mn = fmin(r.mn, s[2][w]);
            r.mx = fmax(r.mx, s[3][w]);
        }
        out[blockIdx.x] = r;
    }
}

}  // namespace sf

using namespace sf;

// ---------------------------------------------------------------------------
// Host side.

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr int kFoamPeriod = 16;          // lo/Γ/pole ping-pong x rebase every 8 steps

template <class T>
T* falloc(size_t count) {
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, count * sizeof(T) + 16);
    if (e != hipSuccess) {
        set_error("hipMalloc of %zu bytes failed: %s", count * sizeof(T), hipGetErrorString(e));
        throw Error{SF_ERR_NOMEM};
    }
    return static_cast<T*>(p);
}

unsigned cdivu(size_t a, size_t b) { return static_cast<unsigned>((a + b - 1) / b); }

int bucket_of(int np) { return np <= 96 ? 0 : (np <= 192 ? 1 : (np <= 384 ? 2 : 3)); }


}  // namespace

struct Foam {
    sf_params p;
    int nb = 0;
    std::vector<uint8_t> active;        // per bubble: stepped (1) or held (0)
    std::vector<sf_bubble> desc;
    std::vector<FoamBubble> bub;
    std::vector<int64_t> offsets;
    int64_t total = 0;
    int total_rows = 0;
    hipStream_t stream = nullptr;
    int n_cu = 256;

    float* hi[2] = {nullptr, nullptr};
    float* lo[2] = {nullptr, nullptr};
    float* g[2] = {nullptr, nullptr};
    int cur = 0, cur_hi = 0, cur_pole = 0, since_rebase = 0;
    bool poles_virtual = false;
    // Two steps per launch (foam_step2_kernel): requested, and possible for
    // this foam (every bubble at least 4 x 8).
    int steps_per_launch = 2;
    bool two_ok = false;
    float2* hpole[2] = {nullptr, nullptr};
    int pole_slots = 0;
    FoamRow* rows = nullptr;
    FoamBubble* d_bub = nullptr;
    FoamItem* d_items = nullptr;
    int n_items = 0;
    int items_cap = 0;
    int pole_cap = 0;
    int path_begin[kPaths + 1] = {0, 0, 0, 0, 0};
    double* d_area = nullptr;
    double* d_lamg = nullptr;
    double* d_sinab = nullptr;
    double* d_coef3 = nullptr;
    FoamStat* d_stat = nullptr;
    int variant_of_bucket[4] = {1, 3, 6, 6};
    std::vector<double> time;
    int64_t steps_taken = 0;

    bool use_graph = true;
    hipGraphExec_t graph = nullptr;
    int graph_base = -1;                  // ping-pong phase the graph was captured at
    KernelTimer timer;

    ~Foam() {
        if (stream) (void)hipStreamSynchronize(stream);
        if (graph) (void)hipGraphExecDestroy(graph);
        for (int k = 0; k < 2; ++k) {
            (void)hipFree(hi[k]); (void)hipFree(lo[k]); (void)hipFree(g[k]); (void)hipFree(hpole[k]);
        }
        (void)hipFree(rows); (void)hipFree(d_bub); (void)hipFree(d_items); (void)hipFree(d_area);
        (void)hipFree(d_lamg); (void)hipFree(d_sinab); (void)hipFree(d_coef3); (void)hipFree(d_stat);
        timer.destroy();
        if (stream) (void)hipStreamDestroy(stream);
    }

    // Tile shape, chunking and pole-partial slots of bubble b.
    bool two_step() const { return steps_per_launch == 2 && two_ok; }

    void shape_bubble(int b) {
        FoamBubble& B = bub[b];
        B.variant = variant_of_bucket[bucket_of(B.np)];
        const int cpl = kShapeCpl[B.variant];
        B.path = cpl == 1 ? 0 : (cpl == 2 ? 1 : (B.np % 4 == 0 ? 2 : 3));
        // Where two steps per launch are possible a chunk's columns sit in
        // lanes 1..62 (lanes 0 and 63 form the halo columns of the two-step
        // kernel), in both launch modes, so that both share one layout and
        // one summation order of the pole ring partials.
        B.lane0 = two_ok ? 1 : 0;
        B.cw = (64 - 2 * B.lane0) * cpl;
        B.nchunks = static_cast<int>(cdivu(B.np, B.cw));
    }

    // Work items of bubble b: column chunks x bands of at most kShapeBand
    // rows, the interior rows split as evenly as the band count allows.
    void append_items(int b, std::vector<FoamItem>& items) const {
        const FoamBubble& B = bub[b];
        const int rows = B.nt - 2;
        const int nbands = (rows + kShapeBand[B.variant] - 1) / kShapeBand[B.variant];
        for (int k = 0; k < nbands; ++k) {
            const int r0 = 1 + static_cast<int>(static_cast<int64_t>(rows) * k / nbands);
            const int r1 = 1 + static_cast<int>(static_cast<int64_t>(rows) * (k + 1) / nbands);
            for (int ch = 0; ch < B.nchunks; ++ch) items.push_back(FoamItem{b, r0, r1, ch * B.cw, ch, {0, 0, 0}});
        }
    }

    // Work items, grouped by code path; within a path bubbles, bands and
    // chunks keep their natural order, so that the blocks running at the
    // same time stream neighbouring rows and share their halo rows and edge
    // columns in cache.  only_bucket >= 0 keeps the bubbles of one size
    // bucket (autotuning trials); pole slots are laid out for all bubbles.
    void layout_items(int only_bucket = -1) {
        SF_HIP(hipStreamSynchronize(stream));     // buffers below may still be in use
        std::vector<FoamItem> per_path[kPaths];
        pole_slots = 0;
        for (int b = 0; b < nb; ++b) {
            shape_bubble(b);
            bub[b].pole_off = pole_slots;
            pole_slots += 2 * bub[b].nchunks;
            bub[b].active = active[b];
            if (active[b] && (only_bucket < 0 || bucket_of(bub[b].np) == only_bucket))
                append_items(b, per_path[bub[b].path]);
        }
        std::vector<FoamItem> items;
        for (int p = 0; p < kPaths; ++p) {
            path_begin[p] = static_cast<int>(items.size());
            items.insert(items.end(), per_path[p].begin(), per_path[p].end());
        }
        path_begin[kPaths] = static_cast<int>(items.size());
        n_items = static_cast<int>(items.size());
        if (n_items > items_cap) {
            (void)hipFree(d_items);
            d_items = falloc<FoamItem>(n_items);
            items_cap = n_items;
        }
        SF_HIP(hipMemcpyAsync(d_items, items.data(), n_items * sizeof(FoamItem), hipMemcpyHostToDevice, stream));
        SF_HIP(hipMemcpyAsync(d_bub, bub.data(), nb * sizeof(FoamBubble), hipMemcpyHostToDevice, stream));
        if (pole_slots > pole_cap) {
            for (int k = 0; k < 2; ++k) {
                (void)hipFree(hpole[k]);
                hpole[k] = falloc<float2>(pole_slots);
            }
            pole_cap = pole_slots;
        }
        SF_HIP(hipStreamSynchronize(stream));
        drop_graph();
    }

    void drop_graph() {
        if (graph) (void)hipGraphExecDestroy(graph);
        graph = nullptr;
        graph_base = -1;
    }

    FoamArgs args_for(int cur_, int cur_hi_, int cur_pole_, bool pole_mean) const {
        FoamArgs a;
        a.hi_in = hi[cur_hi_];
        a.lo_in = lo[cur_];
        a.g_in = g[cur_];
        a.hi_out = hi[1 - cur_hi_];
        a.lo_out = lo[1 - cur_];
        a.g_out = g[1 - cur_];
        a.rows = rows;
        a.bub = d_bub;
        a.hpole_in = hpole[cur_pole_];
        a.hpole_out = hpole[1 - cur_pole_];
        a.pole_mean = pole_mean ? 1 : 0;
        return a;
    }

    // t0 / t1 (may be null): events recorded at the kernel's own start / end.
    template <int PATH>
    void launch_path(bool rebase, const FoamArgs& a, hipEvent_t t0, hipEvent_t t1) {
        const int n = path_begin[PATH + 1] - path_begin[PATH];
        if (n == 0) return;
        const dim3 grid((cdivu(static_cast<size_t>(n), 4) + kXcds - 1) / kXcds * kXcds), blk(256);
        const FoamItem* items = d_items + path_begin[PATH];
        if (rebase)
            hipExtLaunchKernelGGL((foam_step_kernel<PATH, true>), grid, blk, 0, stream, t0, t1, 0, a, items, n);
        else
            hipExtLaunchKernelGGL((foam_step_kernel<PATH, false>), grid, blk, 0, stream, t0, t1, 0, a, items, n);
    }

    template <int PATH>
    void launch_path2(bool rebase, const FoamArgs& a, hipEvent_t t0, hipEvent_t t1) {
        const int n = path_begin[PATH + 1] - path_begin[PATH];
        if (n == 0) return;
        const dim3 grid((cdivu(static_cast<size_t>(n), 4) + kXcds - 1) / kXcds * kXcds), blk(256);
        const FoamItem* items = d_items + path_begin[PATH];
        if (rebase)
            hipExtLaunchKernelGGL((foam_step2_kernel<PATH, true>), grid, blk, 0, stream, t0, t1, 0, a, items, n);
        else
            hipExtLaunchKernelGGL((foam_step2_kernel<PATH, false>), grid, blk, 0, stream, t0, t1, 0, a, items, n);
    }

    // One step (two with two_steps): a launch per code path in use; t0 is
    // recorded at the start of the first kernel and t1 at the end of the last.
    void launch_one(bool rebase, const FoamArgs& a, hipEvent_t t0 = nullptr, hipEvent_t t1 = nullptr,
                    bool two_steps = false) {
        int first = -1, last = -1;
        for (int p = 0; p < kPaths; ++p)
            if (path_begin[p + 1] > path_begin[p]) {
                if (first < 0) first = p;
                last = p;
            }
        if (two_steps) {
            launch_path2<0>(rebase, a, first == 0 ? t0 : nullptr, last == 0 ? t1 : nullptr);
            launch_path2<1>(rebase, a, first == 1 ? t0 : nullptr, last == 1 ? t1 : nullptr);
            launch_path2<2>(rebase, a, first == 2 ? t0 : nullptr, last == 2 ? t1 : nullptr);
            launch_path2<3>(rebase, a, first == 3 ? t0 : nullptr, last == 3 ? t1 : nullptr);
            return;
        }
        launch_path<0>(rebase, a, first == 0 ? t0 : nullptr, last == 0 ? t1 : nullptr);
        launch_path<1>(rebase, a, first == 1 ? t0 : nullptr, last == 1 ? t1 : nullptr);
        launch_path<2>(rebase, a, first == 2 ? t0 : nullptr, last == 2 ? t1 : nullptr);
        launch_path<3>(rebase, a, first == 3 ? t0 : nullptr, last == 3 ? t1 : nullptr);
    }

    // One step with the current buffers; advances the ping-pong state.
    void step_direct() {
        const bool rebase = since_rebase + 1 >= 8;
        const FoamArgs a = args_for(cur, cur_hi, cur_pole, poles_virtual);
        if (timer.enabled) {
            auto& ev = timer.next(stream);
            launch_one(rebase, a, ev.first, ev.second);
        } else {
            launch_one(rebase, a);
        }
        SF_HIP(hipGetLastError());
        advance(rebase);
    }

    // Two steps in one launch, from virtual poles at an even rebase phase
    // (a rebase then only follows the second step).
    void step_two_direct() {
        const bool rebase = since_rebase + 2 >= 8;
        const FoamArgs a = args_for(cur, cur_hi, cur_pole, true);
        if (timer.enabled) {
            auto& ev = timer.next(stream);
            launch_one(rebase, a, ev.first, ev.second, true);
        } else {
            launch_one(rebase, a, nullptr, nullptr, true);
        }
        SF_HIP(hipGetLastError());
        advance(false);
        advance(rebase);
        cur = 1 - cur;              // the launch flips the ping-pong buffers once
        cur_pole = 1 - cur_pole;
    }

    void advance(bool rebase) {
        cur = 1 - cur;
        cur_pole = 1 - cur_pole;
        if (rebase) {
            cur_hi = 1 - cur_hi;
            since_rebase = 0;
        } else {
            ++since_rebase;
        }
        poles_virtual = true;
        ++steps_taken;
        for (int b = 0; b < nb; ++b)
            if (active[b]) time[b] += static_cast<double>(bub[b].dt);
    }

    int phase() const { return (two_step() ? 32 : 0) | (cur_hi << 4) | (cur << 3) | since_rebase; }

    // Captures kFoamPeriod steps from the current phase into a graph.
    void capture() {
        drop_graph();
        hipGraph_t gr = nullptr;
        SF_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        int c = cur, ch = cur_hi, cp = cur_pole, sr = since_rebase;
        const int per_launch = two_step() ? 2 : 1;
        for (int s = 0; s < kFoamPeriod; s += per_launch) {
            const bool rebase = sr + per_launch >= 8;
            launch_one(rebase, args_for(c, ch, cp, true), nullptr, nullptr, per_launch == 2);
            c = 1 - c;
            cp = 1 - cp;
            if (rebase) { ch = 1 - ch; sr = 0; } else { sr += per_launch; }
        }
        hipError_t e = hipStreamEndCapture(stream, &gr);
        if (e == hipSuccess) e = hipGraphInstantiate(&graph, gr, nullptr, nullptr, 0);
        if (gr) (void)hipGraphDestroy(gr);
        SF_HIP(e);
        graph_base = phase();
    }

    void step(int n) {
        int done = 0;
        if (!poles_virtual && done < n) {   // first step reads the stored pole rows
            step_direct();
            ++done;
        }
        const bool two = two_step();
        if (two && (since_rebase & 1) && done < n) {   // pairs start at an even rebase phase
            step_direct();
            ++done;
        }
        if (use_graph && !timer.enabled) {
            while (n - done >= kFoamPeriod) {
                if (graph == nullptr || graph_base != phase()) capture();
                SF_HIP(hipGraphLaunch(graph, stream));
                for (int s = 0; s < kFoamPeriod; ++s) {
                    advance(since_rebase + 1 >= 8);
                    if (two && (s & 1)) {            // one ping-pong flip per launch
                        cur = 1 - cur;
                        cur_pole = 1 - cur_pole;
                    }
                }
                done += kFoamPeriod;
            }
        }
        while (two && n - done >= 2) {
            step_two_direct();
            done += 2;
        }
        while (done < n) {
            step_direct();
            ++done;
        }
    }

    void materialize() {
        if (!poles_virtual) return;
        hipLaunchKernelGGL(foam_poles_kernel, dim3(nb), dim3(256), 0, stream, hi[cur_hi], lo[cur],
                           hpole[cur_pole], d_bub);
        SF_HIP(hipGetLastError());
        poles_virtual = false;
    }

    std::vector<FoamStat> stats() {
        materialize();
        hipLaunchKernelGGL(foam_stats_kernel, dim3(nb), dim3(256), 0, stream, hi[cur_hi], lo[cur], d_bub,
                           d_area, d_lamg, d_sinab, d_coef3, d_stat);
        SF_HIP(hipGetLastError());
        std::vector<FoamStat> s(nb);
        SF_HIP(hipMemcpyAsync(s.data(), d_stat, nb * sizeof(FoamStat), hipMemcpyDeviceToHost, stream));
        SF_HIP(hipStreamSynchronize(stream));
        return s;
    }

    void set_dt(const std::vector<double>& dts) {
#pragma clang fp contract(off)
        for (int b = 0; b < nb; ++b) {
            const sf_bubble& d = desc[b];
            const double R = d.bubble_radius;
            const double dt = dts[b];
            FoamBubble& B = bub[b];
            B.dt = static_cast<float>(dt);
            B.dtD = static_cast<float>(dt * p.diffusion_coeff / (R * R));
            B.dtDs = (p.flags & SF_FLAG_SURFACTANT) ? static_cast<float>(dt * p.surfactant_diffusion / (R * R)) : 0.f;
            B.dtc = static_cast<float>(dt * p.density * p.gravity / (3.0 * d.viscosity));
        }
        SF_HIP(hipMemcpyAsync(d_bub, bub.data(), nb * sizeof(FoamBubble), hipMemcpyHostToDevice, stream));
        SF_HIP(hipStreamSynchronize(stream));
        drop_graph();
    }

    std::vector<double> bubble_dt() const {
        std::vector<double> r(nb);
        for (int b = 0; b < nb; ++b) r[b] = static_cast<double>(bub[b].dt);
        return r;
    }

    // dt of bubbles whose description asks for the CFL step (dt <= 0) from
    // the FP64 statistics kernel.
    void resolve_cfl() {
        bool any = false;
        for (const auto& d : desc) any |= !(d.dt > 0.0);
        std::vector<double> dts(nb);
        std::vector<FoamStat> st;
        if (any) st = stats();
        for (int b = 0; b < nb; ++b)
            dts[b] = desc[b].dt > 0.0 ? desc[b].dt : p.cfl_safety * (st[b].lam > 0.0 ? 2.0 / st[b].lam : 0.0);
        for (int b = 0; b < nb; ++b)
            SF_REQUIRE(dts[b] > 0.0 && std::isfinite(dts[b]), SF_ERR_INVALID, "bubble %d has no finite dt", b);
        set_dt(dts);
    }

    // Sources may be host or device memory.
    void upload(const float* h_hi, const float* h_lo, const float* gam) {
        for (int k = 0; k < 2; ++k) {
            if (h_hi) {
                SF_HIP(hipMemcpyAsync(hi[k], h_hi, total * sizeof(float), hipMemcpyDefault, stream));
                SF_HIP(hipMemcpyAsync(lo[k], h_lo, total * sizeof(float), hipMemcpyDefault, stream));
            }
            if (gam) SF_HIP(hipMemcpyAsync(g[k], gam, total * sizeof(float), hipMemcpyDefault, stream));
        }
        SF_HIP(hipStreamSynchronize(stream));
        cur = cur_hi = cur_pole = since_rebase = 0;
        poles_virtual = false;
        drop_graph();
    }

    // Saves and restores the device state around autotuning trials.
    struct Snapshot {
        std::vector<float*> buf;
    };
    Snapshot snapshot() {
        Snapshot s;
        float* src[6] = {hi[0], hi[1], lo[0], lo[1], g[0], g[1]};
        for (float* p_ : src) {
            float* d = falloc<float>(total);
            SF_HIP(hipMemcpyAsync(d, p_, total * sizeof(float), hipMemcpyDeviceToDevice, stream));
            s.buf.push_back(d);
        }
        SF_HIP(hipStreamSynchronize(stream));
        return s;
    }
    void restore(Snapshot& s) {
        float* dst[6] = {hi[0], hi[1], lo[0], lo[1], g[0], g[1]};
        for (int k = 0; k < 6; ++k) {
            SF_HIP(hipMemcpyAsync(dst[k], s.buf[k], total * sizeof(float), hipMemcpyDeviceToDevice, stream));
        }
        SF_HIP(hipStreamSynchronize(stream));
        for (float* p_ : s.buf) (void)hipFree(p_);
        s.buf.clear();
    }

    // Makes the whole state current in both ping-pong buffers (pole rows
    // written), so that a bubble that is not stepped reads the same values
    // whichever buffer is current.
    void unify_buffers() {
        materialize();
        SF_HIP(hipMemcpyAsync(hi[1 - cur_hi], hi[cur_hi], total * sizeof(float), hipMemcpyDeviceToDevice, stream));
        SF_HIP(hipMemcpyAsync(lo[1 - cur], lo[cur], total * sizeof(float), hipMemcpyDeviceToDevice, stream));
        SF_HIP(hipMemcpyAsync(g[1 - cur], g[cur], total * sizeof(float), hipMemcpyDeviceToDevice, stream));
    }

    // Only bubbles with mask[b] != 0 (all when mask is null) are stepped;
    // the others keep their state and time until they are stepped again.
    void set_active(const uint8_t* mask) {
        unify_buffers();
        for (int b = 0; b < nb; ++b) active[b] = mask ? (mask[b] != 0) : 1;
        layout_items();
    }

    // Times every item shape on each size bucket and keeps the fastest.
    void autotune(int trial_steps, int32_t* chosen, double* trial_ms) {
        if (trial_ms)
            for (int k = 0; k < 4 * kShapes; ++k) trial_ms[k] = -1.0;
        materialize();                  // the state is then whole in the current buffers
        SF_HIP(hipStreamSynchronize(stream));
        Snapshot snap = snapshot();
        const int c0 = cur, ch0 = cur_hi, cp0 = cur_pole, sr0 = since_rebase;
        const bool pv0 = poles_virtual;
        const int64_t st0 = steps_taken;
        const std::vector<double> time0 = time;
        std::vector<FoamBubble> keep = bub;
        const bool g0 = use_graph;
        use_graph = false;
        timer.drain(stream);            // the caller's own timing survives the trials
        const double t_ms = timer.ms;
        const int64_t t_launches = timer.launches;
        const bool t_enabled = timer.enabled;
        for (int bk = 0; bk < 4; ++bk) {
            bool present = false;
            for (const auto& B : bub) present |= bucket_of(B.np) == bk;
            if (!present) continue;
            double best = INFINITY;
            int best_v = variant_of_bucket[bk];
            const int nshapes = total > kShortBandMaxCells ? 8 : kShapes;
            for (int v = 0; v < nshapes; ++v) {
                // Only this bucket's bubbles take part in the trial.
                std::vector<FoamBubble> saved = bub;
                int save[4];
                std::copy(variant_of_bucket, variant_of_bucket + 4, save);
                variant_of_bucket[bk] = v;
                layout_items(bk);
                poles_virtual = false;      // first trial step reads the stored pole rows
                step_direct();              // untimed warm-up
                while (two_step() && (since_rebase & 1)) step_direct();
                timer.enabled = true;
                timer.ms = 0.0;
                timer.launches = 0;
                const int per_launch = two_step() ? 2 : 1;
                for (int s = 0; s < trial_steps; ++s) {
                    if (per_launch == 2)
                        step_two_direct();
                    else
                        step_direct();
                }
                timer.drain(stream);
                // ms per step of the trial's launches
                const double ms = timer.launches ? timer.ms / timer.launches / per_launch : INFINITY;
                timer.enabled = false;
                if (trial_ms) trial_ms[bk * kShapes + v] = ms;
                if (ms < best) {
                    best = ms;
                    best_v = v;
                }
                std::copy(save, save + 4, variant_of_bucket);
                bub = saved;
            }
            variant_of_bucket[bk] = best_v;
        }
        bub = keep;
        use_graph = g0;
        timer.ms = t_ms;
        timer.launches = t_launches;
        timer.enabled = t_enabled;
        restore(snap);
        cur = c0; cur_hi = ch0; cur_pole = cp0; since_rebase = sr0; poles_virtual = pv0;
        steps_taken = st0;
        time = time0;
        layout_items();
        if (chosen)
            for (int k = 0; k < 4; ++k) chosen[k] = variant_of_bucket[k];
    }

};

static Foam* F(sf_foam* f) { return reinterpret_cast<Foam*>(f); }

extern "C" int sf_foam_create(const sf_params* params, int32_t n, const sf_bubble* bubbles, sf_foam** out) {
    return guard([&] {
#pragma clang fp contract(off)
        SF_REQUIRE(out && params && bubbles && n > 0, SF_ERR_INVALID, "bad arguments");
        *out = nullptr;
        SF_REQUIRE((params->flags & ~SF_FLAG_SURFACTANT) == 0, SF_ERR_INVALID,
                   "foam bubbles support SF_FLAG_SURFACTANT only (h poles are ring means, Γ poles fixed; "
                   "no Marangoni term, no fronts); got flags 0x%x", params->flags);
        SF_REQUIRE(params->sin_guard > 0.0 && params->critical_thickness >= 0.0, SF_ERR_INVALID,
                   "sin_guard must be positive and critical_thickness non-negative");
        require_device();
        Foam* f = new Foam();
        try {
            f->p = *params;
            f->nb = n;
            f->desc.assign(bubbles, bubbles + n);
            int dev = 0;
            SF_HIP(hipGetDevice(&dev));
            hipDeviceProp_t prop;
            SF_HIP(hipGetDeviceProperties(&prop, dev));
            f->n_cu = prop.multiProcessorCount;
            SF_HIP(hipStreamCreateWithFlags(&f->stream, hipStreamNonBlocking));
            f->bub.resize(n);
            f->offsets.resize(n + 1);
            int64_t off = 0;
            int rows_total = 0;
            std::vector<FoamRow> rows;
            std::vector<double> area, lamg, sinab, coef3(n);
            const Float3 hc = split3(params->critical_thickness);
            for (int b = 0; b < n; ++b) {
                const sf_bubble& d = bubbles[b];
                SF_REQUIRE(d.n_theta >= 3 && d.n_phi >= 2 && d.n_phi % 2 == 0, SF_ERR_INVALID,
                           "bubble %d: need n_theta >= 3 and even n_phi >= 2 (got %dx%d)", b, d.n_theta, d.n_phi);
                SF_REQUIRE(d.bubble_radius > 0.0 && d.viscosity > 0.0, SF_ERR_INVALID,
                           "bubble %d: radius and viscosity must be positive", b);
                FoamBubble& B = f->bub[b];
                std::memset(&B, 0, sizeof(B));
                B.off = off;
                B.nt = d.n_theta;
                B.np = d.n_phi;
                B.row_off = rows_total;
                const double dth = kPi / static_cast<double>(d.n_theta - 1);
                const double dph = 2.0 * kPi / static_cast<double>(d.n_phi);
                B.G0 = static_cast<float>(1.0 / (dth * dth));
                B.np_f = static_cast<float>(d.n_phi);
                B.hc0 = hc.c0;
                B.hc1 = hc.c1;
                B.hc2 = hc.c2;
                const double R2 = d.bubble_radius * d.bubble_radius;
                double D = params->diffusion_coeff;
                if ((params->flags & SF_FLAG_SURFACTANT) && params->surfactant_diffusion > D)
                    D = params->surfactant_diffusion;
                coef3[b] = 3.0 * (params->density * params->gravity / (3.0 * d.viscosity));
                for (int i = 0; i < d.n_theta; ++i) {
                    const double th = static_cast<double>(i) * dth;
                    const double s = std::sin(th), c = std::cos(th);
                    const double ss = std::fabs(s) < params->sin_guard ? std::copysign(params->sin_guard, s) : s;
                    rows.push_back(FoamRow{static_cast<float>(s), static_cast<float>(c / ss / (2.0 * dth)),
                                           static_cast<float>(1.0 / (ss * ss * dph * dph)), 0.f});
                    const double lo_b = i == 0 ? 0.0 : th - 0.5 * dth;
                    const double hi_b = i == d.n_theta - 1 ? kPi : th + 0.5 * dth;
                    area.push_back(R2 * (std::cos(lo_b) - std::cos(hi_b)) * dph);
                    lamg.push_back(4.0 * D * (1.0 / (dth * dth) + 1.0 / (ss * ss * dph * dph)) / R2);
                    sinab.push_back(std::fabs(s));
                }
                rows_total += d.n_theta;
                f->offsets[b] = off;
                off += static_cast<int64_t>(d.n_theta) * d.n_phi;
                off = (off + 3) & ~int64_t(3);      // keep every bubble 16-byte aligned
            }
            f->offsets[n] = off;
            f->total = off;
            f->total_rows = rows_total;
            for (int k = 0; k < 2; ++k) {
                f->hi[k] = falloc<float>(off);
                f->lo[k] = falloc<float>(off);
                f->g[k] = falloc<float>(off);
            }
            f->rows = falloc<FoamRow>(rows_total);
            f->d_bub = falloc<FoamBubble>(n);
            f->d_area = falloc<double>(rows_total);
            f->d_lamg = falloc<double>(rows_total);
            f->d_sinab = falloc<double>(rows_total);
            f->d_coef3 = falloc<double>(n);
            f->d_stat = falloc<FoamStat>(n);
            SF_HIP(hipMemcpyAsync(f->rows, rows.data(), rows.size() * sizeof(FoamRow), hipMemcpyHostToDevice, f->stream));
            SF_HIP(hipMemcpyAsync(f->d_area, area.data(), area.size() * 8, hipMemcpyHostToDevice, f->stream));
            SF_HIP(hipMemcpyAsync(f->d_lamg, lamg.data(), lamg.size() * 8, hipMemcpyHostToDevice, f->stream));
            SF_HIP(hipMemcpyAsync(f->d_sinab, sinab.data(), sinab.size() * 8, hipMemcpyHostToDevice, f->stream));
            SF_HIP(hipMemcpyAsync(f->d_coef3, coef3.data(), coef3.size() * 8, hipMemcpyHostToDevice, f->stream));
            // uniform initial state from the descriptions
            std::vector<float> hh(off, 0.f), ll(off, 0.f), gg(off, 0.5f);
            for (int b = 0; b < n; ++b) {
                const float h_hi = static_cast<float>(bubbles[b].h0);
                const float h_lo = static_cast<float>(bubbles[b].h0 - static_cast<double>(h_hi));
                const int64_t cells = static_cast<int64_t>(bubbles[b].n_theta) * bubbles[b].n_phi;
                for (int64_t c = 0; c < cells; ++c) {
                    hh[f->offsets[b] + c] = h_hi;
                    ll[f->offsets[b] + c] = h_lo;
                }
            }
            f->time.assign(n, 0.0);
            f->active.assign(n, 1);
            f->two_ok = true;
            for (int b = 0; b < n; ++b) f->two_ok &= bubbles[b].n_theta >= 4 && bubbles[b].n_phi >= 8;
            // Two steps per launch by default where the foam is bandwidth-bound;
            // smaller foams are latency-bound and their recomputed halo rows
            // would outweigh the saved traffic.
            f->steps_per_launch = f->total > kShortBandMaxCells ? 2 : 1;
            // Small foams are latency-bound: short bands give them waves
            // (one 256 x 512 bubble: 4.4 us per step with 2-row bands against
            // 22.7 us with the 32-row default; 16 such bubbles: 11.1 against
            // 25.6 us with 4-row bands).  sf_foam_autotune refines per bucket.
            if (f->total <= (int64_t{1} << 19))
                std::fill(f->variant_of_bucket, f->variant_of_bucket + 4, 9);
            else if (f->total <= kShortBandMaxCells)
                std::fill(f->variant_of_bucket, f->variant_of_bucket + 4, 8);
            f->layout_items();
            f->upload(hh.data(), ll.data(), gg.data());
            f->resolve_cfl();
        } catch (...) {
            delete f;
            throw;
        }
        *out = reinterpret_cast<sf_foam*>(f);
    });
}

extern "C" void sf_foam_destroy(sf_foam* f) { delete F(f); }

extern "C" int sf_foam_layout(sf_foam* f, int64_t* offsets) {
    return guard([&] {
        SF_REQUIRE(f && offsets, SF_ERR_INVALID, "NULL argument");
        std::copy(F(f)->offsets.begin(), F(f)->offsets.end(), offsets);
    });
}

extern "C" int sf_foam_set_fields_f64(sf_foam* f, const double* h, const double* gamma) {
    return guard([&] {
        SF_REQUIRE(f != nullptr, SF_ERR_INVALID, "foam is NULL");
        Foam* fm = F(f);
        fm->materialize();
        std::vector<float> hh(fm->total), ll(fm->total), gg(fm->total);
        if (h) {
            for (int64_t c = 0; c < fm->total; ++c) {
                hh[c] = static_cast<float>(h[c]);
                ll[c] = static_cast<float>(h[c] - static_cast<double>(hh[c]));
            }
        } else {
            SF_HIP(hipMemcpy(hh.data(), fm->hi[fm->cur_hi], fm->total * 4, hipMemcpyDeviceToHost));
            SF_HIP(hipMemcpy(ll.data(), fm->lo[fm->cur], fm->total * 4, hipMemcpyDeviceToHost));
        }
        if (gamma) {
            for (int64_t c = 0; c < fm->total; ++c) gg[c] = static_cast<float>(gamma[c]);
        } else {
            SF_HIP(hipMemcpy(gg.data(), fm->g[fm->cur], fm->total * 4, hipMemcpyDeviceToHost));
        }
        fm->upload(hh.data(), ll.data(), gg.data());
        fm->resolve_cfl();
    });
}

extern "C" int sf_foam_set_fields(sf_foam* f, const float* h, const float* gamma) {
    return guard([&] {
        SF_REQUIRE(f != nullptr && h != nullptr, SF_ERR_INVALID, "NULL argument");
        Foam* fm = F(f);
        std::vector<float> zeros(fm->total, 0.f), gg;
        if (!gamma) {
            gg.assign(fm->total, 0.5f);
            gamma = gg.data();
        }
        fm->upload(h, zeros.data(), gamma);
        fm->resolve_cfl();
    });
}

extern "C" int sf_foam_copy_fields_f64(sf_foam* f, double* h, double* gamma) {
    return guard([&] {
        SF_REQUIRE(f != nullptr, SF_ERR_INVALID, "foam is NULL");
        Foam* fm = F(f);
        fm->materialize();
        std::vector<float> hh(fm->total), ll(fm->total), gg(fm->total);
        SF_HIP(hipMemcpyAsync(hh.data(), fm->hi[fm->cur_hi], fm->total * 4, hipMemcpyDeviceToHost, fm->stream));
        SF_HIP(hipMemcpyAsync(ll.data(), fm->lo[fm->cur], fm->total * 4, hipMemcpyDeviceToHost, fm->stream));
        SF_HIP(hipMemcpyAsync(gg.data(), fm->g[fm->cur], fm->total * 4, hipMemcpyDeviceToHost, fm->stream));
        SF_HIP(hipStreamSynchronize(fm->stream));
        for (int64_t c = 0; c < fm->total; ++c) {
            if (h) h[c] = static_cast<double>(hh[c]) + static_cast<double>(ll[c]);
            if (gamma) gamma[c] = gg[c];
        }
    });
}

extern "C" int sf_foam_step(sf_foam* f, int32_t n_steps) {
    return guard([&] {
        SF_REQUIRE(f != nullptr && n_steps >= 0, SF_ERR_INVALID, "bad arguments");
        F(f)->step(n_steps);
    });
}

extern "C" int sf_foam_synchronize(sf_foam* f) {
    return guard([&] {
        SF_REQUIRE(f != nullptr, SF_ERR_INVALID, "foam is NULL");
        SF_HIP(hipStreamSynchronize(F(f)->stream));
    });
}

extern "C" int sf_foam_use_graph(sf_foam* f, int32_t enable) {
    return guard([&] {
        SF_REQUIRE(f != nullptr, SF_ERR_INVALID, "foam is NULL");
        F(f)->use_graph = enable != 0;
        if (!enable) F(f)->drop_graph();
    });
}

extern "C" int sf_foam_set_steps_per_launch(sf_foam* f, int32_t steps) {
    return guard([&] {
        SF_REQUIRE(f != nullptr, SF_ERR_INVALID, "foam is NULL");
        SF_REQUIRE(steps == 1 || steps == 2, SF_ERR_INVALID, "steps per launch must be 1 or 2 (got %d)", steps);
        Foam* fm = F(f);
        SF_REQUIRE(steps == 1 || fm->two_ok, SF_ERR_INVALID,
                   "two steps per launch need every bubble at least 4 x 8 cells");
        if (steps == fm->steps_per_launch) return;
        fm->steps_per_launch = steps;
        fm->drop_graph();
    });
}

extern "C" int sf_foam_get_steps_per_launch(sf_foam* f, int32_t* steps) {
    return guard([&] {
        SF_REQUIRE(f != nullptr && steps != nullptr, SF_ERR_INVALID, "NULL argument");
        *steps = F(f)->two_step() ? 2 : 1;
    });
}

extern "C" int sf_foam_autotune(sf_foam* f, int32_t trial_steps, int32_t* chosen, double* trial_ms) {
    return guard([&] {
        SF_REQUIRE(f != nullptr && trial_steps > 0, SF_ERR_INVALID, "bad arguments");
        F(f)->autotune(trial_steps, chosen, trial_ms);
    });
}

extern "C" int sf_foam_set_tile_shapes(sf_foam* f, const int32_t* shapes) {
    return guard([&] {
        SF_REQUIRE(f != nullptr && shapes != nullptr, SF_ERR_INVALID, "NULL argument");
        for (int k = 0; k < 4; ++k)
            SF_REQUIRE(shapes[k] >= 0 && shapes[k] < kShapes, SF_ERR_INVALID, "tile shape %d out of range",
                       shapes[k]);
        Foam* fm = F(f);
        fm->materialize();              // pole partials change layout with the shapes
        std::copy(shapes, shapes + 4, fm->variant_of_bucket);
        fm->layout_items();
    });
}

extern "C" int sf_foam_bubble_dt(sf_foam* f, double* dt) {
    return guard([&] {
        SF_REQUIRE(f != nullptr && dt != nullptr, SF_ERR_INVALID, "NULL argument");
        const auto v = F(f)->bubble_dt();
        std::copy(v.begin(), v.end(), dt);
    });
}

extern "C" int sf_foam_stats(sf_foam* f, double* mass, double* dt_cfl, double* hmin, double* hmax) {
    return guard([&] {
        SF_REQUIRE(f != nullptr, SF_ERR_INVALID, "foam is NULL");
        const auto s = F(f)->stats();
        for (size_t b = 0; b < s.size(); ++b) {
            if (mass) mass[b] = s[b].mass;
            if (dt_cfl) dt_cfl[b] = s[b].lam > 0.0 ? 2.0 / s[b].lam : INFINITY;
            if (hmin) hmin[b] = s[b].mn;
            if (hmax) hmax[b] = s[b].mx;
        }
    });
}

extern "C" int sf_foam_kernel_timing(sf_foam* f, int32_t enable, double* avg_ms, int64_t* launches) {
    return guard([&] {
        SF_REQUIRE(f != nullptr, SF_ERR_INVALID, "foam is NULL");
        Foam* fm = F(f);
        fm->timer.drain(fm->stream);
        if (avg_ms) *avg_ms = fm->timer.launches ? fm->timer.ms / static_cast<double>(fm->timer.launches) : 0.0;
        if (launches) *launches = fm->timer.launches;
        fm->timer.ms = 0.0;
        fm->timer.launches = 0;
        fm->timer.enabled = enable != 0;
    });
}

extern "C" int sf_foam_info(sf_foam* f, int64_t* total_cells, int32_t* n_items, int32_t* variants) {
    return guard([&] {
        SF_REQUIRE(f != nullptr, SF_ERR_INVALID, "foam is NULL");
        Foam* fm = F(f);
        int64_t cells = 0;
        for (const auto& d : fm->desc) cells += static_cast<int64_t>(d.n_theta) * d.n_phi;
        if (total_cells) *total_cells = cells;
        if (n_items) *n_items = fm->n_items;
        if (variants)
            for (int k = 0; k < 4; ++k) variants[k] = fm->variant_of_bucket[k];
    });
}

extern "C" int sf_foam_set_active(sf_foam* f, const uint8_t* mask) {
    return guard([&] {
        SF_REQUIRE(f != nullptr, SF_ERR_INVALID, "foam is NULL");
        F(f)->set_active(mask);
    });
}

extern "C" int sf_foam_set_state(sf_foam* f, const float* hi, const float* lo, const float* gamma) {
    return guard([&] {
        SF_REQUIRE(f != nullptr && hi != nullptr && lo != nullptr, SF_ERR_INVALID, "NULL argument");
        Foam* fm = F(f);
        if (gamma == nullptr) fm->unify_buffers();   // keep Γ: current in both buffers
        fm->upload(hi, lo, gamma);
    });
}

extern "C" int sf_foam_device_state(sf_foam* f, const float** hi, const float** lo, const float** gamma) {
    return guard([&] {
        SF_REQUIRE(f != nullptr, SF_ERR_INVALID, "foam is NULL");
        Foam* fm = F(f);
        fm->materialize();
        SF_HIP(hipStreamSynchronize(fm->stream));
        if (hi) *hi = fm->hi[fm->cur_hi];
        if (lo) *lo = fm->lo[fm->cur];
        if (gamma) *gamma = fm->g[fm->cur];
    });
}

extern "C" int sf_foam_bubbles(sf_foam* f, sf_bubble* out) {
    return guard([&] {
        SF_REQUIRE(f != nullptr && out != nullptr, SF_ERR_INVALID, "NULL argument");
        Foam* fm = F(f);
        for (int b = 0; b < fm->nb; ++b) {
            out[b] = fm->desc[b];
            out[b].dt = static_cast<double>(fm->bub[b].dt);
        }
    });
}

extern "C" int sf_foam_stream(sf_foam* f, void** stream) {
    return guard([&] {
        SF_REQUIRE(f != nullptr && stream != nullptr, SF_ERR_INVALID, "NULL argument");
        *stream = F(f)->stream;
    });
}

extern "C" int sf_foam_num_bubbles(sf_foam* f, int32_t* n) {
    return guard([&] {
        SF_REQUIRE(f != nullptr && n != nullptr, SF_ERR_INVALID, "NULL argument");
        *n = F(f)->nb;
    });
}
