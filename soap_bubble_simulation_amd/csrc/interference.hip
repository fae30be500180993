// interference.hip -- thin-film interference kernels (gfx950).
//
//   lut_rgb_kernel      runtime LUT of interference_lut.rs:18-163 (FP32,
//                       7 wavelengths, CIE fits, RGBA8)
//   calc_kernel         InterferenceCalculator of interference.rs:197-512
//                       in FP32 (3 wavelengths, gamma 1/2.2): 1-D/2-D LUTs,
//                       per-point colour, intensity and intermediate terms
//   spectral_kernel     81-wavelength Airy/Fresnel reflectance at every grid
//                       cell's view angle, contracted with cmf[81][3] on BF16
//                       MFMA (v_mfma_f32_16x16x32_bf16) with FP32 accumulation
#include <cmath>
#include <cstring>
#include <vector>

#include <hip/hip_ext.h>

#include "drainage.hpp"

namespace sf {

constexpr int kLutT = 256;          // interference_lut.rs:9
constexpr int kLutA = 64;           // interference_lut.rs:10
constexpr float kLutMaxNm = 2000.f; // interference_lut.rs:13
constexpr float kPiF = 3.14159265358979323846f;

// ---------------------------------------------------------------------------
// Runtime RGBA8 LUT.  Operation order follows interference_lut.rs with FMA
// contraction disabled, so that values land within an ulp or two of the
// CPU build (transcendentals come from the device math library).

__device__ inline float max_f(float a, float b) { return fmaxf(a, b); }

__device__ inline float lut_gauss(float x, float mean, float sigma) {
#pragma clang fp contract(off)
    const float t = (x - mean) / sigma;
    return expf(-0.5f * t * t);
}

__device__ inline void lut_thin_film(float t_nm, float cos_theta, float n_film, float intensity,
                                     float rgb[3]) {
#pragma clang fp contract(off)
    const float W[7] = {400.f, 450.f, 500.f, 550.f, 600.f, 650.f, 700.f};
    // snells_law (96-100)
    const float sin_i = sqrtf(max_f(1.f - cos_theta * cos_theta, 0.f));
    const float sin_t = sin_i / n_film;
    const float cos_t = sqrtf(max_f(1.f - sin_t * sin_t, 0.f));
    // fresnel_unpolarized (103-120), n1 = 1
    const float n1_cos_i = 1.f * cos_theta;
    const float n2_cos_t = n_film * cos_t;
    const float r_s = (n1_cos_i - n2_cos_t) / max_f(n1_cos_i + n2_cos_t, 0.0001f);
    const float n2_cos_i = n_film * cos_theta;
    const float n1_cos_t = 1.f * cos_t;
    const float r_p = (n2_cos_i - n1_cos_t) / max_f(n2_cos_i + n1_cos_t, 0.0001f);
    const float fres = (r_s * r_s + r_p * r_p) * 0.5f;
    const float optical_path = 2.f * n_film * t_nm * cos_t;
    float x = 0.f, y = 0.f, z = 0.f;
    for (int k = 0; k < 7; ++k) {
        const float w = W[k];
        const float phase = 2.f * kPiF * optical_path / w + kPiF;
        // airy_interference (123-133)
        const float omr = max_f(1.f - fres, 0.001f);
        const float f = 4.f * fres / (omr * omr);
        const float s = sinf(phase * 0.5f);
        const float s2 = s * s;
        const float a = (f * s2) / max_f(1.f + f * s2, 0.001f);
        // cie_color_matching (136-148)
        const float cx = max_f(1.056f * lut_gauss(w, 599.8f, 37.9f) + 0.362f * lut_gauss(w, 442.f, 16.f) -
                                   0.065f * lut_gauss(w, 501.1f, 20.4f), 0.f);
        const float cy = max_f(0.821f * lut_gauss(w, 568.8f, 46.9f) + 0.286f * lut_gauss(w, 530.9f, 31.1f), 0.f);
        const float cz = max_f(1.217f * lut_gauss(w, 437.f, 11.8f) + 0.681f * lut_gauss(w, 459.f, 26.f), 0.f);
        x += cx * a;
        y += cy * a;
        z += cz * a;
    }
    x /= 7.f;
    y /= 7.f;
    z /= 7.f;
    // xyz_to_rgb (158-163)
    rgb[0] = (3.2404542f * x - 1.5371385f * y - 0.4985314f * z) * intensity;
    rgb[1] = (-0.969266f * x + 1.8760108f * y + 0.0415560f * z) * intensity;
    rgb[2] = (0.0556434f * x - 0.2040259f * y + 1.0572252f * z) * intensity;
}

// `(v.clamp(0, 1) * 255.0) as u8` (interference_lut.rs:33-35): NaN -> 0.
__device__ inline uint8_t quant_u8(float v) {
#pragma clang fp contract(off)
    if (v != v) return 0;
    const float c = v < 0.f ? 0.f : (v > 1.f ? 1.f : v);
    return static_cast<uint8_t>(c * 255.f);
}

__global__ void lut_rgb_kernel(float n_film, float intensity, float* rgb_out, uint8_t* rgba_out) {
#pragma clang fp contract(off)
    const int t = blockIdx.x * blockDim.x + threadIdx.x;   // thickness index
    const int a = blockIdx.y;                               // angle index
    if (t >= kLutT || a >= kLutA) return;
    const float t_nm = (static_cast<float>(t) / static_cast<float>(kLutT - 1)) * kLutMaxNm;
    const float cos_theta = static_cast<float>(a) / static_cast<float>(kLutA - 1);
    float rgb[3];
    lut_thin_film(t_nm, cos_theta, n_film, intensity, rgb);
    const size_t c = static_cast<size_t>(a) * kLutT + t;
    if (rgb_out) {
        rgb_out[3 * c + 0] = rgb[0];
        rgb_out[3 * c + 1] = rgb[1];
        rgb_out[3 * c + 2] = rgb[2];
    }
    if (rgba_out) {
        uchar4 px;
        px.x = quant_u8(rgb[0]);
        px.y = quant_u8(rgb[1]);
        px.z = quant_u8(rgb[2]);
        px.w = 255;
        reinterpret_cast<uchar4*>(rgba_out)[c] = px;
    }
}

// ---------------------------------------------------------------------------
// InterferenceCalculator in FP32.

struct CalcTerms {
    float cos_t, opd, phase, rs, rp, ravg;
};

// calculate_transmission_angle_cos / optical path / phase / Fresnel
// (interference.rs:197-298), n_medium = 1 (REFRACTIVE_INDEX_AIR).
__device__ inline CalcTerms calc_terms(float n_film, float d_nm, float cos_i, float lambda) {
#pragma clang fp contract(off)
    CalcTerms r;
    const float c = fminf(fmaxf(cos_i, -1.f), 1.f);
    const float sin_i = sqrtf(1.f - c * c);
    const float sin_t = (1.f / n_film) * sin_i;
    r.cos_t = fabsf(sin_t) > 1.f ? 0.f : sqrtf(1.f - sin_t * sin_t);
    r.opd = 2.f * n_film * d_nm * r.cos_t + lambda / 2.f;
    r.phase = 2.f * kPiF * r.opd / lambda;
    const float n1_cos_i = c, n2_cos_t = n_film * r.cos_t;
    const float n2_cos_i = n_film * c, n1_cos_t = r.cos_t;
    const float as = (n1_cos_i - n2_cos_t) / (n1_cos_i + n2_cos_t);
    const float ap = (n2_cos_i - n1_cos_t) / (n2_cos_i + n1_cos_t);
    r.rs = as * as;
    r.rp = ap * ap;
    r.ravg = (r.rs + r.rp) / 2.f;
    return r;
}

// calculate_reflected_intensity (interference.rs:312-353).  The phase is
// reduced to one period as OPD / lambda - round(OPD / lambda) before the
// cosine, which keeps FP32 accurate for thick films.
__device__ inline float calc_intensity(float n_film, float d_nm, float cos_i, float lambda) {
#pragma clang fp contract(off)
    if (d_nm <= 0.f) return 0.f;
    const CalcTerms t = calc_terms(n_film, d_nm, cos_i, lambda);
    const float cyc = t.opd / lambda;
    const float red = cyc - rintf(cyc);
    const float cos_phase = cospif(2.f * red);
    const float R = t.ravg;
    const float num = 2.f * R * (1.f - cos_phase);
    const float den = 1.f + R * R - 2.f * R * cos_phase;
    if (fabsf(den) < 2.220446049250313e-16f) return R;
    return fminf(fmaxf(num / den, 0.f), 1.f);
}

__device__ inline float calc_channel(float n_film, float d_nm, float cos_i, float lambda, int gamma) {
    float v = calc_intensity(n_film, d_nm, cos_i, lambda);
    if (gamma) v = powf(v, 1.f / 2.2f);
    return fminf(fmaxf(v, 0.f), 1.f);      // RgbColor::new (interference.rs:54-60)
}

// mode 0: 2-D LUT [t][a][3]; mode 1: 1-D LUT [t][3] at fixed cos; mode 2:
// per-point colour; mode 3: per-point intensity; mode 4: per-point terms.
// LUT thickness and cos values arrive as arrays formed on the host in FP64
// exactly as interference.rs:469-470 and 497-507 form them, then rounded to
// FP32, so that the LUT addressing matches the reference bit for bit.
struct CalcArgs {
    int mode, nt, na, gamma, count;
    float n_film, l0, l1, l2;
    const float* d_nm;
    const float* cos_i;
    const float* lambda;
    float* out;
};

__global__ void calc_kernel(CalcArgs a) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const float lam[3] = {a.l0, a.l1, a.l2};
    if (a.mode == 0) {
        if (k >= a.nt * a.na) return;
        const int ti = k / a.na, ai = k % a.na;
        const float d = a.d_nm[ti];
        const float c = a.cos_i[ai];
        for (int ch = 0; ch < 3; ++ch)
            a.out[3 * static_cast<size_t>(k) + ch] = calc_channel(a.n_film, d, c, lam[ch], a.gamma);
    } else if (a.mode == 1) {
        if (k >= a.nt) return;
        const float d = a.d_nm[k];
        const float c = a.cos_i[0];
        for (int ch = 0; ch < 3; ++ch)
            a.out[3 * static_cast<size_t>(k) + ch] = calc_channel(a.n_film, d, c, lam[ch], 1);
    } else if (a.mode == 2) {
        if (k >= a.count) return;
        for (int ch = 0; ch < 3; ++ch)
            a.out[3 * static_cast<size_t>(k) + ch] =
                calc_channel(a.n_film, a.d_nm[k], a.cos_i[k], lam[ch], a.gamma);
    } else if (a.mode == 3) {
        if (k >= a.count) return;
        a.out[k] = calc_intensity(a.n_film, a.d_nm[k], a.cos_i[k], a.lambda[k]);
    } else if (a.mode == 4) {
        if (k >= a.count) return;
        const CalcTerms t = calc_terms(a.n_film, a.d_nm[k], a.cos_i[k], a.lambda[k]);
        float* o = a.out + 6 * static_cast<size_t>(k);
        o[0] = t.cos_t; o[1] = t.opd; o[2] = t.phase; o[3] = t.rs; o[4] = t.rp; o[5] = t.ravg;
    } else if (a.mode == 6) {
        // calculate_phase_difference (interference.rs:248-254) of a given OPD.
        if (k >= a.count) return;
        a.out[k] = 2.f * kPiF * a.d_nm[k] / a.lambda[k];
    } else {
        // Optical path difference and phase for a given transmission cosine
        // (interference.rs:226-254).
        if (k >= a.count) return;
        const float lam = a.lambda[k];
        const float opd = 2.f * a.n_film * a.d_nm[k] * a.cos_i[k] + lam / 2.f;
        a.out[2 * static_cast<size_t>(k)] = opd;
        a.out[2 * static_cast<size_t>(k) + 1] = 2.f * kPiF * opd / lam;
    }
}

// ---------------------------------------------------------------------------
// Spectral 81-wavelength interference with the BF16 MFMA contraction.
//
// The Airy reflectance is I = F c^2 / (1 + F c^2) = 1 - r with
// r = 1 / (1 + F c^2), c = cos(pi opd / lambda), so with S_c = sum_k cmf[k][c]
//   sum_k cmf[k][c] I_k = S_c - sum_k cmf[k][c] r_k,
// and the MFMA contracts r with -cmf into an accumulator that starts at S_c.
// Where F <= kSpecFCos2 for the whole wave,
// 1 + F c^2 = (1 + F/2) + (F/2) cos(2 pi opd / lambda) saves the square; past
// that the cancellation in 1 + cos would cost accuracy, and the square form
// runs (with v_fract_f32 range reduction, which also covers long paths).
//
// v_mfma_f32_16x16x32_{f16,bf16}: lane l holds A[m][8 q + v] and B[8 q + v][m]
// (m = l % 16, q = l / 16, v = 0..7) and receives D[4 q + v][m].  A row m is
// a cell of a 16-cell group; lane (m, q) evaluates r for wavelengths
// k = q + 4 s, s = 0..20 (84 slots, 81 used, B zero on the rest), slot s in
// MFMA t = s / 8 at v = s % 8; slots 21..23 are zero.  A wave takes 64 cells
// at a time as four 16-cell groups.  The phases of slots 0..19 come from
// small FP16 MFMAs (v_mfma_f32_4x4x4_16b_f16, see phase_parts) rather than
// one VALU multiply per slot.
//
// The mean over the 81 wavelengths is folded into r: the kernel forms
// r' = r / 81 = 1 / (81 A + 81 B c) with the per-cell coefficients scaled,
// and starts each accumulator at sum_k cmf[k][c] / 81, so the contraction
// gives the RGB with no multiply per output.
//
// Operand formats.  r' lies in [0, 1/81], where FP16 (11-bit significand)
// rounds to nearest within 2^-18, so one FP16 operand changes channel c by
// at most 2^-18 sum_k |cmf[k][c]|; that is the default (one v_cvt_pk_f16_f32
// per two slots, one MFMA per group and curve part) whenever this bound stays
// within kSpecF16Budget and the curves fit FP16.  Otherwise r' is split into
// truncated BF16 hi and lo = r' - hi (exact), each contracted with the BF16
// curves, which keeps about 16 bits of r' at three more VALU operations per
// slot.  Either way the curves are a hi part plus the remainder's lo part in
// the same format, and hi times lo runs only when some lo is nonzero.

typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
typedef _Float16 f16x4 __attribute__((ext_vector_type(4)));

constexpr int kSpecLambdas = 81;
constexpr int kSpecSlots = 21;      // wavelength slots per lane and cell
constexpr int kSpecMfma = 3;        // 16x16x32 MFMAs per group and operand part
constexpr int kSpecPhase = 5;       // 4x4x4 (16-block) phase MFMAs per group (slots 0..19)
constexpr float kSpecFCos2 = 16.f;
// Largest RGB change the FP16 rounding of r' may cause for the FP16 operand
// format to be chosen automatically (a quarter of the 1e-3 RGB tolerance).
constexpr double kSpecF16Budget = 2.5e-4;
// Slot s holds wavelengths 4 s .. 4 s + 3 (slot 20: wavelength 80, 780 nm,
// and padding).  Trailing slots whose wavelengths the curves give no weight
// that matters are not evaluated: slot 20, and then slot 19 (760-775 nm),
// are skipped while sum_{skipped k} |cmf[k][c]| / 81 <= kSpecTailBudget for
// every c.  r is taken as 0 there (I = 1), which moves channel c by at most
// that sum (a hundredth of the RGB tolerance).  CIE 1931 curves skip slot 20
// (x-bar 4.2e-5 at 780 nm); curves that end below 740 nm skip both.
constexpr double kSpecTailBudget = 1e-5;
constexpr int kSpecMaxSkip = 2;     // trailing slots that may be skipped
// Paths below which v_cos_f32 (argument in revolutions, exact to its
// specification for |x| <= 256) takes opd / lambda unreduced.
constexpr float kSpecOpFast = 5.0e4f;

struct SpecArgs {
    const float* hi;
    const float* lo;
    const float4* rows;     // per theta row: sin, cos
    const float2* cols;     // per phi column: sin, cos
    const u32x4* bfrag;     // [2 (hi, lo)][3 MFMAs][64 lanes] B fragments of the operand format
    const float* ulam;      // [21 slots][4 q] 1 / (2 lambda) of wavelength q + 4 s (0 past 80)
    const u32x2* pfrag;     // [5 MFMAs][64 lanes] A fragments of the phase products (FP16 parts of 512 / lambda)
    float* rgb;             // 3 planes of n cells
    int np;
    size_t n;
    float cx, cy, cz, cam2;  // camera and |cam|^2
    float radius, n_film, inv_n_film;
    float s0, s1, s2;       // mean over the wavelengths of each (hi + lo) curve
    int b_lo;               // the curves have a nonzero lo part
    // Pole rows held as ring partials (lazy poles): their h is the ring mean
    // from hpole, as materialize_poles_kernel would write it.
    const float2* hpole;
    int pole_virtual, nt, ntiles;
    // Per-cell (F, cos theta_t) of the current camera and film index; the
    // geometry is fixed while they are, so spectral_view_kernel computes it
    // once; spectral_kernel reads it from here.
    const float2* view;
};

__device__ inline float lane_read(float v, int src_lane) {
    return __int_as_float(__builtin_amdgcn_ds_bpermute(src_lane << 2, __float_as_int(v)));
}

// Per-cell terms of the Airy sum: coefficient of finesse F and optical path
// difference opc (nm) of cell cc at its view angle.
// Row and column of cell cc (FP32 quotient, corrected).
__device__ inline void spec_cell_ij(const SpecArgs& a, size_t cc, float inv_np, int& i, int& j) {
    i = static_cast<int>(static_cast<float>(cc) * inv_np);
    j = static_cast<int>(cc - static_cast<size_t>(i) * a.np);
    if (j < 0) { --i; j += a.np; }
    if (j >= a.np) { ++i; j -= a.np; }
}

// Coefficient of finesse F and cos theta_t of cell (i, j) at its view angle.
__device__ inline void spec_view_terms(const SpecArgs& a, int i, int j, float& Fc, float& cos_t_out) {
    const float4 rw = a.rows[i];
    const float2 cl = a.cols[j];
    // View angle |dot(normalize(cam - p), n)|, p = R n (bubble.wgsl:491-494).
    const float nx = rw.x * cl.y, ny = rw.y, nz = rw.x * cl.x;
    const float cn = a.cx * nx + a.cy * ny + a.cz * nz;
    const float len2 = a.cam2 - 2.f * a.radius * cn + a.radius * a.radius;
    const float cos_i = fminf(fabsf(cn - a.radius) * rsqrtf(len2), 1.f);
    // Snell and unpolarised Fresnel (interference_lut.rs:96-120).
    const float sin_i = sqrtf(fmaxf(1.f - cos_i * cos_i, 0.f));
    const float sin_t = sin_i * a.inv_n_film;
    const float cos_t = sqrtf(fmaxf(1.f - sin_t * sin_t, 0.f));
    const float rs = (cos_i - a.n_film * cos_t) * __builtin_amdgcn_rcpf(fmaxf(cos_i + a.n_film * cos_t, 1e-4f));
    const float rp = (a.n_film * cos_i - cos_t) * __builtin_amdgcn_rcpf(fmaxf(a.n_film * cos_i + cos_t, 1e-4f));
    const float R = (rs * rs + rp * rp) * 0.5f;
    const float omr = fmaxf(1.f - R, 1e-3f);
    Fc = 4.f * R * __builtin_amdgcn_rcpf(omr * omr);
    cos_t_out = cos_t;
}

// Per-cell terms of the Airy sum: coefficient of finesse F and optical path
// difference opc (nm) of cell cc at its view angle, from the view table
// a.view (F, cos theta_t per cell) and the film state.  last_row is the
// index of the first cell of row nt - 1.
__device__ inline void spec_cell_terms(const SpecArgs& a, size_t cc, size_t last_row, float& Fc, float& opc) {
    // Streaming load: this VALU-bound pass has bandwidth to spare, and the
    // film state should keep the cache.
    typedef float f2v __attribute__((ext_vector_type(2)));
    const f2v v = __builtin_nontemporal_load(reinterpret_cast<const f2v*>(a.view) + cc);
    Fc = v.x;
    const float cos_t = v.y;
    const bool top = cc < static_cast<size_t>(a.np);
    float h;
    if (a.pole_virtual && (top || cc >= last_row)) {
        const float2 m = pole_mean(a.hpole + (top ? 0 : a.ntiles), a.ntiles, static_cast<float>(a.np));
        h = m.x + m.y;
    } else {
        h = a.hi[cc] + a.lo[cc];
    }
    const float d_nm = h * 1e9f;
    opc = 2.f * a.n_film * d_nm * cos_t;
}

// r' = r / 81 of one slot from its phase x in revolutions.  FAST:
// x = opd / lambda, A = 81 (1 + F/2), B = 81 F/2 (cos 2x form); otherwise
// x = opd / (2 lambda), A = 81, B = 81 F (square form, reduced argument).
template <bool FAST>
__device__ inline float spec_r_x(float x, float A, float B) {
    if (FAST) return __builtin_amdgcn_rcpf(fmaf(B, __builtin_amdgcn_cosf(x), A));
    const float c = __builtin_amdgcn_cosf(__builtin_amdgcn_fractf(x));
    return __builtin_amdgcn_rcpf(fmaf(B, c * c, A));
}

// High 16 bits of two floats packed as bf16 x 2 (truncation): x0 low, x1 high.
__device__ inline uint32_t pack_hi16(float x0, float x1) {
    return __builtin_amdgcn_perm(__float_as_uint(x1), __float_as_uint(x0), 0x07060302u);
}

// Two FP16 values (round to nearest even) in one register: x0 low, x1 high.
__device__ inline uint32_t pack_f16(float x0, float x1) {
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(f32x2{x0, x1}, f16x2));
}

// The phase products.  x = P / (2 lambda) of a cell and a wavelength slot is
// P' U with P' = P 2^-10 and U = 512 / lambda (nm), each split into two FP16
// parts, P' ~ p0 + 2^-11 p1 and U ~ u0 + 2^-11 u1 (p1, u1 the scaled
// remainders, so the parts are normal FP16 numbers away from tiny values):
//   x = p0 u0 + 2^-11 p0 u1 + 2^-11 p1 u0 + 2^-22 p1 u1,
// the exact product of the two 22-bit approximations (relative error of x
// at most about 2^-22, a phase error below 1e-5 rad for films up to 1 um).
// The four products are the K terms of v_mfma_f32_4x4x4_16b_f16 (16
// independent 4 x 4 x 4 blocks, two passes): lane l holds row l % 4 of
// block l / 4 in A, column l % 4 of block l / 4 in B, and receives column
// l % 4 of block l / 4 of D, rows 0..3.  B columns are cells (lane (m, q)
// supplies cell m of the group: [p0, 2^-11 p0, p1, 2^-11 p1]), A rows are
// the wavelengths q + 4 (4 e + i) of the lane's group q (host-packed:
// [u0, u1, 2^-11 u0, 2^-11 u1]), so lane (m, q) receives x of its own cell
// at slots 4 e .. 4 e + 3, in the registers the contraction reads.
struct PhaseParts {
    uint32_t x0, x1;       // [p0, 2^-11 p0], [p1, 2^-11 p1] as FP16 pairs
};

__device__ inline PhaseParts phase_parts(float P) {
    // P' = P 2^-10 <= 65504 (optical paths below 3.3e7 nm; longer ones are
    // clamped and lose the phase, which FP32 had already lost there).
    const float Pp = fminf(P * 0x1p-10f, 65504.f);
    const float p0 = static_cast<float>(static_cast<_Float16>(Pp));
    const float p1 = (Pp - p0) * 0x1p11f;
    PhaseParts o;
    o.x0 = pack_f16(p0, p0 * 0x1p-11f);
    o.x1 = pack_f16(p1, p1 * 0x1p-11f);
    return o;
}

__device__ inline uint32_t lane_read_u(uint32_t v, int src_lane) {
    return static_cast<uint32_t>(__builtin_amdgcn_ds_bpermute(src_lane << 2, static_cast<int>(v)));
}

// One 16-cell group: the phases of slots 0..19 on the phase MFMAs, slot 20
// (wavelengths 80..83, 81..83 padding) as one multiply; r on the VALU, but
// for the SKIP trailing slots; r contracted with the curves on 16x16x32 MFMAs
// into acc.
template <bool FAST, bool F16, int SKIP>
__device__ inline void spec_group(const u32x2 (&pa)[kSpecPhase], u32x2 bop, float u20, const u32x4 (&bh)[kSpecMfma],
                                  const u32x4 (&bl)[kSpecMfma], bool b_lo, float P, float A, float B,
                                  v4f& acc) {
    const f16x4 bp = __builtin_bit_cast(f16x4, bop);
    const v4f z = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int t = 0; t < kSpecMfma; ++t) {
        float r[8];
        const v4f x0 = __builtin_amdgcn_mfma_f32_4x4x4f16(__builtin_bit_cast(f16x4, pa[2 * t]), bp, z, 0, 0, 0);
#pragma unroll
        for (int v = 0; v < 4; ++v)
            r[v] = 8 * t + v < kSpecSlots - SKIP ? spec_r_x<FAST>(x0[v], A, B) : 0.f;
        if (2 * t + 1 < kSpecPhase) {
            const v4f x1 =
                __builtin_amdgcn_mfma_f32_4x4x4f16(__builtin_bit_cast(f16x4, pa[2 * t + 1]), bp, z, 0, 0, 0);
#pragma unroll
            for (int v = 0; v < 4; ++v) r[4 + v] = spec_r_x<FAST>(x1[v], A, B);
        } else {
            r[4] = SKIP == 0 ? spec_r_x<FAST>(P * u20, A, B) : 0.f;
            r[5] = r[6] = r[7] = 0.f;
        }
        if (F16) {
            u32x4 h;
#pragma unroll
            for (int p = 0; p < 4; ++p) h[p] = pack_f16(r[2 * p], r[2 * p + 1]);
            const f16x8 ah = __builtin_bit_cast(f16x8, h);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, __builtin_bit_cast(f16x8, bh[t]), acc, 0, 0, 0);
            if (b_lo) acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, __builtin_bit_cast(f16x8, bl[t]), acc, 0, 0, 0);
        } else {
            u32x4 h, l;
#pragma unroll
            for (int p = 0; p < 4; ++p) {
                const float h0 = __uint_as_float(__float_as_uint(r[2 * p]) & 0xFFFF0000u);
                const float h1 = __uint_as_float(__float_as_uint(r[2 * p + 1]) & 0xFFFF0000u);
                h[p] = pack_hi16(r[2 * p], r[2 * p + 1]);
                l[p] = pack_hi16(r[2 * p] - h0, r[2 * p + 1] - h1);
            }
            const bf16x8 ah = __builtin_bit_cast(bf16x8, h);
            const bf16x8 al = __builtin_bit_cast(bf16x8, l);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, __builtin_bit_cast(bf16x8, bh[t]), acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, __builtin_bit_cast(bf16x8, bh[t]), acc, 0, 0, 0);
            if (b_lo)
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, __builtin_bit_cast(bf16x8, bl[t]), acc, 0, 0, 0);
        }
    }
}

// Five waves per SIMD (96 VGPRs; in some instantiations the compiler moves a
// few per-batch values to scratch outside the wavelength loop): 1.4 % faster
// than four.
template <bool F16, int SKIP>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5, 8))) void spectral_kernel(const SpecArgs a) {
    const int lane = threadIdx.x & 63;
    const int m = lane & 15;
    const int q = lane >> 4;
    const size_t wave_global = (static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
    const size_t nwaves = (static_cast<size_t>(gridDim.x) * blockDim.x) >> 6;

    // This lane's phase-MFMA A fragments, 1 / (2 lambda) of its slot 20 (0 on
    // the padding) and its B fragments.
    u32x2 pa[kSpecPhase];
#pragma unroll
    for (int e = 0; e < kSpecPhase; ++e) pa[e] = a.pfrag[e * 64 + lane];
    const float u20 = a.ulam[4 * (kSpecSlots - 1) + q];
    u32x4 bh[kSpecMfma], bl[kSpecMfma];
#pragma unroll
    for (int t = 0; t < kSpecMfma; ++t) {
        bh[t] = a.bfrag[t * 64 + lane];
        bl[t] = a.bfrag[(kSpecMfma + t) * 64 + lane];
    }
    const bool b_lo = a.b_lo != 0;
    // The B fragments hold -cmf, so starting the accumulator at the curve's
    // sum S / 81 gives (S - sum_k cmf_k r_k) / 81 directly.
    const float S = m == 0 ? a.s0 : (m == 1 ? a.s1 : a.s2);
    const size_t last_row = static_cast<size_t>(a.nt - 1) * static_cast<size_t>(a.np);
    const bool vec_out = (a.n & 3) == 0;

    for (size_t sg = wave_global; sg * 64 < a.n; sg += nwaves) {
        const size_t cell = sg * 64 + lane;
        float Fc, opc;
        spec_cell_terms(a, cell < a.n ? cell : a.n - 1, last_row, Fc, opc);
        const bool fast = !__any(!(Fc <= kSpecFCos2 && opc <= kSpecOpFast));
        // This lane's cell's (P, 81 A, 81 B) of the chosen form and the FP16
        // parts of P, read by the lanes of its 16-cell group.
        const float Pc = fast ? 2.f * opc : opc;
        const float Ac = fast ? fmaf(40.5f, Fc, 81.f) : 81.f;
        const float Bc = fast ? 40.5f * Fc : 81.f * Fc;
        const PhaseParts pp = phase_parts(Pc);
        v4f acc[4];
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            const int src = 16 * g + m;
            const float P = lane_read(Pc, src);
            const float A = lane_read(Ac, src);
            const float B = lane_read(Bc, src);
            u32x2 bop;
            bop[0] = lane_read_u(pp.x0, src);
            bop[1] = lane_read_u(pp.x1, src);
            acc[g] = v4f{S, S, S, S};
            if (fast)
                spec_group<true, F16, SKIP>(pa, bop, u20, bh, bl, b_lo, P, A, B, acc[g]);
            else
                spec_group<false, F16, SKIP>(pa, bop, u20, bh, bl, b_lo, P, A, B, acc[g]);
        }
        // acc[g][v] = D[4 q + v][m]: cell 64 sg + 16 g + 4 q + v, colour m.
        if (m < 3) {
            float* plane = a.rgb + static_cast<size_t>(m) * a.n;
            if (vec_out && sg * 64 + 64 <= a.n) {
#pragma unroll
                for (int g = 0; g < 4; ++g) {
                    // Streaming store: the RGB planes are not re-read here, so
                    // they should not displace the film state from the caches.
                    typedef float f4v __attribute__((ext_vector_type(4)));
                    __builtin_nontemporal_store(f4v{acc[g][0], acc[g][1], acc[g][2], acc[g][3]},
                                                reinterpret_cast<f4v*>(plane + sg * 64 + 16 * g + 4 * q));
                }
            } else {
#pragma unroll
                for (int g = 0; g < 4; ++g) {
                    const size_t c0 = sg * 64 + 16 * g + 4 * q;
#pragma unroll
                    for (int v = 0; v < 4; ++v)
                        if (c0 + v < a.n) plane[c0 + v] = acc[g][v];
                }
            }
        }
    }
}

__global__ __launch_bounds__(256) void spectral_view_kernel(const SpecArgs a, float2* __restrict__ out) {
    const size_t cc = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
    if (cc >= a.n) return;
    int i, j;
    spec_cell_ij(a, cc, 1.f / static_cast<float>(a.np), i, j);
    float F, cos_t;
    spec_view_terms(a, i, j, F, cos_t);
    out[cc] = make_float2(F, cos_t);
}

static inline unsigned cdiv_u(size_t a, size_t b) { return static_cast<unsigned>((a + b - 1) / b); }

template <class T>
static T* dalloc_i(size_t count) {
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, count * sizeof(T) + 16);
    if (e != hipSuccess) {
        set_error("hipMalloc of %zu bytes failed: %s", count * sizeof(T), hipGetErrorString(e));
        throw Error{SF_ERR_NOMEM};
    }
    return static_cast<T*>(p);
}

// Device scratch for the stateless calls, freed on exit of the call.
struct DevBuf {
    void* p = nullptr;
    explicit DevBuf(size_t bytes) {
        hipError_t e = hipMalloc(&p, bytes + 16);
        if (e != hipSuccess) {
            set_error("hipMalloc of %zu bytes failed: %s", bytes, hipGetErrorString(e));
            throw Error{SF_ERR_NOMEM};
        }
    }
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    template <class T>
    T* as() const { return static_cast<T*>(p); }
};

static void run_calc(CalcArgs a, size_t out_floats, float* host_out, size_t threads,
                     const float* d_nm, size_t n_d, const float* cos_i, size_t n_c,
                     const float* lam, size_t n_l) {
    require_device();
    DevBuf out(out_floats * sizeof(float));
    DevBuf in((n_d + n_c + n_l) * sizeof(float) + 16);
    a.out = out.as<float>();
    float* base = in.as<float>();
    if (n_d) SF_HIP(hipMemcpy(base, d_nm, n_d * sizeof(float), hipMemcpyHostToDevice));
    if (n_c) SF_HIP(hipMemcpy(base + n_d, cos_i, n_c * sizeof(float), hipMemcpyHostToDevice));
    if (n_l) SF_HIP(hipMemcpy(base + n_d + n_c, lam, n_l * sizeof(float), hipMemcpyHostToDevice));
    a.d_nm = base;
    a.cos_i = base + n_d;
    a.lambda = base + n_d + n_c;
    if (threads) {
        hipLaunchKernelGGL(calc_kernel, dim3(cdiv_u(threads, 256)), dim3(256), 0, nullptr, a);
        SF_HIP(hipGetLastError());
    }
    SF_HIP(hipMemcpy(host_out, a.out, out_floats * sizeof(float), hipMemcpyDeviceToHost));
}

}  // namespace sf

using namespace sf;

extern "C" int sf_lut_rgba8(float refractive_index, float intensity, uint8_t* out) {
    return guard([&] {
        SF_REQUIRE(out != nullptr, SF_ERR_INVALID, "out is NULL");
        require_device();
        DevBuf d(static_cast<size_t>(kLutT) * kLutA * 4);
        hipLaunchKernelGGL(lut_rgb_kernel, dim3(kLutT / 64, kLutA), dim3(64), 0, nullptr,
                           refractive_index, intensity, static_cast<float*>(nullptr), d.as<uint8_t>());
        SF_HIP(hipGetLastError());
        SF_HIP(hipMemcpy(out, d.p, static_cast<size_t>(kLutT) * kLutA * 4, hipMemcpyDeviceToHost));
    });
}

extern "C" int sf_lut_rgb_f32(float refractive_index, float intensity, float* out) {
    return guard([&] {
        SF_REQUIRE(out != nullptr, SF_ERR_INVALID, "out is NULL");
        require_device();
        const size_t nb = static_cast<size_t>(kLutT) * kLutA * 3 * sizeof(float);
        DevBuf d(nb);
        hipLaunchKernelGGL(lut_rgb_kernel, dim3(kLutT / 64, kLutA), dim3(64), 0, nullptr,
                           refractive_index, intensity, d.as<float>(), static_cast<uint8_t*>(nullptr));
        SF_HIP(hipGetLastError());
        SF_HIP(hipMemcpy(out, d.p, nb, hipMemcpyDeviceToHost));
    });
}

static CalcArgs calc_base(double n, const double* w) {
    CalcArgs a{};
    a.n_film = static_cast<float>(n);
    a.l0 = static_cast<float>(w ? w[0] : 650.0);   // wavelengths::RED_NM (interference.rs:26)
    a.l1 = static_cast<float>(w ? w[1] : 532.0);   // GREEN_NM (28)
    a.l2 = static_cast<float>(w ? w[2] : 450.0);   // BLUE_NM (30)
    return a;
}

extern "C" int sf_lut_2d(double refractive_index, const double wavelengths_nm[3], double tmin_nm,
                         double tmax_nm, int32_t n_thickness, int32_t n_angle, int32_t gamma,
                         float* out) {
    return guard([&] {
#pragma clang fp contract(off)
        SF_REQUIRE(out != nullptr && n_thickness >= 0 && n_angle >= 0, SF_ERR_INVALID,
                   "bad arguments");
        SF_REQUIRE(refractive_index > 0.0, SF_ERR_INVALID, "refractive index must be positive");
        CalcArgs a = calc_base(refractive_index, wavelengths_nm);
        a.mode = 0;
        a.nt = n_thickness;
        a.na = n_angle;
        a.gamma = gamma != 0;
        const size_t cells = static_cast<size_t>(n_thickness) * n_angle;
        if (cells == 0) return;
        // generate_2d_color_lookup_table addressing (interference.rs:497-507).
        const double tstep = (tmax_nm - tmin_nm) / static_cast<double>(std::max(n_thickness - 1, 1));
        const double astep = 1.0 / static_cast<double>(std::max(n_angle - 1, 1));
        std::vector<float> t(n_thickness), c(n_angle);
        for (int i = 0; i < n_thickness; ++i) t[i] = static_cast<float>(tmin_nm + static_cast<double>(i) * tstep);
        for (int j = 0; j < n_angle; ++j) c[j] = static_cast<float>(static_cast<double>(j) * astep);
        run_calc(a, 3 * cells, out, cells, t.data(), t.size(), c.data(), c.size(), nullptr, 0);
    });
}

extern "C" int sf_lut_1d(double refractive_index, const double wavelengths_nm[3], double tmin_nm,
                         double tmax_nm, int32_t n, double cos_theta, float* out) {
    return guard([&] {
#pragma clang fp contract(off)
        SF_REQUIRE(out != nullptr && n >= 0, SF_ERR_INVALID, "bad arguments");
        SF_REQUIRE(refractive_index > 0.0, SF_ERR_INVALID, "refractive index must be positive");
        CalcArgs a = calc_base(refractive_index, wavelengths_nm);
        a.mode = 1;
        a.nt = n;
        if (n == 0) return;
        // generate_color_lookup_table addressing (interference.rs:469-473).
        const double step = (tmax_nm - tmin_nm) / static_cast<double>(n - 1);
        std::vector<float> t(n);
        for (int i = 0; i < n; ++i) t[i] = static_cast<float>(tmin_nm + static_cast<double>(i) * step);
        const float c = static_cast<float>(cos_theta);
        run_calc(a, 3 * static_cast<size_t>(n), out, n, t.data(), t.size(), &c, 1, nullptr, 0);
    });
}

extern "C" int sf_reflected_intensity(double refractive_index, int32_t n, const float* d_nm,
                                      const float* cos_i, const float* lambda_nm, float* out) {
    return guard([&] {
        SF_REQUIRE(n >= 0 && (n == 0 || (d_nm && cos_i && lambda_nm && out)), SF_ERR_INVALID,
                   "bad arguments");
        if (n == 0) return;
        CalcArgs a = calc_base(refractive_index, nullptr);
        a.mode = 3;
        a.count = n;
        run_calc(a, n, out, n, d_nm, n, cos_i, n, lambda_nm, n);
    });
}

extern "C" int sf_interference_color(double refractive_index, const double wavelengths_nm[3],
                                     int32_t n, const float* d_nm, const float* cos_i,
                                     int32_t gamma, float* out) {
    return guard([&] {
        SF_REQUIRE(n >= 0 && (n == 0 || (d_nm && cos_i && out)), SF_ERR_INVALID, "bad arguments");
        if (n == 0) return;
        CalcArgs a = calc_base(refractive_index, wavelengths_nm);
        a.mode = 2;
        a.count = n;
        a.gamma = gamma != 0;
        run_calc(a, 3 * static_cast<size_t>(n), out, n, d_nm, n, cos_i, n, nullptr, 0);
    });
}

extern "C" int sf_interference_terms(double refractive_index, int32_t n, const float* d_nm,
                                     const float* cos_i, const float* lambda_nm, float* out) {
    return guard([&] {
        SF_REQUIRE(n >= 0 && (n == 0 || (d_nm && cos_i && lambda_nm && out)), SF_ERR_INVALID,
                   "bad arguments");
        if (n == 0) return;
        CalcArgs a = calc_base(refractive_index, nullptr);
        a.mode = 4;
        a.count = n;
        run_calc(a, 6 * static_cast<size_t>(n), out, n, d_nm, n, cos_i, n, lambda_nm, n);
    });
}

extern "C" int sf_optical_path(double refractive_index, int32_t n, const float* d_nm,
                               const float* cos_t, const float* lambda_nm, float* out) {
    return guard([&] {
        SF_REQUIRE(n >= 0 && (n == 0 || (d_nm && cos_t && lambda_nm && out)), SF_ERR_INVALID,
                   "bad arguments");
        if (n == 0) return;
        CalcArgs a = calc_base(refractive_index, nullptr);
        a.mode = 5;
        a.count = n;
        run_calc(a, 2 * static_cast<size_t>(n), out, n, d_nm, n, cos_t, n, lambda_nm, n);
    });
}

extern "C" int sf_phase_difference(int32_t n, const float* opd_nm, const float* lambda_nm,
                                   float* out) {
    return guard([&] {
        SF_REQUIRE(n >= 0 && (n == 0 || (opd_nm && lambda_nm && out)), SF_ERR_INVALID,
                   "bad arguments");
        if (n == 0) return;
        CalcArgs a = calc_base(1.0, nullptr);
        a.mode = 6;
        a.count = n;
        run_calc(a, static_cast<size_t>(n), out, n, opd_nm, n, nullptr, 0, lambda_nm, n);
    });
}

// Colour curves on the device: B fragments of spectral_kernel,
// [2 formats (BF16, FP16)][2 (hi, lo)][3 MFMAs][64 lanes] of 16-bit x 8,
// then the slot table of 1 / (2 lambda) (FP64 quotients, rounded).
constexpr size_t kCmfFragsPerFormat = 2 * kSpecMfma * 64;
constexpr size_t kCmfFrags = 2 * kCmfFragsPerFormat;
constexpr size_t kCmfPhaseOffset = kCmfFrags * 4 + 4 * kSpecSlots;   // in floats; 8-byte aligned
constexpr size_t kCmfFloats = kCmfPhaseOffset + 2 * kSpecPhase * 64;

static uint16_t bf16_rne(float x) {
    uint32_t u;
    std::memcpy(&u, &x, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return static_cast<uint16_t>(u >> 16);
}

static float bf16_value(uint16_t h) {
    const uint32_t u = static_cast<uint32_t>(h) << 16;
    float x;
    std::memcpy(&x, &u, 4);
    return x;
}

// Round-to-nearest-even FP16 (the host compiler's conversion) and back.
static uint16_t f16_rne(float x) {
    const _Float16 h = static_cast<_Float16>(x);
    uint16_t u;
    std::memcpy(&u, &h, 2);
    return u;
}

static float f16_value(uint16_t u) {
    _Float16 h;
    std::memcpy(&h, &u, 2);
    return static_cast<float>(h);
}

struct PackedCmf {
    std::vector<u32x4> frag;
    float ulam[4 * kSpecSlots];
    u32x2 pfrag[kSpecPhase * 64];
    float sum[2][3];          // [format][colour]
    int has_lo[2];
    bool f16_bound;           // 2^-18 sum_k |cmf[k][c]| <= kSpecF16Budget for every c
    bool f16_range;           // every |cmf| <= 2^14
    int skip;                 // trailing slots left out (kSpecTailBudget)
};

// Packs -cmf[81][3] into the B-operand layout of spectral_kernel: lane
// 16 q + m of MFMA t holds -cmf[q + 4 (8 t + v)][m] at v = 0..7 (zero past
// slot 20, wavelength 80 or colour 2), as the round-to-nearest hi part and
// the rounded remainder (lo), once in BF16 and once in FP16.  The sums are
// those of the curves as the MFMA sees them, -(hi + lo).
static PackedCmf pack_cmf(const float* cmf) {
    PackedCmf out;
    out.frag.assign(kCmfFrags, u32x4{0u, 0u, 0u, 0u});
    out.has_lo[0] = out.has_lo[1] = 0;
    double sum[2][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
    double sabs[3] = {0.0, 0.0, 0.0};
    double maxabs = 0.0;
    for (int k = 0; k < kSpecLambdas; ++k)
        for (int c = 0; c < 3; ++c) {
            sabs[c] += std::fabs(static_cast<double>(cmf[3 * k + c]));
            maxabs = std::max(maxabs, std::fabs(static_cast<double>(cmf[3 * k + c])));
        }
    out.f16_range = maxabs <= 16384.0;
    out.skip = 0;
    double tail[3] = {0.0, 0.0, 0.0};
    for (int sk = 1; sk <= kSpecMaxSkip; ++sk) {
        const int slot = kSpecSlots - sk;
        bool ok = true;
        for (int c = 0; c < 3; ++c) {
            for (int k = 4 * slot; k < std::min(4 * slot + 4, kSpecLambdas); ++k)
                tail[c] += std::fabs(static_cast<double>(cmf[3 * k + c])) / kSpecLambdas;
            if (tail[c] > kSpecTailBudget) ok = false;
        }
        if (!ok) break;
        out.skip = sk;
    }
    out.f16_bound = true;
    for (int c = 0; c < 3; ++c)
        if (std::ldexp(sabs[c], -18) > kSpecF16Budget) out.f16_bound = false;
    for (int f = 0; f < 2; ++f)
        for (int t = 0; t < kSpecMfma; ++t)
            for (int lane = 0; lane < 64; ++lane) {
                const int m = lane & 15, q = lane >> 4;
                uint16_t hv[8], lv[8];
                for (int v = 0; v < 8; ++v) {
                    const int s = 8 * t + v;
                    const int k = q + 4 * s;
                    const bool used = s < kSpecSlots && k < kSpecLambdas && m < 3;
                    const float x = used ? -cmf[3 * k + m] : 0.f;
                    double hx, lx;
                    if (f == 0) {
                        hv[v] = bf16_rne(x);
                        lv[v] = bf16_rne(x - bf16_value(hv[v]));
                        hx = bf16_value(hv[v]);
                        lx = bf16_value(lv[v]);
                    } else {
                        hv[v] = out.f16_range ? f16_rne(x) : 0;
                        lv[v] = out.f16_range ? f16_rne(x - f16_value(hv[v])) : 0;
                        hx = f16_value(hv[v]);
                        lx = f16_value(lv[v]);
                    }
                    if (lv[v] & 0x7FFFu) out.has_lo[f] = 1;
                    if (used) sum[f][m] -= hx + lx;
                }
                u32x4 h, l;
                for (int p = 0; p < 4; ++p) {
                    h[p] = hv[2 * p] | (static_cast<uint32_t>(hv[2 * p + 1]) << 16);
                    l[p] = lv[2 * p] | (static_cast<uint32_t>(lv[2 * p + 1]) << 16);
                }
                out.frag[f * kCmfFragsPerFormat + t * 64 + lane] = h;
                out.frag[f * kCmfFragsPerFormat + (kSpecMfma + t) * 64 + lane] = l;
            }
    for (int f = 0; f < 2; ++f)
        for (int c = 0; c < 3; ++c) out.sum[f][c] = static_cast<float>(sum[f][c]);
    for (int s = 0; s < kSpecSlots; ++s)
        for (int q = 0; q < 4; ++q) {
            const int k = q + 4 * s;
            out.ulam[4 * s + q] = k < kSpecLambdas ? static_cast<float>(0.5 / (380.0 + 5.0 * k)) : 0.f;
        }
    // Phase-MFMA A fragments: lane l of MFMA e holds row l % 4 of block l / 4,
    // the wavelength k = l / 16 + 4 (4 e + l % 4), as [u0, u1, 2^-11 u0,
    // 2^-11 u1] with U = 512 / lambda ~ u0 + 2^-11 u1 in FP16 parts.
    for (int e = 0; e < kSpecPhase; ++e)
        for (int lane = 0; lane < 64; ++lane) {
            const int k = (lane >> 4) + 4 * (4 * e + (lane & 3));
            const double U = 512.0 / (380.0 + 5.0 * k);
            const double u0 = f16_value(f16_rne(static_cast<float>(U)));
            const double u1 = f16_value(f16_rne(static_cast<float>((U - u0) * 2048.0)));
            const uint16_t hv[4] = {f16_rne(static_cast<float>(u0)), f16_rne(static_cast<float>(u1)),
                                    f16_rne(static_cast<float>(u0 / 2048.0)), f16_rne(static_cast<float>(u1 / 2048.0))};
            out.pfrag[e * 64 + lane] = u32x2{hv[0] | (static_cast<uint32_t>(hv[1]) << 16),
                                             hv[2] | (static_cast<uint32_t>(hv[3]) << 16)};
        }
    return out;
}

namespace sf {
// Per-row (sin, cos theta) and per-column (sin, cos phi) tables of the view
// geometry, built once per context in FP64 and rounded.
static void ensure_view_tables(Drainage* d) {
    if (!d->spec_rows) {
        d->spec_rows = dalloc_i<float4>(d->nt);
        d->spec_cols = dalloc_i<float2>(d->np);
        std::vector<float4> rw(d->nt);
        std::vector<float2> cl(d->np);
        for (int i = 0; i < d->nt; ++i) {
            const double th = static_cast<double>(i) * d->dtheta;
            rw[i] = make_float4(static_cast<float>(std::sin(th)), static_cast<float>(std::cos(th)), 0.f, 0.f);
        }
        for (int j = 0; j < d->np; ++j) {
            const double ph = static_cast<double>(j) * d->dphi;
            cl[j] = make_float2(static_cast<float>(std::sin(ph)), static_cast<float>(std::cos(ph)));
        }
        SF_HIP(hipMemcpyAsync(d->spec_rows, rw.data(), d->nt * sizeof(float4), hipMemcpyHostToDevice, d->stream));
        SF_HIP(hipMemcpyAsync(d->spec_cols, cl.data(), d->np * sizeof(float2), hipMemcpyHostToDevice, d->stream));
        SF_HIP(hipStreamSynchronize(d->stream));
    }
}

// Output plane, view tables and (when given) the colour curves, on the
// context's stream.
static void spectral_prepare(Drainage* d, const float* cmf) {
    if (!d->rgb) d->rgb = dalloc_i<float>(3 * d->n);
    ensure_view_tables(d);
    if (cmf) {
        for (int k = 0; k < 3 * kSpecLambdas; ++k)
            SF_REQUIRE(std::isfinite(cmf[k]), SF_ERR_INVALID, "colour-matching curve value %d is not finite", k);
        if (!d->cmf_dev) d->cmf_dev = dalloc_i<float>(kCmfFloats);
        const PackedCmf packed = pack_cmf(cmf);
        SF_HIP(hipMemcpyAsync(d->cmf_dev, packed.frag.data(), kCmfFrags * sizeof(u32x4), hipMemcpyHostToDevice,
                              d->stream));
        SF_HIP(hipMemcpyAsync(d->cmf_dev + kCmfFrags * 4, packed.ulam, sizeof(packed.ulam), hipMemcpyHostToDevice,
                              d->stream));
        SF_HIP(hipMemcpyAsync(d->cmf_dev + kCmfPhaseOffset, packed.pfrag, sizeof(packed.pfrag),
                              hipMemcpyHostToDevice, d->stream));
        SF_HIP(hipStreamSynchronize(d->stream));
        for (int f = 0; f < 2; ++f) {
            for (int c = 0; c < 3; ++c) d->cmf_sum[f][c] = packed.sum[f][c];
            d->cmf_lo[f] = packed.has_lo[f];
        }
        d->cmf_f16_bound = packed.f16_bound;
        d->cmf_f16_range = packed.f16_range;
        d->cmf_skip = packed.skip;
    }
}

// Operand format of the next spectral pass: 1 (FP16) or 0 (BF16 hi/lo pair).
static int spectral_format(const Drainage* d) {
    if (d->spec_operand == SF_SPECTRAL_OPERAND_BF16_PAIR) return 0;
    if (d->spec_operand == SF_SPECTRAL_OPERAND_FP16) {
        SF_REQUIRE(d->cmf_f16_range, SF_ERR_INVALID, "colour-matching curves exceed the FP16 operand range");
        return 1;
    }
    return (d->cmf_f16_range && d->cmf_f16_bound) ? 1 : 0;
}

// Launches spectral_kernel on stream st over the context's current state
// (lazy pole rows are taken from the ring partials).
void spectral_launch(Drainage* d, const sf_camera* cam, double n_film, const float* cmf, hipStream_t st,
                     const FieldView* view, float* rgb_out) {
    spectral_prepare(d, cmf);
    SF_REQUIRE(d->cmf_dev != nullptr, SF_ERR_INVALID, "no colour-matching curves given yet");
    SpecArgs a;
    a.hi = view ? view->hi : d->hi[d->cur_hi];
    a.lo = view ? view->lo : d->lo[d->cur];
    a.rows = d->spec_rows;
    a.cols = d->spec_cols;
    const int f = spectral_format(d);
    a.bfrag = reinterpret_cast<const u32x4*>(d->cmf_dev) + f * kCmfFragsPerFormat;
    a.ulam = d->cmf_dev + kCmfFrags * 4;
    a.pfrag = reinterpret_cast<const u32x2*>(d->cmf_dev + kCmfPhaseOffset);
    a.s0 = d->cmf_sum[f][0] / kSpecLambdas;
    a.s1 = d->cmf_sum[f][1] / kSpecLambdas;
    a.s2 = d->cmf_sum[f][2] / kSpecLambdas;
    a.b_lo = d->cmf_lo[f];
    a.hpole = view ? view->hpole : d->hpole[d->cur_pole];
    a.pole_virtual = view ? (view->pole_virtual ? 1 : 0)
                          : ((d->poles_virtual && !(d->p.flags & SF_FLAG_POLES_COPY_H)) ? 1 : 0);
    a.nt = d->nt;
    a.ntiles = d->ntiles;
    a.rgb = rgb_out ? rgb_out : d->rgb;
    a.np = d->np;
    a.n = d->n;
    const sf_camera c = cam ? *cam : sf_camera{{0.0, 0.0, 4.0 * d->p.bubble_radius}};
    a.cx = static_cast<float>(c.position[0]);
    a.cy = static_cast<float>(c.position[1]);
    a.cz = static_cast<float>(c.position[2]);
    a.cam2 = static_cast<float>(c.position[0] * c.position[0] + c.position[1] * c.position[1] +
                                c.position[2] * c.position[2]);
    a.radius = static_cast<float>(d->p.bubble_radius);
    a.n_film = static_cast<float>(n_film);
    a.inv_n_film = static_cast<float>(1.0 / n_film);
    a.view = nullptr;
    // View terms of this camera and film index, cached on the context.
    const float key[6] = {a.cx, a.cy, a.cz, a.radius, a.n_film, a.inv_n_film};
    if (!d->spec_view) d->spec_view = dalloc_i<float2>(d->n);
    if (!d->spec_view_valid || std::memcmp(key, d->spec_view_key, sizeof(key)) != 0) {
        hipLaunchKernelGGL(spectral_view_kernel, dim3(cdiv_u(d->n, 256)), dim3(256), 0, st, a, d->spec_view);
        SF_HIP(hipGetLastError());
        std::memcpy(d->spec_view_key, key, sizeof(key));
        d->spec_view_valid = true;
    }
    a.view = d->spec_view;
    const size_t supergroups = (d->n + 63) / 64;
    // About two 64-cell groups per wave: the short blocks balance the tail
    // (145.8 us at 2048 x 4096, against 154.5 us with 2560 grid-striding
    // blocks and 149.5 us with one group per wave).
    const unsigned blocks = static_cast<unsigned>(std::max<size_t>(1, (supergroups + 7) / 8));
    static void (*const kernels[2][kSpecMaxSkip + 1])(const SpecArgs) = {
        {spectral_kernel<false, 0>, spectral_kernel<false, 1>, spectral_kernel<false, 2>},
        {spectral_kernel<true, 0>, spectral_kernel<true, 1>, spectral_kernel<true, 2>}};
    void (*kern)(const SpecArgs) = kernels[f][d->cmf_skip];
    if (d->spec_timer.enabled) {
        auto& ev = d->spec_timer.next(st);
        hipExtLaunchKernelGGL(kern, dim3(blocks), dim3(256), 0, st, ev.first, ev.second, 0, a);
    } else {
        hipLaunchKernelGGL(kern, dim3(blocks), dim3(256), 0, st, a);
    }
    SF_HIP(hipGetLastError());
}
}  // namespace sf

extern "C" int sf_spectral_rgb(sf_drainage* ctx, const sf_camera* cam, double refractive_index,
                               const float* cmf, float* rgb_host) {
    return guard([&] {
        SF_REQUIRE(ctx != nullptr, SF_ERR_INVALID, "ctx is NULL");
        SF_REQUIRE(refractive_index > 0.0, SF_ERR_INVALID, "refractive index must be positive");
        Drainage* d = reinterpret_cast<Drainage*>(ctx);
        spectral_launch(d, cam, refractive_index, cmf, d->stream);
        if (rgb_host) {
            SF_HIP(hipMemcpyAsync(rgb_host, d->rgb, 3 * d->n * sizeof(float), hipMemcpyDeviceToHost, d->stream));
            SF_HIP(hipStreamSynchronize(d->stream));
        }
    });
}

// n_steps of {drainage step, spectral RGB of the new state}; frame s goes to
// frames + 3 n s when frames (device) is given, else to the context's planes.
static void step_spectral(Drainage* d, int32_t n_steps, double dt, const sf_camera* cam, double refractive_index,
                          const float* cmf, float* frames) {
    SF_REQUIRE(n_steps >= 0, SF_ERR_INVALID, "bad arguments");
    SF_REQUIRE(refractive_index > 0.0, SF_ERR_INVALID, "refractive index must be positive");
    SF_REQUIRE(std::isfinite(dt), SF_ERR_INVALID, "dt must be finite");
    if (n_steps > 0) dt = d->resolve_dt(dt);
    if (n_steps > 0) {
        spectral_prepare(d, cmf);
        // Curves and operand format are checked before any step is queued.
        SF_REQUIRE(d->cmf_dev != nullptr, SF_ERR_INVALID, "no colour-matching curves given yet");
        (void)spectral_format(d);
        if (d->tables_dirty) d->upload_tables();
    }
    auto frame = [&](int s) { return frames ? frames + 3 * d->n * static_cast<size_t>(s) : nullptr; };
    // The pole rows stay lazy (ring partials): the spectral kernel takes
    // their ring mean itself, and the next step reads the partials.  Where
    // two steps go into one stencil launch, that launch also writes the
    // state between them, and both states are rendered.
    for (int s = 0; s < n_steps;) {
        if (d->two_step_next(n_steps - s)) {
            d->launch_two(dt, true);
            const FieldView mid{d->mid_hi, d->lo_mid, d->mid_hpole, !(d->p.flags & SF_FLAG_POLES_COPY_H)};
            spectral_launch(d, cam, refractive_index, nullptr, d->stream, &mid, frame(s));
            spectral_launch(d, cam, refractive_index, nullptr, d->stream, nullptr, frame(s + 1));
            s += 2;
        } else {
            d->launch_one(dt);
            spectral_launch(d, cam, refractive_index, nullptr, d->stream, nullptr, frame(s));
            s += 1;
        }
    }
    d->finalize_diag();
}

extern "C" int sf_drainage_step_spectral(sf_drainage* ctx, int32_t n_steps, double dt,
                                         const sf_camera* cam, double refractive_index,
                                         const float* cmf) {
    return guard([&] {
        SF_REQUIRE(ctx != nullptr, SF_ERR_INVALID, "ctx is NULL");
        step_spectral(reinterpret_cast<Drainage*>(ctx), n_steps, dt, cam, refractive_index, cmf, nullptr);
    });
}

extern "C" int sf_drainage_step_spectral_frames(sf_drainage* ctx, int32_t n_steps, double dt,
                                                const sf_camera* cam, double refractive_index,
                                                const float* cmf, float* frames_dev) {
    return guard([&] {
        SF_REQUIRE(ctx != nullptr && frames_dev != nullptr, SF_ERR_INVALID, "NULL argument");
        step_spectral(reinterpret_cast<Drainage*>(ctx), n_steps, dt, cam, refractive_index, cmf, frames_dev);
    });
}

extern "C" int sf_spectral_device_rgb(sf_drainage* ctx, const float** rgb) {
    return guard([&] {
        SF_REQUIRE(ctx != nullptr && rgb != nullptr, SF_ERR_INVALID, "NULL argument");
        Drainage* d = reinterpret_cast<Drainage*>(ctx);
        SF_REQUIRE(d->rgb != nullptr, SF_ERR_INVALID, "sf_spectral_rgb has not run on this context");
        *rgb = d->rgb;
    });
}

extern "C" int sf_spectral_copy_rgb(sf_drainage* ctx, float* rgb_host) {
    return guard([&] {
        SF_REQUIRE(ctx != nullptr && rgb_host != nullptr, SF_ERR_INVALID, "NULL argument");
        Drainage* d = reinterpret_cast<Drainage*>(ctx);
        SF_REQUIRE(d->rgb != nullptr, SF_ERR_INVALID, "no spectral RGB has been computed on this context");
        SF_HIP(hipMemcpyAsync(rgb_host, d->rgb, 3 * d->n * sizeof(float), hipMemcpyDeviceToHost, d->stream));
        SF_HIP(hipStreamSynchronize(d->stream));
    });
}

// ---------------------------------------------------------------------------
// Rendering of the thickness field through the runtime LUT (the fragment
// path of bubble.wgsl:425-439 evaluated on the simulation grid): per cell the
// view cosine |dot(normalize(cam - p), n)| (bubble.wgsl:491-494), texture
// coordinates u = clamp(t_nm / 2000, 0, 1), v = clamp(cos, 0, 1), bilinear
// sampling of the 256 x 64 RGBA8 LUT with clamp-to-edge addressing
// (pipeline.rs:350-359), times the interference intensity, clamped to [0, 1]
// and written as RGBA8 (alpha 255).

struct LutRenderArgs {
    const float* hi;
    const float* lo;
    const float4* rows;
    const float2* cols;
    const uint8_t* lut;       // [64][256][4]
    uint8_t* out;             // [n][4]
    int np;
    size_t n;
    float cx, cy, cz, cam2, radius, intensity;
};

__device__ inline uchar4 lut_texel4(const uint8_t* lut, int i, int j) {
    return reinterpret_cast<const uchar4*>(lut)[j * kLutT + i];      // 32-bit offset from the LUT base
}

__device__ inline float lut_chan(uchar4 t, int c) {
    const uint8_t b = c == 0 ? t.x : (c == 1 ? t.y : t.z);
    return static_cast<float>(b) * (1.f / 255.f);
}

// One pixel per thread on a (column block, row) grid; each of the four
// bilinear texels is one 32-bit load.
__global__ __launch_bounds__(256) void lut_render_kernel(const LutRenderArgs a) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y;
    if (j >= a.np) return;
    // row bases are uniform in the block: the per-lane offset stays 32-bit
    const size_t row = static_cast<size_t>(i) * a.np;
    const float* rh = a.hi + row;
    const float* rl = a.lo + row;
    const float4 rw = a.rows[i];
    const float2 cl = a.cols[j];
    const float nx = rw.x * cl.y, ny = rw.y, nz = rw.x * cl.x;
    const float cn = a.cx * nx + a.cy * ny + a.cz * nz;
    const float len2 = a.cam2 - 2.f * a.radius * cn + a.radius * a.radius;
    const float cos_v = fminf(fabsf(cn - a.radius) * rsqrtf(len2), 1.f);
    const float t_nm = (rh[j] + rl[j]) * 1e9f;
    const float u = fminf(fmaxf(t_nm / kLutMaxNm, 0.f), 1.f);
    const float v = fminf(fmaxf(cos_v, 0.f), 1.f);
    const float x = u * static_cast<float>(kLutT) - 0.5f, y = v * static_cast<float>(kLutA) - 0.5f;
    const float x0 = floorf(x), y0 = floorf(y);
    const float fx = x - x0, fy = y - y0;
    const int i0 = min(max(static_cast<int>(x0), 0), kLutT - 1), i1 = min(max(static_cast<int>(x0) + 1, 0), kLutT - 1);
    const int j0 = min(max(static_cast<int>(y0), 0), kLutA - 1), j1 = min(max(static_cast<int>(y0) + 1, 0), kLutA - 1);
    const uchar4 t00 = lut_texel4(a.lut, i0, j0), t10 = lut_texel4(a.lut, i1, j0);
    const uchar4 t01 = lut_texel4(a.lut, i0, j1), t11 = lut_texel4(a.lut, i1, j1);
    uchar4 px;
    uint8_t ch[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float top = lut_chan(t00, c) * (1.f - fx) + lut_chan(t10, c) * fx;
        const float bot = lut_chan(t01, c) * (1.f - fx) + lut_chan(t11, c) * fx;
        const float val = fminf(fmaxf((top * (1.f - fy) + bot * fy) * a.intensity, 0.f), 1.f);
        ch[c] = static_cast<uint8_t>(val * 255.f + 0.5f);
    }
    px.x = ch[0];
    px.y = ch[1];
    px.z = ch[2];
    px.w = 255;
    (reinterpret_cast<uchar4*>(a.out) + row)[j] = px;
}

extern "C" int sf_render_lut_rgba8(sf_drainage* ctx, const uint8_t* lut_rgba8, double refractive_index,
                                   float intensity, const sf_camera* cam, uint8_t* rgba_out) {
    return guard([&] {
        SF_REQUIRE(ctx != nullptr && rgba_out != nullptr, SF_ERR_INVALID, "NULL argument");
        SF_REQUIRE(std::isfinite(intensity) && refractive_index > 0.0, SF_ERR_INVALID, "bad intensity or index");
        Drainage* d = reinterpret_cast<Drainage*>(ctx);
        ensure_view_tables(d);
        d->materialize();
        DevBuf lut(static_cast<size_t>(kLutT) * kLutA * 4);
        if (lut_rgba8) {
            SF_HIP(hipMemcpyAsync(lut.p, lut_rgba8, static_cast<size_t>(kLutT) * kLutA * 4, hipMemcpyHostToDevice,
                                  d->stream));
        } else {   // the renderer's LUT: built at intensity 1, scaled at sampling (pipeline.rs:311-314)
            hipLaunchKernelGGL(lut_rgb_kernel, dim3(kLutT / 64, kLutA), dim3(64), 0, d->stream,
                               static_cast<float>(refractive_index), 1.f, static_cast<float*>(nullptr),
                               static_cast<uint8_t*>(lut.p));
            SF_HIP(hipGetLastError());
        }
        DevBuf out(d->n * 4);
        LutRenderArgs a;
        a.hi = d->hi[d->cur_hi];
        a.lo = d->lo[d->cur];
        a.rows = d->spec_rows;
        a.cols = d->spec_cols;
        a.lut = static_cast<const uint8_t*>(lut.p);
        a.out = static_cast<uint8_t*>(out.p);
        a.np = d->np;
        a.n = d->n;
        const sf_camera c = cam ? *cam : sf_camera{{0.0, 0.0, 4.0 * d->p.bubble_radius}};
        a.cx = static_cast<float>(c.position[0]);
        a.cy = static_cast<float>(c.position[1]);
        a.cz = static_cast<float>(c.position[2]);
        a.cam2 = static_cast<float>(c.position[0] * c.position[0] + c.position[1] * c.position[1] +
                                    c.position[2] * c.position[2]);
        a.radius = static_cast<float>(d->p.bubble_radius);
        a.intensity = intensity;
        hipLaunchKernelGGL(lut_render_kernel, dim3(static_cast<unsigned>((d->np + 255) / 256), d->nt), dim3(256),
                           0, d->stream, a);
        SF_HIP(hipGetLastError());
        SF_HIP(hipMemcpyAsync(rgba_out, out.p, d->n * 4, hipMemcpyDeviceToHost, d->stream));
        SF_HIP(hipStreamSynchronize(d->stream));
    });
}
