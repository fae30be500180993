// fno_kernels.hpp -- device kernels of the Fourier neural operator surrogate.
//
// Tensors are FP32 in HBM, channel-last: activations [B][n_theta][n_phi][C].
// Every contraction (truncated DFTs, inverse DFTs, pointwise channel mixing,
// and the transposed products of the backward pass) runs through one
// batched GEMM on BF16 MFMA (v_mfma_f32_16x16x16_bf16) with FP32
// accumulation; operands are rounded to BF16 (nearest even) as they are
// staged into LDS.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "sf_common.hpp"

namespace sf {
namespace fno {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// Network input channels (h, surfactant and three geometry/parameter
// features) and output channels (the two predicted increments).
constexpr int kLiftIn = 8;
constexpr int kProjOut = 2;

__device__ inline unsigned short to_bf16(float x) {
    uint32_t u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);   // round to nearest even (inputs are finite)
    return static_cast<unsigned short>(u >> 16);
}

// C[b] = alpha * op(A[b]) * op(B[b]) + beta * C[b], row-major storage;
// op(A)[m][k] = ta ? A[k * lda + m] : A[m * lda + k], likewise for B.
struct GemmArgs {
    const float* A;
    const float* B;
    float* C;
    int M, N, K;
    int lda, ldb, ldc;
    long long sA, sB, sC;   // batch strides in elements
    int ta, tb;
    float alpha, beta;
};

constexpr int kTM = 64, kTN = 64, kTK = 16, kPad = 4;

// Block: 4 waves, 64 x 64 tile of C; wave w owns the 32 x 32 quadrant
// (w / 2, w % 2) as 2 x 2 MFMA tiles of 16 x 16.  Grid: (M tiles, N tiles,
// batch).  MFMA lane layout (16x16x16, bf16): lane l supplies A[l % 16][4 (l / 16) + v]
// and B[4 (l / 16) + v][l % 16] and receives D[4 (l / 16) + v][l % 16].
__global__ __launch_bounds__(256) void gemm_kernel(const GemmArgs g) {
    __shared__ unsigned short As[kTM][kTK + kPad];   // [m][k]
    __shared__ unsigned short Bs[kTN][kTK + kPad];   // [n][k]
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int m0 = blockIdx.x * kTM, n0 = blockIdx.y * kTN;
    const long long bz = blockIdx.z;
    const float* A = g.A + bz * g.sA;
    const float* B = g.B + bz * g.sB;
    float* C = g.C + bz * g.sC;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    f32x4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

    for (int k0 = 0; k0 < g.K; k0 += kTK) {
        if (!g.ta) {
            const int r = tid >> 2, kq = (tid & 3) * 4, m = m0 + r;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int k = k0 + kq + j;
                const float v = (m < g.M && k < g.K) ? A[static_cast<long long>(m) * g.lda + k] : 0.f;
                As[r][kq + j] = to_bf16(v);
            }
        } else {
            const int kr = tid >> 4, mq = (tid & 15) * 4, k = k0 + kr;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int m = m0 + mq + j;
                const float v = (m < g.M && k < g.K) ? A[static_cast<long long>(k) * g.lda + m] : 0.f;
                As[mq + j][kr] = to_bf16(v);
            }
        }
        if (!g.tb) {
            const int kr = tid >> 4, nq = (tid & 15) * 4, k = k0 + kr;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int n = n0 + nq + j;
                const float v = (n < g.N && k < g.K) ? B[static_cast<long long>(k) * g.ldb + n] : 0.f;
                Bs[nq + j][kr] = to_bf16(v);
            }
        } else {
            const int r = tid >> 2, kq = (tid & 3) * 4, n = n0 + r;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int k = k0 + kq + j;
                const float v = (n < g.N && k < g.K) ? B[static_cast<long long>(n) * g.ldb + k] : 0.f;
                Bs[r][kq + j] = to_bf16(v);
            }
        }
        __syncthreads();
        const int kl = 4 * (lane >> 4);
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const bf16x4 a = *reinterpret_cast<const bf16x4*>(&As[wm + i * 16 + (lane & 15)][kl]);
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                const bf16x4 b = *reinterpret_cast<const bf16x4*>(&Bs[wn + j * 16 + (lane & 15)][kl]);
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[i][j], 0, 0, 0);
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int m = m0 + wm + i * 16 + 4 * (lane >> 4) + v;
                const int n = n0 + wn + j * 16 + (lane & 15);
                if (m < g.M && n < g.N) {
                    float r = g.alpha * acc[i][j][v];
                    const long long o = static_cast<long long>(m) * g.ldc + n;
                    if (g.beta != 0.f) r += g.beta * C[o];
                    C[o] = r;
                }
            }
}

// ---- elementwise -------------------------------------------------------------

__device__ inline float gelu(float z) {      // tanh form
    const float k0 = 0.7978845608028654f, k1 = 0.044715f;
    return 0.5f * z * (1.f + tanhf(k0 * (z + k1 * z * z * z)));
}

// ---- fused Fourier layer, one block per (bubble, theta row) ------------------
//
// For one grid row t of bubble b (P cells x C channels):
//   z  = x W + Iphi Z[b][t] + bias      (pointwise mixing + inverse phi DFT,
//                                        one MFMA product with K = C + J)
//   x' = act(z)                          (stored; z stored when training)
//   X1'[b][t] = Ephi x'                  (phi DFT of the next layer, fed from
//                                        the accumulators: the MFMA result
//                                        layout D[4(l/16)+v][l%16] is the
//                                        B-operand layout B[4(l/16)+v][l%16])
// Inference options: with x16 set, x and x_out hold BF16 activations (the
// values the MFMA reads; a float is rounded the same way on load, so this
// changes no result); with `in` set, x is not read but lifted on the fly from
// the kLiftIn input features (x = b_l + in W_l, FP32); with `proj` set, the
// projection proj[p][j] = q[j] + sum_o x'[p][o] Q[o][j] is formed from the
// accumulators (FP32) and x' is stored only if x_out is set.
// Block: WAVES waves, wave w owns rows p in [w P/WAVES, (w+1) P/WAVES).
// RT = P / (16 WAVES) row tiles per wave, CT = C / 16 channel tiles, JT =
// J / 16 tiles of the 2 Mp real/imaginary phi modes.
struct RowArgs {
    const float* x;       // [rows][P][C]
    const float* Z;       // [rows][J][C]
    const float* W;       // [C][C]
    const float* bias;    // [C]
    const unsigned short* Iphi;   // [P][J] BF16
    const unsigned short* Ephi;   // [J][P] BF16
    float* x_out;         // [rows][P][C]
    float* z_out;         // [rows][P][C] or null
    float* X1_out;        // [rows][J][C] or null
    int act;
    const float* in;      // [rows][P][kLiftIn] or null: lift x from these
    const float* Wl;      // [kLiftIn][C] lift weights
    const float* bl;      // [C] lift bias
    const float* Q;       // [C][kProjOut] projection weights
    const float* qb;      // [kProjOut] projection bias
    float* proj;          // [rows][P][kProjOut] or null
    int x16;              // x and x_out hold BF16 values (inference)
};

template <int RT, int CT, int JT, int WAVES>
__global__ __launch_bounds__(64 * WAVES) void fno_row_kernel(const RowArgs a) {
    constexpr int C = 16 * CT, J = 16 * JT, K = C + J, P = 16 * RT * WAVES, NT = 64 * WAVES;
    __shared__ unsigned short Bs[C][K + 4];       // [out channel][k]
    __shared__ float red[WAVES][J][C];
    __shared__ __align__(16) float s_lift[(kLiftIn + 1) * C];   // W_l rows, then b_l
    const long long row = blockIdx.x;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int mr = lane & 15, kq = 4 * (lane >> 4);
    for (int q = threadIdx.x; q < K * C; q += NT) {
        const int k = q / C, n = q % C;
        const float v = k < C ? a.W[k * C + n] : a.Z[row * J * C + static_cast<long long>(k - C) * C + n];
        Bs[n][k] = to_bf16(v);
    }
    const bool lift = a.in != nullptr;
    if (lift)
        for (int q = threadIdx.x; q < (kLiftIn + 1) * C; q += NT)
            s_lift[q] = q < kLiftIn * C ? a.Wl[q] : a.bl[q - kLiftIn * C];
    __syncthreads();
    const float* x = a.x + row * P * C;
    f32x4 acc[RT][CT];
#pragma unroll
    for (int rt = 0; rt < RT; ++rt)
#pragma unroll
        for (int ct = 0; ct < CT; ++ct) acc[rt][ct] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int rt = 0; rt < RT; ++rt) {
        const int p = wave * RT * 16 + rt * 16 + mr;
        float fin[kLiftIn];
        if (lift) {
            const float* ip = a.in + (row * P + p) * kLiftIn;
#pragma unroll
            for (int i = 0; i < kLiftIn; ++i) fin[i] = ip[i];
        }
#pragma unroll
        for (int ks = 0; ks < CT; ++ks) {
            bf16x4 A;
            if (!lift && a.x16) {
                A = *reinterpret_cast<const bf16x4*>(reinterpret_cast<const unsigned short*>(a.x) + row * P * C +
                                                     static_cast<long long>(p) * C + ks * 16 + kq);
            } else {
                float4 v;
                if (lift) {
                    v = *reinterpret_cast<const float4*>(&s_lift[kLiftIn * C + ks * 16 + kq]);
#pragma unroll
                    for (int i = 0; i < kLiftIn; ++i) {
                        const float4 wv = *reinterpret_cast<const float4*>(&s_lift[i * C + ks * 16 + kq]);
                        v.x = fmaf(fin[i], wv.x, v.x);
                        v.y = fmaf(fin[i], wv.y, v.y);
                        v.z = fmaf(fin[i], wv.z, v.z);
                        v.w = fmaf(fin[i], wv.w, v.w);
                    }
                } else {
                    v = *reinterpret_cast<const float4*>(x + static_cast<long long>(p) * C + ks * 16 + kq);
                }
                A = bf16x4{static_cast<short>(to_bf16(v.x)), static_cast<short>(to_bf16(v.y)),
                           static_cast<short>(to_bf16(v.z)), static_cast<short>(to_bf16(v.w))};
            }
#pragma unroll
            for (int ct = 0; ct < CT; ++ct) {
                const bf16x4 B = *reinterpret_cast<const bf16x4*>(&Bs[ct * 16 + mr][ks * 16 + kq]);
                acc[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(A, B, acc[rt][ct], 0, 0, 0);
            }
        }
#pragma unroll
        for (int js = 0; js < JT; ++js) {
            const bf16x4 A = *reinterpret_cast<const bf16x4*>(a.Iphi + static_cast<long long>(p) * J + js * 16 + kq);
#pragma unroll
            for (int ct = 0; ct < CT; ++ct) {
                const bf16x4 B = *reinterpret_cast<const bf16x4*>(&Bs[ct * 16 + mr][C + js * 16 + kq]);
                acc[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(A, B, acc[rt][ct], 0, 0, 0);
            }
        }
    }
    // epilogue: bias, activation, stores
    float* xo = (a.x_out && !a.x16) ? a.x_out + row * P * C : nullptr;
    unsigned short* xo16 = (a.x_out && a.x16) ? reinterpret_cast<unsigned short*>(a.x_out) + row * P * C : nullptr;
    float* zo = a.z_out ? a.z_out + row * P * C : nullptr;
#pragma unroll
    for (int rt = 0; rt < RT; ++rt)
#pragma unroll
        for (int ct = 0; ct < CT; ++ct) {
            const int o = ct * 16 + mr;
            const float bo = a.bias[o];
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int p = wave * RT * 16 + rt * 16 + kq + v;
                const float z = acc[rt][ct][v] + bo;
                const float xv = a.act ? gelu(z) : z;
                if (xo) xo[static_cast<long long>(p) * C + o] = xv;
                if (xo16) xo16[static_cast<long long>(p) * C + o] = to_bf16(xv);
                if (zo) zo[static_cast<long long>(p) * C + o] = z;
                acc[rt][ct][v] = xv;
            }
        }
    if (a.proj) {
        // Channel o = 16 ct + mr of rows p = base + kq + v sits in lane
        // (mr, kq): per-lane partial sums over ct, then a sum over the 16
        // lanes of the kq group; lane mr = 2 v + j stores output j of row v.
        float qv[CT][kProjOut];
#pragma unroll
        for (int ct = 0; ct < CT; ++ct)
#pragma unroll
            for (int j = 0; j < kProjOut; ++j) qv[ct][j] = a.Q[(ct * 16 + mr) * kProjOut + j];
        float* po = a.proj + row * P * kProjOut;
#pragma unroll
        for (int rt = 0; rt < RT; ++rt) {
            float mine = 0.f;
#pragma unroll
            for (int v = 0; v < 4; ++v)
#pragma unroll
                for (int j = 0; j < kProjOut; ++j) {
                    float s = 0.f;
#pragma unroll
                    for (int ct = 0; ct < CT; ++ct) s = fmaf(acc[rt][ct][v], qv[ct][j], s);
                    s += __shfl_xor(s, 8, 16);
                    s += __shfl_xor(s, 4, 16);
                    s += __shfl_xor(s, 2, 16);
                    s += __shfl_xor(s, 1, 16);
                    if (mr == kProjOut * v + j) mine = s + a.qb[j];
                }
            if (mr < 4 * kProjOut) {
                const int p = wave * RT * 16 + rt * 16 + kq + mr / kProjOut;
                po[static_cast<long long>(p) * kProjOut + mr % kProjOut] = mine;
            }
        }
    }
    if (a.X1_out == nullptr) return;
    // phi DFT of the new activations: X1'[j][o] = sum_p Ephi[j][p] x'[p][o]
    f32x4 acc2[JT][CT];
#pragma unroll
    for (int jt = 0; jt < JT; ++jt)
#pragma unroll
        for (int ct = 0; ct < CT; ++ct) acc2[jt][ct] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int rt = 0; rt < RT; ++rt) {
        const int p0 = wave * RT * 16 + rt * 16;
        bf16x4 Bx[CT];
#pragma unroll
        for (int ct = 0; ct < CT; ++ct)
            Bx[ct] = bf16x4{static_cast<short>(to_bf16(acc[rt][ct][0])), static_cast<short>(to_bf16(acc[rt][ct][1])),
                            static_cast<short>(to_bf16(acc[rt][ct][2])), static_cast<short>(to_bf16(acc[rt][ct][3]))};
#pragma unroll
        for (int jt = 0; jt < JT; ++jt) {
            const bf16x4 A = *reinterpret_cast<const bf16x4*>(a.Ephi + static_cast<long long>(jt * 16 + mr) * P + p0 + kq);
#pragma unroll
            for (int ct = 0; ct < CT; ++ct)
                acc2[jt][ct] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(A, Bx[ct], acc2[jt][ct], 0, 0, 0);
        }
    }
#pragma unroll
    for (int jt = 0; jt < JT; ++jt)
#pragma unroll
        for (int ct = 0; ct < CT; ++ct)
#pragma unroll
            for (int v = 0; v < 4; ++v) red[wave][jt * 16 + kq + v][ct * 16 + mr] = acc2[jt][ct][v];
    __syncthreads();
    float* X1 = a.X1_out + row * J * C;
    for (int q = threadIdx.x; q < J * C; q += NT) {
        const int j = q / C, o = q % C;
        float sum = 0.f;
#pragma unroll
        for (int w = 0; w < WAVES; ++w) sum += red[w][j][o];
        X1[q] = sum;
    }
}
// ---- inference layer, one wave per (bubble, theta row) ----------------------
//
// The inference form of fno_row_kernel (BF16 activations, no saved
// pre-activations).  A wave owns a whole grid row and walks its P cells in
// 16-cell tiles:
//   z  = [x | Iphi] [W ; Z[b][t]] + bias    on v_mfma_f32_16x16x32_bf16 (K = C + J
//                                           in 32-wide steps; lane (m, g) supplies
//                                           A[m][8 g .. 8 g + 7], B[8 g .. 8 g + 7][m])
//   x' = act(z)
//   X1'[b][t] += Ephi[:, tile] x'[tile]     on v_mfma_f32_16x16x16_bf16, x' taken
//                                           straight from the first product's
//                                           result layout D[4(l/16)+v][l%16],
//                                           which is that product's B layout
// so the next layer's phi DFT accumulates inside the wave: no cross-wave
// reduction, no barrier after the set-up.  The B operand of the first
// product (W stacked on the row's inverse-DFT input Z) is held in registers
// for the whole row.  x' tiles go out through a per-wave LDS transpose as
// 16-byte stores.  Layer 0 lifts x from the kLiftIn input features on the
// fly (LIFT: with x = in W_l + b_l, x W = in (W_l W) + b_l W, so it runs on
// the folded weights of lift_fold_kernel and its A operand reads the
// features); the last layer forms the projection instead of the next phi
// DFT (PROJ) and stores x' only when x_out is set.
// Block: 4 waves = 4 rows.  Needs P % 16 == 0 and (C + J) % 32 == 0.
// Layer-0 weights with the linear lift folded in, for the LIFT form of
// fno_row_wave_kernel: Wf[k][n] = sum_c W_l[k][c] W[c][n] for k < kLiftIn
// (rows kLiftIn .. C - 1 zero), bf[n] = b[n] + sum_c b_l[c] W[c][n]; FP32.
__global__ __launch_bounds__(256) void lift_fold_kernel(const float* __restrict__ Wl, const float* __restrict__ bl,
                                                        const float* __restrict__ W, const float* __restrict__ b,
                                                        float* __restrict__ Wf, float* __restrict__ bf, int C) {
    for (int t = threadIdx.x; t < (C + 1) * C; t += blockDim.x) {
        const int k = t / C, n = t - k * C;
        if (k < C) {
            float v = 0.f;
            if (k < kLiftIn)
                for (int c = 0; c < C; ++c) v = fmaf(Wl[k * C + c], W[c * C + n], v);
            Wf[t] = v;
        } else {
            float v = b[n];
            for (int c = 0; c < C; ++c) v = fmaf(bl[c], W[c * C + n], v);
            bf[n] = v;
        }
    }
}

__device__ inline float gelu_exp(float z) {
    // 0.5 z (1 + tanh u) = z / (1 + exp(-2u)), u = k0 (z + k1 z^3); hardware
    // exp2 and reciprocal (about 1e-6 relative, far below the BF16 rounding
    // the result is stored with)
    const float k0 = 0.7978845608028654f, k1 = 0.044715f;
    const float u = k0 * fmaf(k1 * z, z * z, z);
    const float e = __builtin_amdgcn_exp2f(-2.8853900817779268f * u);
    return z * __builtin_amdgcn_rcpf(1.f + e);
}

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

template <int CT, int JT, bool LIFT, bool PROJ>
__global__ __launch_bounds__(256) void fno_row_wave_kernel(const RowArgs a, int P, long long rows) {
    constexpr int C = 16 * CT, J = 16 * JT, K = C + J, KS = K / 32;
    static_assert(K % 32 == 0, "K = C + J must be a multiple of 32");
    constexpr int TRW = C + 8;                           // padded row of the transpose tile (bf16)
    __shared__ __align__(16) unsigned short tr[4][16 * TRW];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int mr = lane & 15, g = lane >> 4;
    constexpr bool lift = LIFT;
    const long long row = blockIdx.x * 4LL + wave;
    if (row >= rows) return;                  // no block barrier: waves are independent

    // B fragments: lane (n = mr, g) holds B[32 s + 8 g + i][16 ct + mr], i < 8.
    // In layer 0 (LIFT) W and bias are the lift-folded weights of
    // lift_fold_kernel: rows k < kLiftIn of W multiply the features.
    static_assert(!LIFT || kLiftIn == 8, "the lifted features fill one 8-wide lane group");
    bf16x8 Bf[KS][CT];
#pragma unroll
    for (int st = 0; st < KS; ++st)
#pragma unroll
        for (int ct = 0; ct < CT; ++ct) {
            u32x4 pk;
#pragma unroll
            for (int h = 0; h < 4; ++h) {
                unsigned short e2[2];
#pragma unroll
                for (int u = 0; u < 2; ++u) {
                    const int k = 32 * st + 8 * g + 2 * h + u, n = 16 * ct + mr;
                    const float v = k < C ? a.W[k * C + n] : a.Z[row * J * C + static_cast<long long>(k - C) * C + n];
                    e2[u] = to_bf16(v);
                }
                pk[h] = static_cast<unsigned>(e2[0]) | (static_cast<unsigned>(e2[1]) << 16);
            }
            Bf[st][ct] = __builtin_bit_cast(bf16x8, pk);
        }
    float bias[CT];
#pragma unroll
    for (int ct = 0; ct < CT; ++ct) bias[ct] = a.bias[16 * ct + mr];
    float qv[CT][kProjOut];
    if (PROJ)
#pragma unroll
        for (int ct = 0; ct < CT; ++ct)
#pragma unroll
            for (int j = 0; j < kProjOut; ++j) qv[ct][j] = a.Q[(16 * ct + mr) * kProjOut + j];
    f32x4 acc2[JT][CT];
#pragma unroll
    for (int jt = 0; jt < JT; ++jt)
#pragma unroll
        for (int ct = 0; ct < CT; ++ct) acc2[jt][ct] = f32x4{0.f, 0.f, 0.f, 0.f};

    const unsigned short* x16 = reinterpret_cast<const unsigned short*>(a.x) + row * P * C;
    unsigned short* xo16 = a.x_out ? reinterpret_cast<unsigned short*>(a.x_out) + row * P * C : nullptr;
    unsigned short* t = tr[wave];
    for (int p0 = 0; p0 < P; p0 += 16) {
        // ---- z = [x | Iphi] B + bias: A[m][k], m = cell p0 + mr
        const int p = p0 + mr;
        f32x4 acc[CT];
#pragma unroll
        for (int ct = 0; ct < CT; ++ct) acc[ct] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int st = 0; st < KS; ++st) {
            const int k0 = 32 * st + 8 * g;                 // this lane's 8 k values
            u32x4 av;
            if (k0 < C) {
                if (lift) {
                    // the kLiftIn features of cell p in the first lane group
                    av = u32x4{0u, 0u, 0u, 0u};
                    if (k0 == 0) {
                        const float4* ip = reinterpret_cast<const float4*>(a.in + (row * P + p) * kLiftIn);
                        const float4 f0 = ip[0], f1 = ip[1];
                        av[0] = static_cast<unsigned>(to_bf16(f0.x)) | (static_cast<unsigned>(to_bf16(f0.y)) << 16);
                        av[1] = static_cast<unsigned>(to_bf16(f0.z)) | (static_cast<unsigned>(to_bf16(f0.w)) << 16);
                        av[2] = static_cast<unsigned>(to_bf16(f1.x)) | (static_cast<unsigned>(to_bf16(f1.y)) << 16);
                        av[3] = static_cast<unsigned>(to_bf16(f1.z)) | (static_cast<unsigned>(to_bf16(f1.w)) << 16);
                    }
                } else {
                    av = *reinterpret_cast<const u32x4*>(x16 + static_cast<long long>(p) * C + k0);
                }
            } else {
                av = *reinterpret_cast<const u32x4*>(a.Iphi + static_cast<long long>(p) * J + (k0 - C));
            }
            const bf16x8 A = __builtin_bit_cast(bf16x8, av);
#pragma unroll
            for (int ct = 0; ct < CT; ++ct) acc[ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, Bf[st][ct], acc[ct], 0, 0, 0);
        }
        // ---- bias, activation; lane (mr, g) holds x'[p0 + 4 g + v][16 ct + mr]
        bf16x4 Bx[CT];
#pragma unroll
        for (int ct = 0; ct < CT; ++ct) {
            unsigned short bv[4];
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const float z = acc[ct][v] + bias[ct];
                const float xv = a.act ? gelu_exp(z) : z;
                acc[ct][v] = xv;
                bv[v] = to_bf16(xv);
                if (xo16) t[(4 * g + v) * TRW + 16 * ct + mr] = bv[v];
            }
            Bx[ct] = bf16x4{static_cast<short>(bv[0]), static_cast<short>(bv[1]), static_cast<short>(bv[2]),
                            static_cast<short>(bv[3])};
        }
        if (xo16) {
            // transpose through the wave's LDS tile: 16-byte stores of 8 channels
            __builtin_amdgcn_wave_barrier();
            constexpr int LPC = C / 8;                  // lanes per cell
            constexpr int CPP = 64 / LPC;               // cells per pass (more than 16 when C = 16)
#pragma unroll
            for (int c0 = 0; c0 < 16; c0 += CPP) {
                const int cell = c0 + lane / LPC, ch = 8 * (lane % LPC);
                if (cell < 16) {
                    const u32x4 v = *reinterpret_cast<const u32x4*>(&t[cell * TRW + ch]);
                    *reinterpret_cast<u32x4*>(xo16 + static_cast<long long>(p0 + cell) * C + ch) = v;
                }
            }
            __builtin_amdgcn_wave_barrier();
        }
        if (PROJ) {
            // proj[p][j] = q[j] + sum_o x'[p][o] Q[o][j]: per-lane partials
            // e = 2 v + j over ct, then a reduce-scatter over the 16 lanes of
            // the group (halving the partials at each of the xor-8/4/2 steps,
            // 7 shuffles in all): lane mr ends with output e = mr / 2, and the
            // even lanes store outputs 8 g .. 8 g + 7 of the tile
            static_assert(kProjOut == 2, "the reduce-scatter pairs 8 partials with 16 lanes");
            float e8[8];
#pragma unroll
            for (int v = 0; v < 4; ++v)
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    float sum = 0.f;
#pragma unroll
                    for (int ct = 0; ct < CT; ++ct) sum = fmaf(acc[ct][v], qv[ct][j], sum);
                    e8[2 * v + j] = sum;
                }
            const bool b8 = (mr & 8) != 0, b4 = (mr & 4) != 0, b2 = (mr & 2) != 0;
            float e4[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const float send = b8 ? e8[k] : e8[k + 4];
                const float keep = b8 ? e8[k + 4] : e8[k];
                e4[k] = keep + __shfl_xor(send, 8, 16);
            }
            float e2[2];
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                const float send = b4 ? e4[k] : e4[k + 2];
                const float keep = b4 ? e4[k + 2] : e4[k];
                e2[k] = keep + __shfl_xor(send, 4, 16);
            }
            const float send = b2 ? e2[0] : e2[1];
            float e1 = (b2 ? e2[1] : e2[0]) + __shfl_xor(send, 2, 16);
            e1 += __shfl_xor(e1, 1, 16);
            const int e = mr >> 1;
            if ((mr & 1) == 0) a.proj[(row * P + p0 + 4 * g) * kProjOut + e] = e1 + a.qb[e & 1];
        }
        if (!PROJ) {
            // next layer's phi DFT: A[j][k] = Ephi[16 jt + mr][p0 + 4 g + v]
#pragma unroll
            for (int jt = 0; jt < JT; ++jt) {
                const bf16x4 E = *reinterpret_cast<const bf16x4*>(a.Ephi + static_cast<long long>(16 * jt + mr) * P + p0 +
                                                                  4 * g);
#pragma unroll
                for (int ct = 0; ct < CT; ++ct)
                    acc2[jt][ct] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(E, Bx[ct], acc2[jt][ct], 0, 0, 0);
            }
        }
    }
    if (!PROJ) {
        float* X1 = a.X1_out + row * J * C;
#pragma unroll
        for (int jt = 0; jt < JT; ++jt)
#pragma unroll
            for (int ct = 0; ct < CT; ++ct)
#pragma unroll
                for (int v = 0; v < 4; ++v) X1[(16 * jt + 4 * g + v) * C + 16 * ct + mr] = acc2[jt][ct][v];
    }
}

__device__ inline float gelu_grad(float z) {
    const float k0 = 0.7978845608028654f, k1 = 0.044715f;
    const float u = k0 * (z + k1 * z * z * z);
    const float t = tanhf(u);
    return 0.5f * (1.f + t) + 0.5f * z * (1.f - t * t) * k0 * (1.f + 3.f * k1 * z * z);
}

// x_out = act(z), z = z_in + bias[c] (+ y), over n rows of C channels.
__global__ void bias_act_kernel(float* z, const float* y, const float* bias, float* x_out, long long total, int C,
                                int act) {
    for (long long k = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; k < total;
         k += static_cast<long long>(gridDim.x) * blockDim.x) {
        float v = z[k] + bias[k % C];
        if (y) v += y[k];
        z[k] = v;
        x_out[k] = act ? gelu(v) : v;
    }
}

// Lift: x[r][c] = b[c] + sum_i in[r][i] W[i][c] for the NIN input channels,
// in FP32.  Replaces a K = NIN GEMM and a bias pass over x.  When C is a
// multiple of 4 whose quarter divides the block, a thread writes four
// channels of one row with one 16-byte store (its weights held in registers)
// and the block covers blockDim / (C/4) whole rows; otherwise one thread per
// element.
template <int NIN>
__global__ __launch_bounds__(256) void lift_kernel(const float* __restrict__ in, const float* __restrict__ W,
                                                   const float* __restrict__ b, float* __restrict__ x, long long rows,
                                                   int C) {
    const int C4 = C >> 2;
    if (C4 > 0 && (C & 3) == 0 && static_cast<int>(blockDim.x) % C4 == 0) {
        const int per = static_cast<int>(blockDim.x) / C4;
        const int rl = static_cast<int>(threadIdx.x) / C4;
        const int c0 = 4 * (static_cast<int>(threadIdx.x) - rl * C4);
        float4 w[NIN];
#pragma unroll
        for (int i = 0; i < NIN; ++i) w[i] = *reinterpret_cast<const float4*>(W + i * C + c0);
        const float4 bias = *reinterpret_cast<const float4*>(b + c0);
        for (long long r = blockIdx.x * static_cast<long long>(per) + rl; r < rows;
             r += static_cast<long long>(gridDim.x) * per) {
            float4 v = bias;
#pragma unroll
            for (int i = 0; i < NIN; ++i) {
                const float a = in[r * NIN + i];
                v.x = fmaf(a, w[i].x, v.x);
                v.y = fmaf(a, w[i].y, v.y);
                v.z = fmaf(a, w[i].z, v.z);
                v.w = fmaf(a, w[i].w, v.w);
            }
            *reinterpret_cast<float4*>(x + r * C + c0) = v;
        }
        return;
    }
    const long long total = rows * C;
    for (long long k = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; k < total;
         k += static_cast<long long>(gridDim.x) * blockDim.x) {
        const long long r = k / C;
        const int c = static_cast<int>(k - r * C);
        float v = b[c];
#pragma unroll
        for (int i = 0; i < NIN; ++i) v = fmaf(in[r * NIN + i], W[i * C + c], v);
        x[k] = v;
    }
}

// Layer-0 phi DFT straight from the input features: with the lift
// x = in W_l + 1 b_l, X1 = Ephi x = (Ephi in) W_l + (Ephi 1) b_l per
// (bubble, theta row), in FP32.  One block per row, P <= kLiftDftMaxP.
// Lane l holds the inputs of columns p = l + 64 k; wave w forms the products
// of modes j = w, w + waves, ... from one coalesced read of the Ephi row (all
// loads of a row issued before use) and a butterfly over the wave.
constexpr int kLiftDftMaxP = 512;

template <int NIN>
__global__ __launch_bounds__(256) void lift_dft_kernel(const float* __restrict__ in, const float* __restrict__ Ephi,
                                                       const float* __restrict__ Wl, const float* __restrict__ bl,
                                                       float* __restrict__ X1, int P, int J, int C) {
    constexpr int KP = kLiftDftMaxP / 64;
    extern __shared__ float lds[];
    float* s_in = lds;                  // [P][NIN]
    float* s_e = lds + P * NIN;         // [J][NIN + 1]
    const long long row = blockIdx.x;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    for (int t = threadIdx.x; t < P * NIN; t += blockDim.x) s_in[t] = in[row * P * NIN + t];
    __syncthreads();
    float fin[KP][NIN];
#pragma unroll
    for (int k = 0; k < KP; ++k) {
        const int p = lane + 64 * k;
#pragma unroll
        for (int i = 0; i < NIN; ++i) fin[k][i] = p < P ? s_in[p * NIN + i] : 0.f;
    }
    for (int j = wave; j < J; j += nw) {
        const float* e = Ephi + static_cast<long long>(j) * P;
        float ev[KP];
#pragma unroll
        for (int k = 0; k < KP; ++k) ev[k] = lane + 64 * k < P ? e[lane + 64 * k] : 0.f;
        float acc[NIN + 1];
#pragma unroll
        for (int i = 0; i <= NIN; ++i) acc[i] = 0.f;
#pragma unroll
        for (int k = 0; k < KP; ++k) {
#pragma unroll
            for (int i = 0; i < NIN; ++i) acc[i] = fmaf(ev[k], fin[k][i], acc[i]);
            acc[NIN] += ev[k];
        }
#pragma unroll
        for (int i = 0; i <= NIN; ++i)
#pragma unroll
            for (int m = 32; m >= 1; m >>= 1) acc[i] += __shfl_xor(acc[i], m, 64);
        if (lane == 0)
#pragma unroll
            for (int i = 0; i <= NIN; ++i) s_e[j * (NIN + 1) + i] = acc[i];
    }
    __syncthreads();
    for (int t = threadIdx.x; t < J * C; t += blockDim.x) {
        const int j = t / C, c = t - j * C;
        const float* ej = s_e + j * (NIN + 1);
        float s = ej[NIN] * bl[c];
#pragma unroll
        for (int i = 0; i < NIN; ++i) s = fmaf(ej[i], Wl[i * C + c], s);
        X1[row * J * C + t] = s;
    }
}

// The same product on FP32 MFMA (v_mfma_f32_16x16x4_f32, lane l supplies
// A[l % 16][l / 16] and B[l / 16][l % 16]), one wave per kLiftRows rows of a
// bubble's grid, so each 16-byte load of the Ephi rows (A) serves kLiftRows
// products: S_r = Ephi in_r (J x NIN, K = P split so that lane group g sums
// cells g P/4 .. (g + 1) P/4 - 1), two rows per 16-wide tile (column
// n = 8 (r % 2) + f), then X1 = S [W_l] + Esum b_l from a per-wave LDS copy of
// S, Esum[j] = sum_p Ephi[j][p] (the DFT of the constant input, host-summed).
// Needs P % 16 == 0, J == 32, NIN == 8 and 64 % C == 0.
constexpr int kLiftWaveJ = 32;
constexpr int kLiftRows = 4;

template <int NIN>
__global__ __launch_bounds__(256) void lift_dft_wave_kernel(const float* __restrict__ in, const float* __restrict__ Ephi,
                                                            const float* __restrict__ Esum,
                                                            const float* __restrict__ Wl, const float* __restrict__ bl,
                                                            float* __restrict__ X1, int P, int C, long long rows) {
    static_assert(NIN == 8, "two rows' inputs fill one 16-wide tile");
    constexpr int SW = NIN + 1;                           // padded row of S in LDS
    constexpr int J = kLiftWaveJ, JT = J / 16, RT = kLiftRows / 2;
    __shared__ float s_S[4][kLiftRows][J * SW];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int mr = lane & 15, g = lane >> 4;
    const long long row0 = (blockIdx.x * 4LL + wave) * kLiftRows;
    if (row0 >= rows) return;                             // no barrier below: waves are independent
    const int nr = static_cast<int>(rows - row0 < kLiftRows ? rows - row0 : kLiftRows);
    const int P4 = P >> 2;
    f32x4 acc[RT][JT];
#pragma unroll
    for (int rt = 0; rt < RT; ++rt)
#pragma unroll
        for (int jt = 0; jt < JT; ++jt) acc[rt][jt] = f32x4{0.f, 0.f, 0.f, 0.f};
    const float* e0 = Ephi + static_cast<long long>(mr) * P + g * P4;
    const int rl = mr >> 3, f = mr & 7;                   // this lane's row of the pair and feature
    for (int kk = 0; kk < P4; kk += 4) {
        float4 ev[JT];
#pragma unroll
        for (int jt = 0; jt < JT; ++jt) ev[jt] = *reinterpret_cast<const float4*>(e0 + static_cast<long long>(16 * jt) * P + kk);
#pragma unroll
        for (int rt = 0; rt < RT; ++rt) {
            if (2 * rt < nr) {
                const int r = 2 * rt + rl;
                float bv[4] = {0.f, 0.f, 0.f, 0.f};
                if (r < nr) {
                    const float* inr = in + (row0 + r) * P * NIN + static_cast<long long>(g * P4 + kk) * NIN + f;
#pragma unroll
                    for (int u = 0; u < 4; ++u) bv[u] = inr[u * NIN];
                }
#pragma unroll
                for (int jt = 0; jt < JT; ++jt) {
                    acc[rt][jt] = __builtin_amdgcn_mfma_f32_16x16x4f32(ev[jt].x, bv[0], acc[rt][jt], 0, 0, 0);
                    acc[rt][jt] = __builtin_amdgcn_mfma_f32_16x16x4f32(ev[jt].y, bv[1], acc[rt][jt], 0, 0, 0);
                    acc[rt][jt] = __builtin_amdgcn_mfma_f32_16x16x4f32(ev[jt].z, bv[2], acc[rt][jt], 0, 0, 0);
                    acc[rt][jt] = __builtin_amdgcn_mfma_f32_16x16x4f32(ev[jt].w, bv[3], acc[rt][jt], 0, 0, 0);
                }
            }
        }
    }
    // lane (mr, g) holds S_{2 rt + mr / 8}[16 jt + 4 g + v][mr % 8]
#pragma unroll
    for (int rt = 0; rt < RT; ++rt)
#pragma unroll
        for (int jt = 0; jt < JT; ++jt)
#pragma unroll
            for (int v = 0; v < 4; ++v) s_S[wave][2 * rt + rl][(16 * jt + 4 * g + v) * SW + f] = acc[rt][jt][v];
    __builtin_amdgcn_wave_barrier();
    // X1[j][c] = Esum[j] b_l[c] + sum_f S[j][f] W_l[f][c]; c = lane % C is
    // fixed per lane (64 % C == 0), its W_l column held in registers
    const int c = lane % C;
    float wl[NIN];
#pragma unroll
    for (int q = 0; q < NIN; ++q) wl[q] = Wl[q * C + c];
    const float b = bl[c];
    for (int r = 0; r < nr; ++r) {
        const float* S = s_S[wave][r];
        float* xr = X1 + (row0 + r) * J * C;
        for (int t = lane; t < J * C; t += 64) {
            const int j = t / C;
            const float* sj = S + j * SW;
            float sum = Esum[j] * b;
#pragma unroll
            for (int q = 0; q < NIN; ++q) sum = fmaf(sj[q], wl[q], sum);
            xr[t] = sum;
        }
    }
}

// Projection: out[r][j] = q[j] + sum_c x[r][c] Q[c][j] for the NOUT output
// channels, in FP32, one thread per row.  Replaces an N = NOUT GEMM (a 64-wide
// tile for two columns) and a bias pass.
template <int NOUT>
__global__ __launch_bounds__(256) void project_kernel(const float* __restrict__ x, const float* __restrict__ Q,
                                                      const float* __restrict__ q, float* __restrict__ out,
                                                      long long rows, int C) {
    __shared__ float sQ[64 * NOUT + NOUT];
    for (int t = threadIdx.x; t < C * NOUT; t += blockDim.x) sQ[t] = Q[t];
    for (int t = threadIdx.x; t < NOUT; t += blockDim.x) sQ[64 * NOUT + t] = q[t];
    __syncthreads();
    for (long long r = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; r < rows;
         r += static_cast<long long>(gridDim.x) * blockDim.x) {
        float o[NOUT];
#pragma unroll
        for (int j = 0; j < NOUT; ++j) o[j] = sQ[64 * NOUT + j];
        const float* xr = x + r * C;
        if ((C & 3) == 0) {
            for (int c = 0; c < C; c += 4) {
                const float4 v = *reinterpret_cast<const float4*>(xr + c);
#pragma unroll
                for (int j = 0; j < NOUT; ++j) {
                    o[j] = fmaf(v.x, sQ[(c + 0) * NOUT + j], o[j]);
                    o[j] = fmaf(v.y, sQ[(c + 1) * NOUT + j], o[j]);
                    o[j] = fmaf(v.z, sQ[(c + 2) * NOUT + j], o[j]);
                    o[j] = fmaf(v.w, sQ[(c + 3) * NOUT + j], o[j]);
                }
            }
        } else {
            for (int c = 0; c < C; ++c)
#pragma unroll
                for (int j = 0; j < NOUT; ++j) o[j] = fmaf(xr[c], sQ[c * NOUT + j], o[j]);
        }
#pragma unroll
        for (int j = 0; j < NOUT; ++j) out[r * NOUT + j] = o[j];
    }
}

// dz = dx * act'(z) (in place on dx when act, else a copy).
__global__ void act_grad_kernel(const float* z, const float* dx, float* dz, long long total, int act) {
    for (long long k = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; k < total;
         k += static_cast<long long>(gridDim.x) * blockDim.x)
        dz[k] = act ? dx[k] * gelu_grad(z[k]) : dx[k];
}

// Compensated (Kahan) FP32 sum: s + c carries about twice FP32 precision,
// so the reductions below need no FP64.
struct Kahan {
    float s = 0.f, c = 0.f;
    __device__ inline void add(float x) {
#pragma clang fp contract(off)
        const float y = x - c;
        const float t = s + y;
        c = (t - s) - y;
        s = t;
    }
    __device__ inline float value() const { return s - c; }
};

// out[j] (+)= sum over rows of part[row][j], compensated FP32.  Block of
// 256 threads: 32 consecutive columns x 8 row lanes, folded in LDS.
__global__ __launch_bounds__(256) void column_sum_kernel(const float* part, long long rows, int cols, float* out,
                                                         int accumulate) {
    const int tx = threadIdx.x & 31, rl = threadIdx.x >> 5;
    const int j = blockIdx.x * 32 + tx;
    Kahan s;
    if (j < cols)
        for (long long r = rl; r < rows; r += 8) s.add(part[r * cols + j]);
    __shared__ float red[8][32];
    red[rl][tx] = s.value();
    __syncthreads();
    if (rl == 0 && j < cols) {
        Kahan t;
        if (accumulate) t.add(out[j]);
        for (int q = 0; q < 8; ++q) t.add(red[q][tx]);
        out[j] = t.value();
    }
}

// Per-block column sums of a [rows][C] tensor: partial[block][C].
__global__ __launch_bounds__(256) void column_partial_kernel(const float* x, long long rows, int C, float* partial) {
    // 256 threads: thread t sums column (t % C) over rows congruent to t / C.
    const int per = 256 / C;
    const int c = threadIdx.x % C, lane_row = threadIdx.x / C;
    Kahan s;
    if (lane_row < per) {
        for (long long r = blockIdx.x * static_cast<long long>(per) + lane_row; r < rows;
             r += static_cast<long long>(gridDim.x) * per)
            s.add(x[r * C + c]);
    }
    __shared__ float red[256];
    red[threadIdx.x] = s.value();
    __syncthreads();
    if (threadIdx.x < C) {
        Kahan t;
        for (int q = 0; q < per; ++q) t.add(red[q * C + threadIdx.x]);
        partial[blockIdx.x * C + threadIdx.x] = t.value();
    }
}

// Complex combine after the theta DFT.  G[b] = [Er; Ei] (2Mt x T) * X1[b]
// (T x (2 Mp C)), laid out [2][Mt][2][Mp][C] as (E part, k, re/im of X1, m, c).
// Xs[b][k][m][2][c] = (Er Xr - Ei Xi, Er Xi + Ei Xr).
__global__ void theta_combine_kernel(const float* G, float* Xs, int Mt, int Mp, int C, long long total) {
    for (long long q = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; q < total;
         q += static_cast<long long>(gridDim.x) * blockDim.x) {
        // q indexes [b][k][m][c]
        const int c = static_cast<int>(q % C);
        long long r = q / C;
        const int m = static_cast<int>(r % Mp);
        r /= Mp;
        const int k = static_cast<int>(r % Mt);
        const long long b = r / Mt;
        const long long base = b * (2LL * Mt * 2 * Mp * C);
        auto at = [&](int part, int ri) { return G[base + ((static_cast<long long>(part) * Mt + k) * 2 + ri) * Mp * C + static_cast<long long>(m) * C + c]; };
        const float re = at(0, 0) - at(1, 1);
        const float im = at(0, 1) + at(1, 0);
        const long long o = (((b * Mt + k) * Mp + m) * 2) * C + c;
        Xs[o] = re;
        Xs[o + C] = im;
    }
}

// Adjoint of theta_combine_kernel: dG from dXs.
__global__ void theta_combine_adjoint_kernel(const float* dXs, float* dG, int Mt, int Mp, int C, long long total) {
    for (long long q = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; q < total;
         q += static_cast<long long>(gridDim.x) * blockDim.x) {
        const int c = static_cast<int>(q % C);
        long long r = q / C;
        const int m = static_cast<int>(r % Mp);
        r /= Mp;
        const int k = static_cast<int>(r % Mt);
        const long long b = r / Mt;
        const long long o = (((b * Mt + k) * Mp + m) * 2) * C + c;
        const float dre = dXs[o], dim = dXs[o + C];
        const long long base = b * (2LL * Mt * 2 * Mp * C);
        auto at = [&](int part, int ri) -> float& { return dG[base + ((static_cast<long long>(part) * Mt + k) * 2 + ri) * Mp * C + static_cast<long long>(m) * C + c]; };
        at(0, 0) = dre;
        at(1, 1) = -dre;
        at(0, 1) = dim;
        at(1, 0) = dim;
    }
}

// Spectral mixing per retained mode (k, m): Y[b][k][m][o] = sum_c Xs[b][k][m][c] R[k][m][c][o]
// (complex).  Xs: [B][Mt][Mp][2][C]; R: [Mt][Mp][2][C][C]; Y written in the
// layout of the inverse theta GEMM: [B][2][Mt][Mp][C].  One block per mode.
// The mode's weights (2 C C) and, when stage_x is set, the Bn input vectors
// (2 C each) are staged in dynamic LDS first, so the products read no global
// memory.
__global__ void mode_mix_kernel(const float* Xs, const float* R, float* Y, int Bn, int Mt, int Mp, int C,
                                int stage_x) {
    extern __shared__ float mm_lds[];
    float* sR = mm_lds;                        // [2][C][C]
    float* sX = mm_lds + 2 * C * C;            // [Bn][2][C]
    const int mode = blockIdx.x;               // k * Mp + m
    const float* Rm = R + static_cast<long long>(mode) * 2 * C * C;
    for (int t = threadIdx.x; t < 2 * C * C; t += blockDim.x) sR[t] = Rm[t];
    if (stage_x)
        for (int t = threadIdx.x; t < Bn * 2 * C; t += blockDim.x) {
            const int b = t / (2 * C), e = t - b * 2 * C;
            sX[t] = Xs[(static_cast<long long>(b) * Mt * Mp + mode) * 2 * C + e];
        }
    __syncthreads();
    for (int q = threadIdx.x; q < Bn * C; q += blockDim.x) {
        const int b = q / C, o = q % C;
        const float* x = stage_x ? sX + b * 2 * C : Xs + (static_cast<long long>(b) * Mt * Mp + mode) * 2 * C;
        float re = 0.f, im = 0.f;
        for (int c = 0; c < C; ++c) {
            const float xr = x[c], xi = x[C + c];
            const float rr = sR[c * C + o], ri = sR[C * C + c * C + o];
            re += xr * rr - xi * ri;
            im += xr * ri + xi * rr;
        }
        const long long plane = static_cast<long long>(Mt) * Mp * C;
        const long long o_re = static_cast<long long>(b) * 2 * plane + static_cast<long long>(mode) * C + o;
        Y[o_re] = re;
        Y[o_re + plane] = im;
    }
}

// Fused theta stage of one layer for inference, one block (8 waves) per
// (phi mode m, bubble b): the theta DFT of the m-th phi coefficients of all
// T rows, the complex combination, the per-mode channel mixing and the
// inverse theta DFT, with the intermediate spectrum in LDS.  It computes the
// same products as the gemm_kernel / theta_combine_kernel / mode_mix_kernel /
// gemm_kernel sequence, in the same order and precision (BF16 operands of
// 16x16x16 MFMAs accumulated in FP32 over increasing k; FP32 combination and
// mixing), so the results match that sequence.
//   X1  [B T][J][C]     phi DFT of the layer input, J = 2 Mp (re rows, then im)
//   Eth [2 Mt][T] BF16  forward theta DFT (cos rows, then -sin rows)
//   R   [Mt Mp][2][C][C] mode weights (re, im)
//   Ath [2 T][2 Mt] BF16 inverse theta DFT
//   Z   [B T][J][C]     output (the columns j = m and Mp + m of this block)
// Dynamic LDS: X slice [2C][T + 8] BF16, G [2 Mt][2C + 1] FP32, Y [C][2 Mt + 8]
// BF16.  Where every wave holds at most two G tiles and G and Y fit in the
// X slice, G and Y take the slice's place once it has been read (each wave
// keeps its G tiles in registers across a barrier), so the block needs only
// the slice: 34 KB at T = 256, C = 32, Mt = 32, four blocks per CU instead of
// two.  Requires C, 2 Mt and T multiples of 16 and 512 % 2C == 0.
// TT, CT, MTT, MPT fix T, C, Mt and Mp at compile time (0: taken from the
// arguments): with them fixed the index arithmetic folds into constants and
// immediate offsets, where it otherwise costs more instructions than the
// products.
constexpr int kThetaThreads = 512;

__host__ __device__ inline bool theta_fused_alias(int T, int Mt, int C) {
    const long sx = 2L * C * (T + 8) * 2, sg = 2L * Mt * (2 * C + 1) * 4, sy = 1L * C * (2 * Mt + 8) * 2;
    return ((2 * Mt) / 16) * ((2 * C) / 16) <= 2 * (kThetaThreads / 64) && sg + sy <= sx;
}
__host__ __device__ inline long theta_fused_lds_bytes(int T, int Mt, int C) {
    const long sx = 2L * C * (T + 8) * 2, sg = 2L * Mt * (2 * C + 1) * 4, sy = 1L * C * (2 * Mt + 8) * 2;
    return theta_fused_alias(T, Mt, C) ? sx : sx + sg + sy;
}

template <int TT, int CT, int MTT, int MPT>
__global__ __launch_bounds__(kThetaThreads) __attribute__((amdgpu_waves_per_eu(8))) void theta_fused_kernel(const float* __restrict__ X1,
                                                                    const unsigned short* __restrict__ Eth,
                                                                    const float* __restrict__ R,
                                                                    const unsigned short* __restrict__ Ath,
                                                                    float* __restrict__ Z, int T_, int Mt_, int Mp_,
                                                                    int C_) {
    extern __shared__ __align__(16) unsigned char th_lds[];
    const int T = TT ? TT : T_, Mt = MTT ? MTT : Mt_, Mp = MPT ? MPT : Mp_, C = CT ? CT : C_;
    const int m = blockIdx.x;
    const long long b = blockIdx.y;
    const int J = 2 * Mp, C2 = 2 * C, M2 = 2 * Mt;
    const int xs_ld = T + 8, g_ld = C2 + 1, y_ld = M2 + 8;
    const bool alias = theta_fused_alias(T, Mt, C);
    unsigned short* sx = reinterpret_cast<unsigned short*>(th_lds);                        // [C2][xs_ld]
    float* sg = reinterpret_cast<float*>(alias ? th_lds : th_lds + static_cast<size_t>(C2) * xs_ld * 2);   // [M2][g_ld]
    unsigned short* sy = reinterpret_cast<unsigned short*>(sg + static_cast<size_t>(M2) * g_ld);   // [C][y_ld]
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    constexpr int nwaves = kThetaThreads / 64;
    const int mr = lane & 15, kq = 4 * (lane >> 4);

    // 1. the m-th coefficients of every row, transposed to [q][t], q = ri C + c:
    // a thread keeps one q and takes every (512 / 2C)-th row; sixteen loads
    // per thread are in flight before their stores
    {
        const int q = tid % C2, tstep = kThetaThreads / C2;
        const int ri = q / C, c = q - ri * C;
        const int JC = J * C;
        const float* src = X1 + ((b * T) * J + ri * Mp + m) * C + c;
        unsigned short* dst = sx + q * xs_ld;
        for (int t0 = tid / C2; t0 < T; t0 += 16 * tstep) {
            float v[16];
#pragma unroll
            for (int s = 0; s < 16; ++s) {
                const int t = t0 + s * tstep;
                v[s] = t < T ? src[t * JC] : 0.f;
            }
#pragma unroll
            for (int s = 0; s < 16; ++s) {
                const int t = t0 + s * tstep;
                if (t < T) dst[t] = to_bf16(v[s]);
            }
        }
    }
    __syncthreads();

    // 2. G[p][q] = sum_t Eth[p][t] X[t][q]
    const int gt_p = M2 / 16, gt_q = C2 / 16;
    f32x4 keep0 = f32x4{0.f, 0.f, 0.f, 0.f}, keep1 = keep0;     // with alias: the wave's tiles
    for (int tile = wave; tile < gt_p * gt_q; tile += nwaves) {
        const int pt = tile / gt_q, qt = tile - pt * gt_q;
        f32x4 acc = f32x4{0.f, 0.f, 0.f, 0.f};
        const unsigned short* erow = Eth + static_cast<long long>(pt * 16 + mr) * T + kq;
        // Eight k-steps' A fragments are loaded before their MFMAs, so the
        // global loads overlap instead of each waiting in turn.
        for (int k0 = 0; k0 < T; k0 += 128) {
            bf16x4 Af[8];
#pragma unroll
            for (int s = 0; s < 8; ++s)
                Af[s] = k0 + 16 * s < T ? *reinterpret_cast<const bf16x4*>(erow + k0 + 16 * s) : bf16x4{0, 0, 0, 0};
#pragma unroll
            for (int s = 0; s < 8; ++s) {
                if (k0 + 16 * s >= T) break;
                const bf16x4 Bf = *reinterpret_cast<const bf16x4*>(sx + (qt * 16 + mr) * xs_ld + k0 + 16 * s + kq);
                acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(Af[s], Bf, acc, 0, 0, 0);
            }
        }
        if (alias) {
            if (tile == wave)
                keep0 = acc;
            else
                keep1 = acc;
        } else {
#pragma unroll
            for (int v = 0; v < 4; ++v) sg[(pt * 16 + kq + v) * g_ld + qt * 16 + mr] = acc[v];
        }
    }
    if (alias) {
        __syncthreads();                          // every wave has read the X slice
        for (int i = 0; i < 2; ++i) {
            const int tile = wave + i * nwaves;
            if (tile >= gt_p * gt_q) break;
            const int pt = tile / gt_q, qt = tile - pt * gt_q;
            const f32x4 acc = i == 0 ? keep0 : keep1;
#pragma unroll
            for (int v = 0; v < 4; ++v) sg[(pt * 16 + kq + v) * g_ld + qt * 16 + mr] = acc[v];
        }
    }
    __syncthreads();

    // 3-4. spectrum of mode (k, m) (theta_combine_kernel) mixed over channels
    // (mode_mix_kernel); Y stored transposed [o][q], q = plane Mt + k
    for (int idx = tid; idx < Mt * C; idx += blockDim.x) {
        const int k = idx / C, o = idx - k * C;
        const float* Rm = R + static_cast<long long>(k * Mp + m) * 2 * C * C;
        float re = 0.f, im = 0.f;
        // Eight channels' weights are loaded before their products (C is a
        // multiple of 16).
        for (int c0 = 0; c0 < C; c0 += 8) {
            float rr[8], rim[8];
#pragma unroll
            for (int s = 0; s < 8; ++s) {
                rr[s] = Rm[(c0 + s) * C + o];
                rim[s] = Rm[C * C + (c0 + s) * C + o];
            }
#pragma unroll
            for (int s = 0; s < 8; ++s) {
                const int c = c0 + s;
                const float xr = sg[k * g_ld + c] - sg[(Mt + k) * g_ld + C + c];
                const float xi = sg[k * g_ld + C + c] + sg[(Mt + k) * g_ld + c];
                re += xr * rr[s] - xi * rim[s];
                im += xr * rim[s] + xi * rr[s];
            }
        }
        sy[o * y_ld + k] = to_bf16(re);
        sy[o * y_ld + Mt + k] = to_bf16(im);
    }
    __syncthreads();

    // 5. Z[r][o] = sum_q Ath[r][q] Y[q][o], r = 2 t + ri
    const int zt_r = (2 * T) / 16, zt_o = C / 16;
    for (int tile = wave; tile < zt_r * zt_o; tile += nwaves) {
        const int rt = tile / zt_o, ot = tile - rt * zt_o;
        f32x4 acc = f32x4{0.f, 0.f, 0.f, 0.f};
        const unsigned short* arow = Ath + static_cast<long long>(rt * 16 + mr) * M2 + kq;
        for (int k0 = 0; k0 < M2; k0 += 128) {
            bf16x4 Af[8];
#pragma unroll
            for (int s = 0; s < 8; ++s)
                Af[s] = k0 + 16 * s < M2 ? *reinterpret_cast<const bf16x4*>(arow + k0 + 16 * s) : bf16x4{0, 0, 0, 0};
#pragma unroll
            for (int s = 0; s < 8; ++s) {
                if (k0 + 16 * s >= M2) break;
                const bf16x4 Bf = *reinterpret_cast<const bf16x4*>(sy + (ot * 16 + mr) * y_ld + k0 + 16 * s + kq);
                acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(Af[s], Bf, acc, 0, 0, 0);
            }
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int r = rt * 16 + kq + v, t = r >> 1, ri = r & 1;
            Z[((b * T + t) * J + ri * Mp + m) * C + ot * 16 + mr] = acc[v];
        }
    }
}

// Backward of mode_mix_kernel: dXs (overwritten) and dR (accumulated).
__global__ void mode_mix_grad_kernel(const float* Xs, const float* R, const float* dY, float* dXs, float* dR, int Bn,
                                     int Mt, int Mp, int C) {
    const int mode = blockIdx.x;
    const float* Rm = R + static_cast<long long>(mode) * 2 * C * C;
    float* dRm = dR + static_cast<long long>(mode) * 2 * C * C;
    const long long plane = static_cast<long long>(Mt) * Mp * C;
    // dX[b][c] = sum_o (dYr R_re + dYi R_im, -dYr R_im + dYi R_re)
    for (int q = threadIdx.x; q < Bn * C; q += blockDim.x) {
        const int b = q / C, c = q % C;
        const float* dyr = dY + static_cast<long long>(b) * 2 * plane + static_cast<long long>(mode) * C;
        const float* dyi = dyr + plane;
        float xr = 0.f, xi = 0.f;
        for (int o = 0; o < C; ++o) {
            const float rr = Rm[c * C + o], ri = Rm[C * C + c * C + o];
            xr += dyr[o] * rr + dyi[o] * ri;
            xi += -dyr[o] * ri + dyi[o] * rr;
        }
        float* dx = dXs + (static_cast<long long>(b) * Mt * Mp + mode) * 2 * C;
        dx[c] = xr;
        dx[C + c] = xi;
    }
    // dR[c][o] += sum_b (dYr Xr + dYi Xi, -dYr Xi + dYi Xr)
    for (int q = threadIdx.x; q < C * C; q += blockDim.x) {
        const int c = q / C, o = q % C;
        float gr = 0.f, gi = 0.f;
        for (int b = 0; b < Bn; ++b) {
            const float* x = Xs + (static_cast<long long>(b) * Mt * Mp + mode) * 2 * C;
            const float dyr = dY[static_cast<long long>(b) * 2 * plane + static_cast<long long>(mode) * C + o];
            const float dyi = dY[static_cast<long long>(b) * 2 * plane + plane + static_cast<long long>(mode) * C + o];
            gr += dyr * x[c] + dyi * x[C + c];
            gi += -dyr * x[C + c] + dyi * x[c];
        }
        dRm[c * C + o] += gr;
        dRm[C * C + c * C + o] += gi;
    }
}

// Adam on FP32 master weights (bias-corrected step size passed in).
__global__ void adam_kernel(float* w, const float* g, float* m, float* v, long long n, float lr_t, float b1, float b2,
                            float eps) {
    for (long long k = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; k < n;
         k += static_cast<long long>(gridDim.x) * blockDim.x) {
        const float gk = g[k];
        const float mk = b1 * m[k] + (1.f - b1) * gk;
        const float vk = b2 * v[k] + (1.f - b2) * gk * gk;
        m[k] = mk;
        v[k] = vk;
        w[k] -= lr_t * mk / (sqrtf(vk) + eps);
    }
}

// Loss L = mean((out - target)^2) pieces: dout = 2 (out - target) / count and
// per-block partial sums of the squared error (compensated FP32; the host
// adds the partials).
__global__ __launch_bounds__(256) void mse_kernel(const float* out, const float* target, float* dout, long long n,
                                                  float two_inv_count, float* partial) {
    Kahan s;
    for (long long k = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; k < n;
         k += static_cast<long long>(gridDim.x) * blockDim.x) {
        const float d = out[k] - target[k];
        s.add(d * d);
        dout[k] = two_inv_count * d;
    }
    __shared__ float red[256];
    red[threadIdx.x] = s.value();
    __syncthreads();
    if (threadIdx.x == 0) {
        Kahan t;
        for (int q = 0; q < 256; ++q) t.add(red[q]);
        partial[blockIdx.x] = t.value();
    }
}

}  // namespace fno
}  // namespace sf
