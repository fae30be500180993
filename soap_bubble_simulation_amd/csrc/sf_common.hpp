// sf_common.hpp -- shared helpers of the soapfilm HIP library (gfx950).
//
// Error plumbing for the C ABI (thread-local last error), HIP call checking,
// and the float-float arithmetic the drainage kernels use to carry film
// thickness at ~48-bit precision in FP32 registers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "soapfilm.h"

namespace sf {

// ---- error handling ---------------------------------------------------------

void set_error(const char* fmt, ...);
const char* last_error();

struct Error {
    int code;
};

#define SF_HIP(call)                                                                   \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess) {                                                        \
            ::sf::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_),     \
                            __FILE__, __LINE__);                                       \
            throw ::sf::Error{SF_ERR_DEVICE};                                          \
        }                                                                              \
    } while (0)

#define SF_REQUIRE(cond, code, ...)                                                    \
    do {                                                                               \
        if (!(cond)) {                                                                 \
            ::sf::set_error(__VA_ARGS__);                                              \
            throw ::sf::Error{code};                                                   \
        }                                                                              \
    } while (0)

// Runs `body` and converts thrown sf::Error into the C status code.
template <class F>
int guard(F&& body) {
    try {
        body();
        return SF_OK;
    } catch (const Error& e) {
        return e.code;
    } catch (...) {
        set_error("unexpected C++ exception");
        return SF_ERR_INTERNAL;
    }
}

// Makes sure a HIP device exists; sets the error and throws otherwise.
void require_device();

// ---- float-float arithmetic ----------------------------------------------

// Knuth's TwoSum: a + b == s + e exactly (round-to-nearest, no overflow).
__host__ __device__ inline void two_sum(float a, float b, float& s, float& e) {
    s = a + b;
    float bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}

// Dekker's FastTwoSum, valid when exponent(a) >= exponent(b).
__host__ __device__ inline void fast_two_sum(float a, float b, float& s, float& e) {
    s = a + b;
    e = b - (s - a);
}

// (ah + al) + (bh + bl) as a normalised pair.
__host__ __device__ inline float2 ff_add(float2 a, float2 b) {
    float s, e;
    two_sum(a.x, b.x, s, e);
    e += a.y + b.y;
    float h, l;
    fast_two_sum(s, e, h, l);
    return make_float2(h, l);
}

// Sign of a+b+c+d+e computed exactly (Shewchuk grow-expansion); returns
// -1, 0 or +1.  Used only on the rare slow path of the rupture test.
__host__ __device__ inline int exact_sign5(float a, float b, float c, float d, float e) {
    float x[5];
    int n = 1;
    x[0] = a;
    const float add[4] = {b, c, d, e};
    for (int k = 0; k < 4; ++k) {
        float q = add[k];
        for (int m = 0; m < n; ++m) {
            float s, err;
            two_sum(q, x[m], s, err);
            x[m] = err;
            q = s;
        }
        x[n++] = q;
    }
    for (int m = n - 1; m >= 0; --m) {
        if (x[m] > 0.0f) return 1;
        if (x[m] < 0.0f) return -1;
    }
    return 0;
}

// FP64 value split into a non-overlapping three-float expansion c0+c1+c2
// (exact for every finite double whose exponent fits FP32).
struct Float3 {
    float c0, c1, c2;
};

inline Float3 split3(double v) {
    Float3 r;
    r.c0 = static_cast<float>(v);
    double rem = v - static_cast<double>(r.c0);
    r.c1 = static_cast<float>(rem);
    rem -= static_cast<double>(r.c1);
    r.c2 = static_cast<float>(rem);
    return r;
}

// Exact test hi + lo < c0 + c1 + c2, kept out of line: it only runs when
// hi + lo lies within 2^-10 of c0.
static __device__ __noinline__ bool ff_less_exact(float hi, float lo, float c0, float c1, float c2) {
    return exact_sign5(hi, lo, -c0, -c1, -c2) < 0;
}

// True when hi + lo < c0 + c1 + c2 holds exactly.  The fast path decides
// whenever the rounded sum is further than 2^-10 c0 from c0 (its rounding
// error is at most 2^-24 of the sum).
__device__ inline bool ff_less(float hi, float lo, float c0, float c1, float c2) {
    const float approx = hi + lo;
    bool r = approx < c0;
    if (fabsf(approx - c0) <= fabsf(c0) * 0x1.0p-10f) r = ff_less_exact(hi, lo, c0, c1, c2);
    return r;
}

inline bool ff_less_host(float hi, float lo, float c0, float c1, float c2) {
    return exact_sign5(hi, lo, -c0, -c1, -c2) < 0;
}

}  // namespace sf
