/*
 * sf_oracle_caustics.c -- CPU restatement of the caustic-intensity stencil
 * (src/render/shaders/caustics_compute.wgsl:27-126, parameters
 * src/render/caustics.rs:20-78).  TEST INFRASTRUCTURE ONLY.
 *
 * The shader computes in FP32; this restatement keeps FP32 and the shader's
 * operation order (no contraction, -ffp-contract=off), so the branch
 * decision hessian_det < -branch_threshold is bit-exact with any FP32
 * implementation of the same order.  length() and pow() use libm in FP64
 * on the FP32 values and round once.
 */
#include <math.h>
#include <stddef.h>

#include "sf_oracle.h"

static float sample(const float* h, int nt, int np, int x, int y) /* caustics_compute.wgsl:28-41 */
{
    int sx = x % np;
    if (sx < 0) sx += np;
    int sy = y < 0 ? 0 : (y > nt - 1 ? nt - 1 : y);
    return h[(size_t)sy * np + sx];
}

void orc_caustics(const float* h, int n_theta, int n_phi, float focal_length, float caustic_intensity,
                  float caustic_sharpness, float branch_threshold, float* out, uint8_t* branch)
{
    for (int y = 0; y < n_theta; ++y)
        for (int x = 0; x < n_phi; ++x) {
            const float hh = 1.0f, h2 = hh * hh;
            /* thickness_gradient (44-49) */
            const float gx = (sample(h, n_theta, n_phi, x + 1, y) - sample(h, n_theta, n_phi, x - 1, y)) / (2.0f * hh);
            const float gy = (sample(h, n_theta, n_phi, x, y + 1) - sample(h, n_theta, n_phi, x, y - 1)) / (2.0f * hh);
            /* thickness_laplacian (52-63) */
            const float c = sample(h, n_theta, n_phi, x, y);
            const float l = sample(h, n_theta, n_phi, x - 1, y);
            const float r = sample(h, n_theta, n_phi, x + 1, y);
            const float u = sample(h, n_theta, n_phi, x, y - 1);
            const float d = sample(h, n_theta, n_phi, x, y + 1);
            const float lap = (l + r + u + d - 4.0f * c) / h2;
            /* thickness_hessian_det (66-86) */
            const float d2x = (l + r - 2.0f * c) / h2;
            const float d2y = (u + d - 2.0f * c) / h2;
            const float ul = sample(h, n_theta, n_phi, x - 1, y - 1);
            const float ur = sample(h, n_theta, n_phi, x + 1, y - 1);
            const float dl = sample(h, n_theta, n_phi, x - 1, y + 1);
            const float dr = sample(h, n_theta, n_phi, x + 1, y + 1);
            const float dxy = (dr - dl - ur + ul) / (4.0f * h2);
            const float det = d2x * d2y - dxy * dxy;
            /* compute_caustic_intensity (89-114) */
            const float curv = -lap * focal_length;
            const float a = fabsf(1.0f - curv);
            const float focusing = 1.0f / (a > 0.01f ? a : 0.01f);
            const int br = det < -branch_threshold;
            const float branch_factor = br ? 1.5f : 1.0f;
            const float gmag = (float)sqrt((double)gx * gx + (double)gy * gy);
            const float refraction = 1.0f + gmag * 0.5f;
            float I = focusing * branch_factor * refraction;
            I = (float)pow((double)I, (double)caustic_sharpness);
            I *= caustic_intensity;
            out[(size_t)y * n_phi + x] = I < 0.0f ? 0.0f : (I > 10.0f ? 10.0f : I);
            if (branch) branch[(size_t)y * n_phi + x] = (uint8_t)br;
        }
}
