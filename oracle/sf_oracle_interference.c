/*
 * sf_oracle_interference.c -- restatement of the reference thin-film
 * interference code.
 *
 * TEST INFRASTRUCTURE ONLY (see sf_oracle.h).  Compiled with -O2
 * -ffp-contract=off.
 *
 *   FP64, 3 wavelengths, gamma 1/2.2 : src/physics/interference.rs:197-512
 *   FP32, 7 wavelengths, CIE, RGBA8  : src/render/interference_lut.rs:18-163
 *   81-wavelength spectral RGB       : build-defined generalisation of
 *       interference_lut.rs:45-93 to 81 wavelengths (380-780 nm, 5 nm step)
 *       at each cell's view angle (bubble.wgsl:491-494), contracted with
 *       caller-supplied colour-matching curves; evaluated here in FP64.
 */
#include "sf_oracle.h"

#include <math.h>
#include <stddef.h>
#include <stdint.h>

#define ORC_PI 3.14159265358979323846

/* ---------------- FP64 InterferenceCalculator ---------------------------- */

/* calculate_transmission_angle_cos (interference.rs:197-211) */
double orc_transmission_cos(double n_film, double n_medium, double cos_i)
{
    double c = cos_i < -1.0 ? -1.0 : (cos_i > 1.0 ? 1.0 : cos_i);
    double sin_i = sqrt(1.0 - c * c);                 /* powi(2) == x*x */
    double sin_t = (n_medium / n_film) * sin_i;
    if (fabs(sin_t) > 1.0) return 0.0;
    return sqrt(1.0 - sin_t * sin_t);
}

/* calculate_optical_path_difference (interference.rs:226-238) */
double orc_optical_path_difference(double n_film, double d_nm, double cos_t, double lambda_nm)
{
    double geometric_path = 2.0 * n_film * d_nm * cos_t;
    return geometric_path + lambda_nm / 2.0;
}

/* calculate_phase_difference (interference.rs:248-254) */
double orc_phase_difference(double opd_nm, double lambda_nm)
{
    return 2.0 * ORC_PI * opd_nm / lambda_nm;
}

/* calculate_fresnel_reflection (interference.rs:266-298) */
void orc_fresnel(double n_film, double n_medium, double cos_i, double* rs, double* rp, double* ravg)
{
    double c = cos_i < -1.0 ? -1.0 : (cos_i > 1.0 ? 1.0 : cos_i);
    double cos_t = orc_transmission_cos(n_film, n_medium, c);
    double n_incident = n_medium, n_transmitted = n_film;
    double n1_cos_i = n_incident * c;
    double n2_cos_t = n_transmitted * cos_t;
    double n2_cos_i = n_transmitted * c;
    double n1_cos_t = n_incident * cos_t;
    double as = (n1_cos_i - n2_cos_t) / (n1_cos_i + n2_cos_t);
    double r_s = as * as;
    double ap = (n2_cos_i - n1_cos_t) / (n2_cos_i + n1_cos_t);
    double r_p = ap * ap;
    if (rs) *rs = r_s;
    if (rp) *rp = r_p;
    if (ravg) *ravg = (r_s + r_p) / 2.0;
}

/* calculate_reflected_intensity (interference.rs:312-353) */
double orc_reflected_intensity(double n_film, double n_medium, double d_nm, double cos_i, double lambda_nm)
{
    if (d_nm <= 0.0) return 0.0;
    double cos_t = orc_transmission_cos(n_film, n_medium, cos_i);
    double opd = orc_optical_path_difference(n_film, d_nm, cos_t, lambda_nm);
    double phase = orc_phase_difference(opd, lambda_nm);
    double r;
    orc_fresnel(n_film, n_medium, cos_i, 0, 0, &r);
    double cos_phase = cos(phase);
    double numerator = 2.0 * r * (1.0 - cos_phase);
    double denominator = 1.0 + r * r - 2.0 * r * cos_phase;
    if (fabs(denominator) < 2.220446049250313e-16) return r;   /* f64::EPSILON */
    double v = numerator / denominator;
    return v < 0.0 ? 0.0 : (v > 1.0 ? 1.0 : v);
}

static double clamp01(double x) { return x < 0.0 ? 0.0 : (x > 1.0 ? 1.0 : x); }

/* calculate_interference_color[_linear] (interference.rs:380-426):
 * gamma != 0 applies powf(1/2.2); RgbColor::new clamps to [0,1] (54-60). */
void orc_interference_color(double n_film, const double lambdas[3], double d_nm, double cos_i,
                            int gamma, double rgb[3])
{
    for (int c = 0; c < 3; ++c) {
        double v = orc_reflected_intensity(n_film, 1.0, d_nm, cos_i, lambdas[c]);
        if (gamma) v = pow(v, 1.0 / 2.2);
        rgb[c] = clamp01(v);
    }
}

/* generate_color_lookup_table (interference.rs:461-477) */
void orc_lut_1d(double n_film, const double lambdas[3], double tmin, double tmax, int n,
                double cos_theta, double* out)
{
    double step = (tmax - tmin) / (double)(n - 1);
    for (int i = 0; i < n; ++i) {
        double t = tmin + (double)i * step;
        orc_interference_color(n_film, lambdas, t, cos_theta, 1, out + 3 * (size_t)i);
    }
}

/* generate_2d_color_lookup_table (interference.rs:491-512): out[t][a][rgb] */
void orc_lut_2d(double n_film, const double lambdas[3], double tmin, double tmax, int nt, int na,
                int gamma, double* out)
{
    double thickness_step = (tmax - tmin) / (double)(nt - 1 > 1 ? nt - 1 : 1);
    double angle_step = 1.0 / (double)(na - 1 > 1 ? na - 1 : 1);
    for (int ti = 0; ti < nt; ++ti) {
        double thickness = tmin + (double)ti * thickness_step;
        for (int ai = 0; ai < na; ++ai) {
            double cos_theta = (double)ai * angle_step;
            orc_interference_color(n_film, lambdas, thickness, cos_theta, gamma,
                                   out + 3 * ((size_t)ti * na + ai));
        }
    }
}

/* ---------------- FP32 runtime LUT (interference_lut.rs) ----------------- */

#define LUT_T 256
#define LUT_A 64
#define LUT_MAX_NM 2000.0f
#define PI_F 3.14159265358979323846f

static float maxf_(float a, float b) { return a > b ? a : (b != b ? a : b); }

/* snells_law (interference_lut.rs:96-100) */
static float lut_snell(float cos_i, float n)
{
    float sin_i = sqrtf(maxf_(1.0f - cos_i * cos_i, 0.0f));
    float sin_t = sin_i / n;
    return sqrtf(maxf_(1.0f - sin_t * sin_t, 0.0f));
}

/* fresnel_unpolarized (interference_lut.rs:103-120) */
static float lut_fresnel(float cos_i, float cos_t, float n1, float n2)
{
    float n1_cos_i = n1 * cos_i;
    float n2_cos_t = n2 * cos_t;
    float r_s = (n1_cos_i - n2_cos_t) / maxf_(n1_cos_i + n2_cos_t, 0.0001f);
    float n2_cos_i = n2 * cos_i;
    float n1_cos_t = n1 * cos_t;
    float r_p = (n2_cos_i - n1_cos_t) / maxf_(n2_cos_i + n1_cos_t, 0.0001f);
    return (r_s * r_s + r_p * r_p) * 0.5f;
}

/* airy_interference (interference_lut.rs:123-133) */
static float lut_airy(float phase, float r)
{
    float one_minus_r = maxf_(1.0f - r, 0.001f);
    float f = 4.0f * r / (one_minus_r * one_minus_r);
    float s = sinf(phase * 0.5f);
    float s2 = s * s;
    float num = f * s2;
    float den = 1.0f + f * s2;
    return num / maxf_(den, 0.001f);
}

/* gaussian (interference_lut.rs:151-155) */
static float lut_gauss(float x, float mean, float sigma)
{
    float t = (x - mean) / sigma;
    return expf(-0.5f * t * t);
}

/* cie_color_matching (interference_lut.rs:136-148) */
static void lut_cie(float w, float out[3])
{
    float x = 1.056f * lut_gauss(w, 599.8f, 37.9f) + 0.362f * lut_gauss(w, 442.0f, 16.0f)
              - 0.065f * lut_gauss(w, 501.1f, 20.4f);
    float y = 0.821f * lut_gauss(w, 568.8f, 46.9f) + 0.286f * lut_gauss(w, 530.9f, 31.1f);
    float z = 1.217f * lut_gauss(w, 437.0f, 11.8f) + 0.681f * lut_gauss(w, 459.0f, 26.0f);
    out[0] = maxf_(x, 0.0f);
    out[1] = maxf_(y, 0.0f);
    out[2] = maxf_(z, 0.0f);
}

static const float LUT_W[7] = {400.0f, 450.0f, 500.0f, 550.0f, 600.0f, 650.0f, 700.0f};

/* The reflectance stage of thin_film_interference (interference_lut.rs:54-72):
 * refraction cosine, unpolarised Fresnel reflectance R, the coefficient of
 * finesse F of airy_interference and the Airy reflectance a[k] at the seven
 * LUT wavelengths. */
void orc_lut_airy_terms(float t_nm, float cos_theta, float n_film, float* cos_t_out, float* r_out,
                        float* f_out, float a[7])
{
    float cos_t = lut_snell(cos_theta, n_film);
    float fres = lut_fresnel(cos_theta, cos_t, 1.0f, n_film);
    float optical_path = 2.0f * n_film * t_nm * cos_t;
    for (int k = 0; k < 7; ++k) {
        float phase = 2.0f * PI_F * optical_path / LUT_W[k] + PI_F;
        a[k] = lut_airy(phase, fres);
    }
    float one_minus_r = maxf_(1.0f - fres, 0.001f);
    if (cos_t_out) *cos_t_out = cos_t;
    if (r_out) *r_out = fres;
    if (f_out) *f_out = 4.0f * fres / (one_minus_r * one_minus_r);
}

/* thin_film_interference (interference_lut.rs:45-93) + xyz_to_rgb (158-163) */
void orc_lut_thin_film(float t_nm, float cos_theta, float n_film, float intensity, float rgb[3])
{
    float a[7];
    orc_lut_airy_terms(t_nm, cos_theta, n_film, 0, 0, 0, a);
    float xyz[3] = {0.0f, 0.0f, 0.0f};
    for (int k = 0; k < 7; ++k) {
        float cie[3];
        lut_cie(LUT_W[k], cie);
        xyz[0] += cie[0] * a[k];
        xyz[1] += cie[1] * a[k];
        xyz[2] += cie[2] * a[k];
    }
    xyz[0] /= 7.0f;
    xyz[1] /= 7.0f;
    xyz[2] /= 7.0f;
    float r = 3.2404542f * xyz[0] - 1.5371385f * xyz[1] - 0.4985314f * xyz[2];
    float g = -0.969266f * xyz[0] + 1.8760108f * xyz[1] + 0.0415560f * xyz[2];
    float b = 0.0556434f * xyz[0] - 0.2040259f * xyz[1] + 1.0572252f * xyz[2];
    rgb[0] = r * intensity;
    rgb[1] = g * intensity;
    rgb[2] = b * intensity;
}

/* Pre-quantisation LUT, angle-major [64][256][3] (interference_lut.rs:22-31). */
void orc_lut_rgb_f32(float n_film, float intensity, float* out)
{
    for (uint32_t a = 0; a < LUT_A; ++a)
        for (uint32_t t = 0; t < LUT_T; ++t) {
            float t_nm = ((float)t / (float)(LUT_T - 1)) * LUT_MAX_NM;
            float cos_theta = (float)a / (float)(LUT_A - 1);
            orc_lut_thin_film(t_nm, cos_theta, n_film, intensity, out + 3 * ((size_t)a * LUT_T + t));
        }
}

/* `(x.clamp(0,1) * 255.0) as u8`: clamp, scale, truncate; NaN -> 0. */
static uint8_t quant_u8(float v)
{
    if (v != v) return 0;
    float c = v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);
    float s = c * 255.0f;
    return (uint8_t)s;
}

/* generate_interference_lut (interference_lut.rs:18-41) */
void orc_lut_rgba8(float n_film, float intensity, uint8_t* out)
{
    float rgb[3];
    for (uint32_t a = 0; a < LUT_A; ++a)
        for (uint32_t t = 0; t < LUT_T; ++t) {
            float t_nm = ((float)t / (float)(LUT_T - 1)) * LUT_MAX_NM;
            float cos_theta = (float)a / (float)(LUT_A - 1);
            orc_lut_thin_film(t_nm, cos_theta, n_film, intensity, rgb);
            uint8_t* px = out + 4 * ((size_t)a * LUT_T + t);
            px[0] = quant_u8(rgb[0]);
            px[1] = quant_u8(rgb[1]);
            px[2] = quant_u8(rgb[2]);
            px[3] = 255;
        }
}

/* ---------------- 81-wavelength spectral RGB (FP64) ---------------------- */

/* Unit normal of grid point (theta, phi), y up so that theta = 0 is the top of
 * the bubble (drainage runs towards theta = pi). */
static void sphere_normal(double theta, double phi, double n[3])
{
    n[0] = sin(theta) * cos(phi);
    n[1] = cos(theta);
    n[2] = sin(theta) * sin(phi);
}

/* |dot(normalize(cam - p), n)| with p = R n (bubble.wgsl:491-494). */
double orc_view_cos(double theta, double phi, double radius, const double cam[3])
{
    double n[3];
    sphere_normal(theta, phi, n);
    double v[3] = {cam[0] - radius * n[0], cam[1] - radius * n[1], cam[2] - radius * n[2]};
    double len = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    return fabs((v[0] * n[0] + v[1] * n[1] + v[2] * n[2]) / len);
}

/* rgb_c = (1/81) Σ_k Airy(h, cosθ_view, λ_k) cmf[k][c], λ_k = 380 + 5k nm,
 * with the Snell / Fresnel / Airy forms of interference_lut.rs:54-72, at one
 * point (theta, phi) of thickness h (m). */
static void spectral_point(double theta, double phi, double h, double radius, const double cam[3],
                           double n_film, const double* cmf, double out[3])
{
    double cos_i = orc_view_cos(theta, phi, radius, cam);
    double d_nm = h * 1e9;
    double sin_i = sqrt(fmax(1.0 - cos_i * cos_i, 0.0));
    double sin_t = sin_i / n_film;
    double cos_t = sqrt(fmax(1.0 - sin_t * sin_t, 0.0));
    double rs = (cos_i - n_film * cos_t) / fmax(cos_i + n_film * cos_t, 1e-4);
    double rp = (n_film * cos_i - cos_t) / fmax(n_film * cos_i + cos_t, 1e-4);
    double R = (rs * rs + rp * rp) * 0.5;
    double omr = fmax(1.0 - R, 1e-3);
    double F = 4.0 * R / (omr * omr);
    double op = 2.0 * n_film * d_nm * cos_t;
    double acc[3] = {0.0, 0.0, 0.0};
    for (int k = 0; k < 81; ++k) {
        double lam = 380.0 + 5.0 * (double)k;
        double phase = 2.0 * ORC_PI * op / lam + ORC_PI;
        double s = sin(0.5 * phase);
        double fs2 = F * s * s;
        double I = fs2 / fmax(1.0 + fs2, 1e-3);
        for (int cc = 0; cc < 3; ++cc) acc[cc] += I * cmf[3 * k + cc];
    }
    for (int cc = 0; cc < 3; ++cc) out[cc] = acc[cc] / 81.0;
}

void orc_spectral_rgb(const double* h, int n_theta, int n_phi, double radius, const double cam[3],
                      double n_film, const double* cmf, double* rgb)
{
    const double dtheta = ORC_PI / (double)(n_theta - 1);
    const double dphi = 2.0 * ORC_PI / (double)n_phi;
    const size_t plane = (size_t)n_theta * n_phi;
    for (int i = 0; i < n_theta; ++i)
        for (int j = 0; j < n_phi; ++j) {
            size_t c = (size_t)i * n_phi + j;
            double out[3];
            spectral_point((double)i * dtheta, (double)j * dphi, h[c], radius, cam, n_film, cmf, out);
            for (int cc = 0; cc < 3; ++cc) rgb[cc * plane + c] = out[cc];
        }
}

void orc_spectral_cells(const double* h, int n_theta, int n_phi, const int64_t* cells, int n_cells,
                        double radius, const double cam[3], double n_film, const double* cmf,
                        double* rgb)
{
    const double dtheta = ORC_PI / (double)(n_theta - 1);
    const double dphi = 2.0 * ORC_PI / (double)n_phi;
#pragma omp parallel for schedule(static) num_threads(orc_threads) if (orc_threads > 1)
    for (int k = 0; k < n_cells; ++k) {
        int64_t c = cells[k];
        int i = (int)(c / n_phi), j = (int)(c % n_phi);
        spectral_point((double)i * dtheta, (double)j * dphi, h[c], radius, cam, n_film, cmf, rgb + 3 * k);
    }
}

/* Thickness field rendered through the RGBA8 runtime LUT (bubble.wgsl:425-439
 * sampled with the linear clamp-to-edge sampler of pipeline.rs:350-359), one
 * pixel per grid point; FP64 texture coordinates and weights, output
 * quantised by round-to-nearest. */
void orc_render_lut(const double* h, int n_theta, int n_phi, double radius, const double cam[3],
                    const uint8_t* lut, double intensity, uint8_t* out)
{
    const double dth = ORC_PI / (double)(n_theta - 1), dph = 2.0 * ORC_PI / (double)n_phi;
    for (int i = 0; i < n_theta; ++i)
        for (int j = 0; j < n_phi; ++j) {
            const size_t c = (size_t)i * n_phi + j;
            const double cv = orc_view_cos(i * dth, j * dph, radius, cam);
            double u = h[c] * 1e9 / 2000.0, v = cv;
            u = u < 0.0 ? 0.0 : (u > 1.0 ? 1.0 : u);
            v = v < 0.0 ? 0.0 : (v > 1.0 ? 1.0 : v);
            const double x = u * 256.0 - 0.5, y = v * 64.0 - 0.5;
            const double x0 = floor(x), y0 = floor(y), fx = x - x0, fy = y - y0;
            int i0 = (int)x0, i1 = (int)x0 + 1, j0 = (int)y0, j1 = (int)y0 + 1;
            i0 = i0 < 0 ? 0 : (i0 > 255 ? 255 : i0);
            i1 = i1 < 0 ? 0 : (i1 > 255 ? 255 : i1);
            j0 = j0 < 0 ? 0 : (j0 > 63 ? 63 : j0);
            j1 = j1 < 0 ? 0 : (j1 > 63 ? 63 : j1);
            for (int ch = 0; ch < 3; ++ch) {
                const double t00 = lut[((size_t)j0 * 256 + i0) * 4 + ch] / 255.0;
                const double t10 = lut[((size_t)j0 * 256 + i1) * 4 + ch] / 255.0;
                const double t01 = lut[((size_t)j1 * 256 + i0) * 4 + ch] / 255.0;
                const double t11 = lut[((size_t)j1 * 256 + i1) * 4 + ch] / 255.0;
                double val = ((t00 * (1 - fx) + t10 * fx) * (1 - fy) + (t01 * (1 - fx) + t11 * fx) * fy) * intensity;
                val = val < 0.0 ? 0.0 : (val > 1.0 ? 1.0 : val);
                out[c * 4 + ch] = (uint8_t)floor(val * 255.0 + 0.5);
            }
            out[c * 4 + 3] = 255;
        }
}
