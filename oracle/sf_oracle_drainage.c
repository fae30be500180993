/*
 * sf_oracle_drainage.c -- FP64 restatement of the reference drainage solver.
 *
 * TEST INFRASTRUCTURE ONLY (see sf_oracle.h).  Compiled with -O2
 * -ffp-contract=off; each expression keeps the operand order of the Rust /
 * WGSL source it restates so that the FP64 rounding sequence is the same.
 *
 * Unified semantics (one code path, selected by flag bits):
 *   h part    src/physics/drainage.rs:298-392   (gravity h^3 sink, surface
 *             Laplacian, rupture skip, clamp >= 0, pole ring mean, TFE fronts)
 *   Γ part    src/render/shaders/drainage.wgsl:150-201 (Marangoni term on h,
 *             surfactant diffusion with optional thickness coupling, clamp [0,1])
 *   poles     ORC_POLES_COPY_H selects the WGSL copy-through for h
 *             (drainage.wgsl:85-89) instead of drainage.rs:480-503;
 *             ORC_POLES_AVG_GAMMA selects the ring mean for Γ instead of the
 *             WGSL copy-through.
 */
#include "sf_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ORC_PI        3.14159265358979323846
#define ORC_FRAC_PI_2 1.57079632679489661923
#define ORC_MAX_FRONTS 8

struct orc_sim {
    orc_params p;
    int nt, np;
    double dtheta, dphi;
    double *h, *g, *sh, *sg;        /* state and scratch copies (drainage.rs:216, 311) */
    double current_time;
    orc_front fronts[ORC_MAX_FRONTS + 1];
    int n_fronts;
    int tfe_enabled;
    double next_tfe_spawn_time;
};

/* SimulationConfig::default() (config.rs:305-316) fed through
 * DrainageSimulator::new (drainage.rs:263-288); Γ/Marangoni constants from
 * DrainageParams::default() (gpu_drainage.rs:46-68). */
void orc_params_cpu_default(orc_params* p)
{
    p->gravity = 9.81;
    p->viscosity = 0.001;
    p->density = 1000.0;
    p->diffusion_coeff = 1e-9;
    p->bubble_radius = 0.05 / 2.0;
    p->critical_thickness = 10.0 * 1e-9;
    p->gamma_air = 0.072;
    p->gamma_reduction = 0.045;
    p->surfactant_diffusion = 1e-9;
    p->marangoni_coeff = 0.01;
    p->sin_guard = 1e-10;
    p->tfe_thickness_ratio = 0.85;
    p->tfe_first_spawn = 0.5;
    p->tfe_spawn_interval = 2.0;
    p->flags = ORC_TFE;
}

/* ThicknessField::new (drainage.rs:121-133) */
orc_sim* orc_create(const orc_params* p, int n_theta, int n_phi, double h0, double gamma0)
{
    if (n_theta < 3 || n_phi < 1) return NULL;
    orc_sim* s = (orc_sim*)calloc(1, sizeof(orc_sim));
    size_t n = (size_t)n_theta * (size_t)n_phi;
    s->p = *p;
    s->nt = n_theta;
    s->np = n_phi;
    s->dtheta = ORC_PI / (double)(n_theta - 1);
    s->dphi = 2.0 * ORC_PI / (double)n_phi;
    s->h = (double*)malloc(n * sizeof(double));
    s->g = (double*)malloc(n * sizeof(double));
    s->sh = (double*)malloc(n * sizeof(double));
    s->sg = (double*)malloc(n * sizeof(double));
    for (size_t k = 0; k < n; ++k) { s->h[k] = h0; s->g[k] = gamma0; }
    s->current_time = 0.0;
    s->n_fronts = 0;
    s->tfe_enabled = (p->flags & ORC_TFE) != 0;
    s->next_tfe_spawn_time = p->tfe_first_spawn;
    return s;
}

void orc_destroy(orc_sim* s)
{
    if (!s) return;
    free(s->h); free(s->g); free(s->sh); free(s->sg);
    free(s);
}

void orc_set_params(orc_sim* s, const orc_params* p)
{
    s->p = *p;
    s->tfe_enabled = (p->flags & ORC_TFE) != 0;
    if (!s->tfe_enabled) s->n_fronts = 0;
}

void orc_set_fields(orc_sim* s, const double* h, const double* g)
{
    size_t n = (size_t)s->nt * s->np;
    if (h) memcpy(s->h, h, n * sizeof(double));
    if (g) memcpy(s->g, g, n * sizeof(double));
}

void orc_get_fields(const orc_sim* s, double* h, double* g)
{
    size_t n = (size_t)s->nt * s->np;
    if (h) memcpy(h, s->h, n * sizeof(double));
    if (g) memcpy(g, s->g, n * sizeof(double));
}

/* TfeFront::contains (drainage.rs:55-71) */
static int front_contains(const orc_front* f, double theta, double phi)
{
    double dist_from_equator = fabs(theta - ORC_FRAC_PI_2);
    if (dist_from_equator > f->extent) return 0;
    double phi_diff = fabs(phi - f->phi_center);
    if (phi_diff > ORC_PI) phi_diff = 2.0 * ORC_PI - phi_diff;
    return phi_diff < f->width / 2.0;
}

/* DrainageSimulator::spawn_tfe_front (drainage.rs:464-477) */
static void spawn_tfe_front(orc_sim* s)
{
    double t = s->current_time;
    orc_front f;
    f.phi_center = fabs(sin(t * 7.89 + 1.23)) * 2.0 * ORC_PI;
    f.width = 0.3 + fabs(sin(t * 3.21)) * 0.4;
    f.velocity = 0.05 + fabs(cos(t * 5.67)) * 0.05;
    f.extent = 0.0;
    s->fronts[s->n_fronts++] = f;
    if (s->n_fronts > ORC_MAX_FRONTS) {
        memmove(&s->fronts[0], &s->fronts[1], (size_t)(s->n_fronts - 1) * sizeof(orc_front));
        s->n_fronts -= 1;
    }
}

/* DrainageSimulator::update_marginal_regeneration (drainage.rs:395-461) */
static void update_marginal_regeneration(orc_sim* s, double dt)
{
    s->next_tfe_spawn_time -= dt;
    if (s->next_tfe_spawn_time <= 0.0) {
        spawn_tfe_front(s);
        double random_factor = 0.5 + fabs(sin(s->current_time * 12.345));
        s->next_tfe_spawn_time = s->p.tfe_spawn_interval * random_factor;
    }
    for (int k = 0; k < s->n_fronts; ++k) {            /* TfeFront::advance (47-49) */
        double e = s->fronts[k].extent + s->fronts[k].velocity * dt;
        s->fronts[k].extent = e < ORC_FRAC_PI_2 ? e : ORC_FRAC_PI_2;
    }
    int w = 0;                                          /* retain (411) */
    for (int k = 0; k < s->n_fronts; ++k)
        if (s->fronts[k].extent < ORC_FRAC_PI_2 * 0.95) s->fronts[w++] = s->fronts[k];
    s->n_fronts = w;
    if (s->n_fronts == 0) return;

    const int nt = s->nt, np = s->np;
    const double dtheta = s->dtheta, dphi = s->dphi;
    long min_idx_all = nt, max_idx_all = 0;
    for (int k = 0; k < s->n_fronts; ++k) {
        double min_theta = ORC_FRAC_PI_2 - s->fronts[k].extent;   /* theta_range (76-79) */
        double max_theta = ORC_FRAC_PI_2 + s->fronts[k].extent;
        double a = floor(min_theta / dtheta);
        double b = ceil(max_theta / dtheta);
        long min_idx = a > 0.0 ? (long)a : 0;                  /* `as usize` saturates */
        long max_idx = b > 0.0 ? (long)b : 0;
        if (max_idx > nt - 1) max_idx = nt - 1;
        if (min_idx < min_idx_all) min_idx_all = min_idx;
        if (max_idx > max_idx_all) max_idx_all = max_idx;
    }
    if (min_idx_all > nt - 1) min_idx_all = nt - 1;

    double reduction_rate = (1.0 - s->p.tfe_thickness_ratio) * 0.5;
    for (long i = min_idx_all; i <= max_idx_all; ++i) {
        double theta = (double)i * dtheta;
        for (int j = 0; j < np; ++j) {
            double phi = (double)j * dphi;
            for (int k = 0; k < s->n_fronts; ++k) {
                if (front_contains(&s->fronts[k], theta, phi)) {
                    double current = s->h[(size_t)i * np + j];
                    s->h[(size_t)i * np + j] = current * (1.0 - reduction_rate * dt);
                    break;
                }
            }
        }
    }
}

/* Ring mean of one row (drainage.rs:484-488 / 495-499), sequential FP64 sum. */
/* Threads of the interior loop and of the sampled spectral cells; 1 (the
 * reference's sequential loop) unless a CPU baseline sets more. */
int orc_threads = 1;

void orc_set_threads(int n) { orc_threads = n < 1 ? 1 : n; }

static double ring_mean(const double* row, int np)
{
    double sum = 0.0;
    for (int j = 0; j < np; ++j) sum += row[j];
    return sum / (double)np;
}

/* DrainageSimulator::step (drainage.rs:298-392) with the Γ/Marangoni terms of
 * drainage_step (drainage.wgsl:150-201). */
void orc_step(orc_sim* s, double dt)
{
    const int nt = s->nt, np = s->np;
    const double dtheta = s->dtheta, dphi = s->dphi;
    const orc_params* p = &s->p;
    const int marangoni = (p->flags & ORC_MARANGONI) != 0;
    const int surfactant = (p->flags & ORC_SURFACTANT) != 0;
    const double drainage_coefficient = p->density * p->gravity / (3.0 * p->viscosity);
    const double diffusion_scaled = p->diffusion_coeff;
    const double R = p->bubble_radius;
    const size_t n = (size_t)nt * np;

    const double* sh = s->sh;
    const double* sg = s->sg;
    (void)n;

    /* Rows are independent once the scratch copy is complete: with more than
     * one thread (orc_set_threads) the copy and the interior rows are split
     * over threads, with the barrier of the first loop between them.  The
     * arithmetic of each cell is unchanged. */
#pragma omp parallel num_threads(orc_threads) if (orc_threads > 1)
    {
#pragma omp for schedule(static)
    for (int i = 0; i < nt; ++i) {
        memcpy(s->sh + (size_t)i * np, s->h + (size_t)i * np, (size_t)np * sizeof(double));
        memcpy(s->sg + (size_t)i * np, s->g + (size_t)i * np, (size_t)np * sizeof(double));
    }
#pragma omp for schedule(static)
    for (int i = 1; i < nt - 1; ++i) {
        double theta = (double)i * dtheta;
        double sin_theta = sin(theta);
        double cos_theta = cos(theta);
        double sin_theta_safe = fabs(sin_theta) < p->sin_guard ? copysign(p->sin_guard, sin_theta)
                                                                : sin_theta;
        for (int j = 0; j < np; ++j) {
            size_t c = (size_t)i * np + j;
            double hc = sh[c];
            if (hc < p->critical_thickness) continue;      /* rupture (331-333, wgsl 108-112) */

            double htm = sh[c - np];
            double htp = sh[c + np];
            int jm = (j + np - 1) % np;
            int jp = (j + 1) % np;
            double hpm = sh[(size_t)i * np + jm];
            double hpp = sh[(size_t)i * np + jp];

            double thickness_cubed = hc * hc * hc;
            double drainage_term = -drainage_coefficient * thickness_cubed * sin_theta;
            double radius_squared = R * R;
            double d2h_dtheta2 = (htp - 2.0 * hc + htm) / (dtheta * dtheta);
            double dh_dtheta = (htp - htm) / (2.0 * dtheta);
            double d2h_dphi2 = (hpp - 2.0 * hc + hpm) / (dphi * dphi);
            double laplacian = (d2h_dtheta2 + cos_theta / sin_theta_safe * dh_dtheta
                                + d2h_dphi2 / (sin_theta_safe * sin_theta_safe))
                               / radius_squared;
            double diffusion_term = diffusion_scaled * laplacian;
            double dh_dt = drainage_term + diffusion_term;

            double gc = sg[c];
            double gtm = sg[c - np], gtp = sg[c + np];
            double gpm = sg[(size_t)i * np + jm], gpp = sg[(size_t)i * np + jp];
            if (marangoni) {                                 /* wgsl 150-171 */
                double dc_dtheta = (gtp - gtm) / (2.0 * dtheta);
                double dc_dphi = (gpp - gpm) / (2.0 * dphi);
                double grad_c_theta = dc_dtheta / R;
                double grad_c_phi = dc_dphi / (R * sin_theta_safe);
                double grad_c_mag = sqrt(grad_c_theta * grad_c_theta + grad_c_phi * grad_c_phi);
                double marangoni_stress = p->gamma_reduction * grad_c_mag;
                double marangoni_term = p->marangoni_coeff * marangoni_stress * hc * hc;
                dh_dt = dh_dt + marangoni_term;
            }
            double new_h = hc + dt * dh_dt;
            s->h[c] = new_h > 0.0 ? new_h : 0.0;          /* .max(0.0) (378) */

            if (surfactant) {                                /* wgsl 180-201 */
                double d2c_dtheta2 = (gtp - 2.0 * gc + gtm) / (dtheta * dtheta);
                double dc_dtheta = (gtp - gtm) / (2.0 * dtheta);
                double d2c_dphi2 = (gpp - 2.0 * gc + gpm) / (dphi * dphi);
                double c_laplacian = (d2c_dtheta2 + cos_theta / sin_theta_safe * dc_dtheta
                                      + d2c_dphi2 / (sin_theta_safe * sin_theta_safe))
                                     / radius_squared;
                double dc_dt = p->surfactant_diffusion * c_laplacian;
                if (marangoni) {
                    double coupling = 0.001 * (laplacian * R * R);
                    dc_dt += coupling * gc;
                }
                double new_c = gc + dt * dc_dt;
                new_c = new_c < 0.0 ? 0.0 : (new_c > 1.0 ? 1.0 : new_c);
                s->g[c] = new_c;
            }
        }
    }
    }

    /* update_poles (480-503) or copy-through (wgsl 85-89). */
    if (!(p->flags & ORC_POLES_COPY_H)) {
        double top = ring_mean(s->h + (size_t)1 * np, np);
        for (int j = 0; j < np; ++j) s->h[j] = top;
        double bottom = ring_mean(s->h + (size_t)(nt - 2) * np, np);
        for (int j = 0; j < np; ++j) s->h[(size_t)(nt - 1) * np + j] = bottom;
    }
    if ((p->flags & ORC_POLES_AVG_GAMMA) && surfactant) {
        double top = ring_mean(s->g + (size_t)1 * np, np);
        for (int j = 0; j < np; ++j) s->g[j] = top;
        double bottom = ring_mean(s->g + (size_t)(nt - 2) * np, np);
        for (int j = 0; j < np; ++j) s->g[(size_t)(nt - 1) * np + j] = bottom;
    }

    if (s->tfe_enabled) update_marginal_regeneration(s, dt);   /* 387-389 */
    s->current_time += dt;                                      /* 391 */
}

/* Rust f64::rem_euclid: r = a % b; r < 0 ? r + |b| : r */
static double rem_euclid(double a, double b)
{
    double r = fmod(a, b);
    return r < 0.0 ? r + fabs(b) : r;
}

/* `f as usize` for a non-negative-or-NaN double: saturating, NaN -> 0. */
static size_t as_usize(double x)
{
    if (!(x > 0.0)) return 0;
    if (x >= 18446744073709551615.0) return (size_t)-1;
    return (size_t)x;
}

/* DrainageSimulator::get_thickness (drainage.rs:516-552) */
double orc_get_thickness(const orc_sim* s, double theta, double phi)
{
    const size_t nt = (size_t)s->nt, np = (size_t)s->np;
    double theta_clamped = theta < 0.0 ? 0.0 : (theta > ORC_PI ? ORC_PI : theta);
    double phi_normalized = rem_euclid(phi, 2.0 * ORC_PI);
    double theta_continuous = theta_clamped / s->dtheta;
    double phi_continuous = phi_normalized / s->dphi;
    size_t ti_lo = as_usize(floor(theta_continuous));
    if (ti_lo > nt - 2) ti_lo = nt - 2;
    size_t pj_lo = as_usize(floor(phi_continuous)) % np;
    size_t ti_hi = ti_lo + 1;
    size_t pj_hi = (pj_lo + 1) % np;
    double tf = theta_continuous - (double)ti_lo;
    double pf = phi_continuous - (double)pj_lo;
    double v00 = s->h[ti_lo * np + pj_lo];
    double v01 = s->h[ti_lo * np + pj_hi];
    double v10 = s->h[ti_hi * np + pj_lo];
    double v11 = s->h[ti_hi * np + pj_hi];
    double v0 = v00 * (1.0 - pf) + v01 * pf;
    double v1 = v10 * (1.0 - pf) + v11 * pf;
    return v0 * (1.0 - tf) + v1 * tf;
}

/* ThicknessField::min_thickness / max_thickness (drainage.rs:180-193):
 * fold with f64::min / f64::max, which ignore NaN like fmin / fmax. */
double orc_min_thickness(const orc_sim* s)
{
    double m = INFINITY;
    size_t n = (size_t)s->nt * s->np;
    for (size_t k = 0; k < n; ++k) m = fmin(m, s->h[k]);
    return m;
}

double orc_max_thickness(const orc_sim* s)
{
    double m = -INFINITY;
    size_t n = (size_t)s->nt * s->np;
    for (size_t k = 0; k < n; ++k) m = fmax(m, s->h[k]);
    return m;
}

int orc_has_critical_region(const orc_sim* s)           /* drainage.rs:565-567 */
{
    return orc_min_thickness(s) < s->p.critical_thickness;
}

double orc_current_time(const orc_sim* s) { return s->current_time; }

/* DrainageSimulator::reset (drainage.rs:575-583); Γ is reset to gamma0 as
 * GPUDrainageSimulator::reset does (gpu_drainage.rs:423-426). */
void orc_reset(orc_sim* s, double h0, double gamma0)
{
    size_t n = (size_t)s->nt * s->np;
    for (size_t k = 0; k < n; ++k) { s->h[k] = h0; s->g[k] = gamma0; }
    s->current_time = 0.0;
    s->n_fronts = 0;
    s->next_tfe_spawn_time = s->p.tfe_first_spawn;
}

void orc_set_marginal_regeneration(orc_sim* s, int enabled)   /* 588-593 */
{
    s->tfe_enabled = enabled != 0;
    if (enabled) s->p.flags |= ORC_TFE; else s->p.flags &= ~ORC_TFE;
    if (!enabled) s->n_fronts = 0;
}

void orc_set_tfe_thickness_ratio(orc_sim* s, double ratio)   /* 601-603: clamp(0.5, 0.95) */
{
    s->p.tfe_thickness_ratio = ratio < 0.5 ? 0.5 : (ratio > 0.95 ? 0.95 : ratio);
}

int orc_active_tfe_count(const orc_sim* s) { return s->n_fronts; }

int orc_get_fronts(const orc_sim* s, orc_front* out, int max_out)
{
    int n = s->n_fronts < max_out ? s->n_fronts : max_out;
    for (int k = 0; k < n; ++k) out[k] = s->fronts[k];
    return s->n_fronts;
}

double orc_next_spawn_time(const orc_sim* s) { return s->next_tfe_spawn_time; }

void orc_tfe_mask(const orc_sim* s, uint8_t* mask)
{
    for (int i = 0; i < s->nt; ++i) {
        double theta = (double)i * s->dtheta;
        for (int j = 0; j < s->np; ++j) {
            double phi = (double)j * s->dphi;
            uint8_t m = 0;
            for (int k = 0; k < s->n_fronts && !m; ++k)
                m = (uint8_t)front_contains(&s->fronts[k], theta, phi);
            mask[(size_t)i * s->np + j] = m;
        }
    }
}

void orc_rupture_mask(const orc_sim* s, uint8_t* mask)
{
    size_t n = (size_t)s->nt * s->np;
    for (size_t k = 0; k < n; ++k) mask[k] = s->h[k] < s->p.critical_thickness;
}

/* Spherical-band area of grid row i (build-defined, no reference counterpart):
 * interior rows span [theta_i - dθ/2, theta_i + dθ/2]; the pole rows are the
 * caps [0, dθ/2] and [π - dθ/2, π].  Summed over all rows and n_phi columns
 * the areas telescope to 4πR². */
double orc_cell_area(int n_theta, int n_phi, double radius, int i)
{
    double dtheta = ORC_PI / (double)(n_theta - 1);
    double dphi = 2.0 * ORC_PI / (double)n_phi;
    double theta = (double)i * dtheta;
    double lo = i == 0 ? 0.0 : theta - 0.5 * dtheta;
    double hi = i == n_theta - 1 ? ORC_PI : theta + 0.5 * dtheta;
    return radius * radius * (cos(lo) - cos(hi)) * dphi;
}

/* Film volume M = Σ h_ij A_i (m^3).  Pinned by ΣA = 4πR² and, for uniform h,
 * M = 4πR²h. */
double orc_mass(const orc_sim* s)
{
    double m = 0.0;
    for (int i = 0; i < s->nt; ++i) {
        double a = orc_cell_area(s->nt, s->np, s->p.bubble_radius, i);
        double row = 0.0;
        for (int j = 0; j < s->np; ++j) row += s->h[(size_t)i * s->np + j];
        m += row * a;
    }
    return m;
}

/* Forward-Euler stability limit (build-defined):
 *   dt_cfl = 2 / max_ij( 4 D (1/dθ² + 1/(sin²θ_i dφ²)) / R² + 3 c h_ij² |sin θ_i| )
 * over interior rows, D = max(D_h, D_s when Γ evolves), c = ρg/(3η).  For h -> 0
 * it reduces to R² / (2 D max_i(1/dθ² + 1/(sin²θ_i dφ²))). */
double orc_cfl_dt(const orc_sim* s)
{
    const orc_params* p = &s->p;
    double D = p->diffusion_coeff;
    if ((p->flags & ORC_SURFACTANT) && p->surfactant_diffusion > D) D = p->surfactant_diffusion;
    double c = p->density * p->gravity / (3.0 * p->viscosity);
    double R2 = p->bubble_radius * p->bubble_radius;
    double lam_max = 0.0;
    for (int i = 1; i < s->nt - 1; ++i) {
        double theta = (double)i * s->dtheta;
        double st = sin(theta);
        double ss = fabs(st) < p->sin_guard ? copysign(p->sin_guard, st) : st;
        double grid = 4.0 * D * (1.0 / (s->dtheta * s->dtheta) + 1.0 / (ss * ss * s->dphi * s->dphi)) / R2;
        for (int j = 0; j < s->np; ++j) {
            double h = s->h[(size_t)i * s->np + j];
            double lam = grid + 3.0 * c * h * h * fabs(st);
            if (lam > lam_max) lam_max = lam;
        }
    }
    return lam_max > 0.0 ? 2.0 / lam_max : INFINITY;
}
