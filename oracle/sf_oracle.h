/*
 * sf_oracle.h -- FP64 CPU restatement of the reference drainage and
 * interference physics.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links this code;
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * it, and there only as the checker or as the timed CPU baseline.
 *
 * Reference (pjt222/soap_bubble_simulation, Rust crate soap-bubble-sim 0.1.0):
 *   drainage      src/physics/drainage.rs:96-613
 *   surfactant    src/render/shaders/drainage.wgsl:72-224 (Γ and Marangoni terms)
 *   GPU params    src/render/gpu_drainage.rs:12-68
 *   interference  src/physics/interference.rs:197-512   (FP64, 3 wavelengths)
 *   runtime LUT   src/render/interference_lut.rs:18-163  (FP32, 7 wavelengths)
 *
 * Every function is compiled with -O2 -ffp-contract=off and no fast-math so
 * that each arithmetic operation rounds the way the Rust source does.
 *
 * Build-defined quantities with no reference counterpart (mass, CFL time step,
 * spectral 81-wavelength RGB) are defined here from their closed forms; the
 * file that implements each one says which closed form pins it.
 */
#ifndef SF_ORACLE_H
#define SF_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Flag bits: identical values to SF_FLAG_* in include/soapfilm.h. */
#define ORC_MARANGONI        (1u << 0)
#define ORC_SURFACTANT       (1u << 1)
#define ORC_TFE              (1u << 2)
#define ORC_POLES_COPY_H     (1u << 3)
#define ORC_POLES_AVG_GAMMA  (1u << 4)

typedef struct orc_params {
    double gravity;              /* m/s^2                                  */
    double viscosity;            /* Pa s                                   */
    double density;              /* kg/m^3                                 */
    double diffusion_coeff;      /* D_h, m^2/s (drainage.rs:277)           */
    double bubble_radius;        /* m                                      */
    double critical_thickness;   /* m (drainage.rs:278)                    */
    double gamma_air;            /* N/m (drainage.wgsl:25)                 */
    double gamma_reduction;      /* drainage.wgsl:26                       */
    double surfactant_diffusion; /* D_s, m^2/s (drainage.wgsl:27)          */
    double marangoni_coeff;      /* drainage.wgsl:28                       */
    double sin_guard;            /* 1e-10 CPU (drainage.rs:320), 1e-4 WGSL */
    double tfe_thickness_ratio;  /* drainage.rs:283                        */
    double tfe_first_spawn;      /* drainage.rs:285                        */
    double tfe_spawn_interval;   /* drainage.rs:286                        */
    uint32_t flags;
} orc_params;

typedef struct orc_front {
    double phi_center, extent, width, velocity;   /* drainage.rs:23-33 */
} orc_front;

typedef struct orc_sim orc_sim;

/* ---- drainage ---------------------------------------------------------- */
void      orc_params_cpu_default(orc_params* p);
orc_sim*  orc_create(const orc_params* p, int n_theta, int n_phi, double h0, double gamma0);
void      orc_destroy(orc_sim* s);
void      orc_set_params(orc_sim* s, const orc_params* p);
void      orc_set_fields(orc_sim* s, const double* h, const double* gamma);
void      orc_get_fields(const orc_sim* s, double* h, double* gamma);
void      orc_step(orc_sim* s, double dt);
double    orc_get_thickness(const orc_sim* s, double theta, double phi);
double    orc_min_thickness(const orc_sim* s);
double    orc_max_thickness(const orc_sim* s);
int       orc_has_critical_region(const orc_sim* s);
double    orc_current_time(const orc_sim* s);
void      orc_reset(orc_sim* s, double h0, double gamma0);
void      orc_set_marginal_regeneration(orc_sim* s, int enabled);
void      orc_set_tfe_thickness_ratio(orc_sim* s, double ratio);
int       orc_active_tfe_count(const orc_sim* s);
int       orc_get_fronts(const orc_sim* s, orc_front* out, int max_out);
double    orc_next_spawn_time(const orc_sim* s);
/* Cell-wise TFE membership of the fronts that are active right now:
 * mask[i*n_phi+j] = 1 when some front contains (theta_i, phi_j). */
void      orc_tfe_mask(const orc_sim* s, uint8_t* mask);
/* Rupture mask of the current state: h < h_crit (drainage.rs:331). */
void      orc_rupture_mask(const orc_sim* s, uint8_t* mask);
double    orc_mass(const orc_sim* s);
double    orc_cfl_dt(const orc_sim* s);
double    orc_cell_area(int n_theta, int n_phi, double radius, int i);

/* ---- FP64 interference (interference.rs) -------------------------------- */
double orc_transmission_cos(double n_film, double n_medium, double cos_i);
double orc_optical_path_difference(double n_film, double d_nm, double cos_t, double lambda_nm);
double orc_phase_difference(double opd_nm, double lambda_nm);
void   orc_fresnel(double n_film, double n_medium, double cos_i,
                   double* rs, double* rp, double* ravg);
double orc_reflected_intensity(double n_film, double n_medium, double d_nm,
                               double cos_i, double lambda_nm);
void   orc_interference_color(double n_film, const double lambdas[3], double d_nm,
                              double cos_i, int gamma, double rgb[3]);
void   orc_lut_1d(double n_film, const double lambdas[3], double tmin, double tmax,
                  int n, double cos_theta, double* out /* n*3 */);
void   orc_lut_2d(double n_film, const double lambdas[3], double tmin, double tmax,
                  int nt, int na, int gamma, double* out /* nt*na*3 */);

/* ---- FP32 runtime LUT (interference_lut.rs) ------------------------------ */
/* Host threads of orc_step's interior rows and orc_spectral_cells (default 1). */
extern int orc_threads;
void   orc_set_threads(int n);
void   orc_lut_airy_terms(float t_nm, float cos_theta, float n_film, float* cos_t, float* r, float* f,
                          float a[7]);
void   orc_lut_thin_film(float t_nm, float cos_theta, float n_film, float intensity, float rgb[3]);
void   orc_lut_rgb_f32(float n_film, float intensity, float* out /* 64*256*3 */);
void   orc_lut_rgba8(float n_film, float intensity, uint8_t* out /* 64*256*4 */);

/* ---- spectral 81-wavelength RGB at per-cell view angle (FP64) ----------- */
double orc_view_cos(double theta, double phi, double radius, const double cam[3]);
void   orc_spectral_rgb(const double* h, int n_theta, int n_phi, double radius,
                        const double cam[3], double n_film, const double* cmf /* 81*3 */,
                        double* rgb /* 3 planes of n_theta*n_phi */);
/* The same at selected flat cell indices: rgb[k][3]. */
void   orc_spectral_cells(const double* h, int n_theta, int n_phi, const int64_t* cells,
                          int n_cells, double radius, const double cam[3], double n_film,
                          const double* cmf, double* rgb);

void   orc_render_lut(const double* h, int n_theta, int n_phi, double radius, const double cam[3],
                      const uint8_t* lut, double intensity, uint8_t* out /* n*4 */);

/* ---- caustic intensity (caustics_compute.wgsl:27-126), FP32 ---------------- */
void   orc_caustics(const float* h, int n_theta, int n_phi, float focal_length, float caustic_intensity,
                    float caustic_sharpness, float branch_threshold, float* out, uint8_t* branch);

#ifdef __cplusplus
}
#endif
#endif
