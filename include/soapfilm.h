/*
 * soapfilm.h -- C ABI of the MI355X (gfx950) soap-film physics library.
 *
 * This is the drop-in boundary for the data-parallel physics path of
 * pjt222/soap_bubble_simulation: the film-thickness drainage step on the
 * bubble's latitude-longitude grid and the thin-film interference evaluation
 * and LUT generation.  The reference has no FFI; each entry point below names
 * the Rust API it replaces (file:line in the reference crate).  All compute
 * runs in hand-written HIP kernels; the library fails with SF_ERR_DEVICE when
 * no GPU is present (there is no CPU fallback).
 *
 * Conventions
 *   - Every function returns SF_OK (0) or a negative SF_ERR_* code;
 *     sf_last_error() gives a message for the calling thread.
 *   - Grids are row-major [theta][phi]: index = i * n_phi + j,
 *     theta_i = i * pi / (n_theta - 1) (both poles included),
 *     phi_j = j * 2 pi / n_phi (drainage.rs:121-157).
 *   - Thickness is in metres (drainage.rs:103-105); LUT thickness in nm.
 *   - One sf_drainage context belongs to one host thread.  It owns its
 *     device buffers and one HIP stream; sf_drainage_step only enqueues
 *     work, every copy/query/stats call synchronises that stream.
 */
#ifndef SOAPFILM_H
#define SOAPFILM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------- */
#define SF_OK             0
#define SF_ERR_INVALID   -1   /* bad argument                                */
#define SF_ERR_DEVICE    -2   /* HIP runtime error or no GPU present         */
#define SF_ERR_NOMEM     -3   /* device allocation failed                    */
#define SF_ERR_INTERNAL  -4

/* ---- drainage flags ------------------------------------------------------ */
/* Marangoni term on h and thickness coupling on Γ (drainage.wgsl:150-171, 193-198). */
#define SF_FLAG_MARANGONI        (1u << 0)
/* Evolve the surfactant field Γ (drainage.wgsl:178-201). */
#define SF_FLAG_SURFACTANT       (1u << 1)
/* Marginal-regeneration fronts (drainage.rs:387-389, 395-477). */
#define SF_FLAG_TFE              (1u << 2)
/* h poles copied through (drainage.wgsl:85-89) instead of ring mean (drainage.rs:480-503). */
#define SF_FLAG_POLES_COPY_H     (1u << 3)
/* Γ poles set to the ring mean instead of copied through. */
#define SF_FLAG_POLES_AVG_GAMMA  (1u << 4)
/* Per-step FP64 film-volume (mass) and CFL reductions fused into the step. */
#define SF_FLAG_DIAG             (1u << 5)

/*
 * Physical and numerical parameters of one bubble.  Union of the GPU uniform
 * DrainageParams (gpu_drainage.rs:12-44) and the DrainageSimulator fields
 * (drainage.rs:212-256), in FP64; the kernels round them to FP32 where the
 * stencil computes in FP32.
 */
typedef struct sf_params {
    double dt;                    /* default step, s (gpu_drainage.rs:14)            */
    double gravity;               /* m/s^2                                           */
    double viscosity;             /* Pa s                                            */
    double density;               /* kg/m^3                                          */
    double diffusion_coeff;       /* D_h, m^2/s (drainage.rs:277)                    */
    double bubble_radius;         /* m                                               */
    double critical_thickness;    /* m; cells below it are frozen (drainage.rs:331)  */
    double gamma_air;             /* N/m (drainage.wgsl:25)                          */
    double gamma_reduction;       /* N/m per unit Γ (drainage.wgsl:26)               */
    double surfactant_diffusion;  /* D_s, m^2/s (drainage.wgsl:27)                   */
    double marangoni_coeff;       /* drainage.wgsl:28                                */
    double sin_guard;             /* |sin θ| floor: 1e-10 CPU, 1e-4 WGSL             */
    double tfe_thickness_ratio;   /* drainage.rs:283, clamped to [0.5, 0.95]         */
    double tfe_first_spawn;       /* s until the first front (drainage.rs:285)       */
    double tfe_spawn_interval;    /* s (drainage.rs:286)                             */
    double cfl_safety;            /* dt = cfl_safety * dt_cfl when step dt <= 0      */
    uint32_t flags;               /* SF_FLAG_*                                       */
    uint32_t rebase_interval;     /* steps between float-float renormalisations      */
} sf_params;

typedef struct sf_stats {
    double min_thickness;         /* over all cells incl. poles (drainage.rs:180-186) */
    double max_thickness;         /* drainage.rs:188-193                              */
    double mass;                  /* FP64 film volume  Σ h_ij A_i, m^3                */
    double dt_cfl;                /* FP64 forward-Euler stability limit, s            */
    double current_time;          /* s (drainage.rs:560)                              */
    int64_t n_ruptured;           /* cells with h < critical_thickness                */
    int32_t n_tfe;                /* active fronts (drainage.rs:611)                  */
    int32_t has_critical_region;  /* drainage.rs:565-567                              */
    int64_t steps;                /* steps taken since create/reset                   */
} sf_stats;

typedef struct sf_drainage sf_drainage;

/* Message for the last failing call on this thread ("" when none). */
const char* sf_last_error(void);
/* Library version string. */
const char* sf_version(void);
/* 1 when a HIP device is usable, 0 otherwise (never fails). */
int sf_device_available(void);

/* Defaults of DrainageSimulator::new(&SimulationConfig::default())
 * (drainage.rs:263-288, config.rs:305-316): 10 nm rupture, TFE on, h poles
 * averaged, Γ frozen, sin guard 1e-10. */
int sf_params_cpu_default(sf_params* out);
/* Defaults of DrainageParams::default() (gpu_drainage.rs:46-68): dt 1e-3,
 * 30 nm rupture, Γ evolving, poles copied through, sin guard 1e-4, no TFE. */
int sf_params_gpu_default(sf_params* out);

/* DrainageSimulator::new (drainage.rs:263-288) / GPUDrainageSimulator::new
 * (gpu_drainage.rs:118-331): uniform h = h0 (m) and Γ = gamma0.
 * Needs n_theta >= 3 and n_phi >= 1. */
int sf_drainage_create(const sf_params* params, int32_t n_theta, int32_t n_phi, double h0,
                       double gamma0, sf_drainage** out);
void sf_drainage_destroy(sf_drainage* ctx);

/* GPUDrainageSimulator::update_params / set_dt / set_physics /
 * set_bubble_radius / set_marangoni_enabled / set_marangoni_params
 * (gpu_drainage.rs:377-456). */
int sf_drainage_set_params(sf_drainage* ctx, const sf_params* params);
int sf_drainage_get_params(const sf_drainage* ctx, sf_params* out);

/* Upload h and Γ from host arrays of n_theta*n_phi values (either may be
 * NULL to keep the field).  The FP64 variant keeps h to ~48 bits. */
int sf_drainage_set_fields(sf_drainage* ctx, const float* h, const float* gamma);
int sf_drainage_set_fields_f64(sf_drainage* ctx, const double* h, const double* gamma);

/* DrainageSimulator::step (drainage.rs:298-392) repeated n_steps times, or
 * GPUDrainageSimulator::step's dispatch loop (gpu_drainage.rs:354-369).
 * dt > 0: fixed step.  dt <= 0: dt = cfl_safety * FP64 CFL estimate of the
 * state at the start of the call, held for the call's n_steps.  Asynchronous. */
int sf_drainage_step(sf_drainage* ctx, int32_t n_steps, double dt);

/* Download the current fields (either pointer may be NULL).  The FP32 form
 * rounds h to float; the FP64 form returns hi + lo. */
int sf_drainage_copy_field(sf_drainage* ctx, float* h, float* gamma);
int sf_drainage_copy_field_f64(sf_drainage* ctx, double* h, double* gamma);

/* DrainageSimulator::get_thickness (drainage.rs:516-552): bilinear query at
 * (theta, phi) radians; index and weight math in FP64 on the host, identical
 * to the reference; the four corner values are gathered on the device. */
int sf_drainage_get_thickness(sf_drainage* ctx, double theta, double phi, double* out);
/* Batched form of the same query (n points). */
int sf_drainage_get_thickness_batch(sf_drainage* ctx, int32_t n, const double* theta,
                                    const double* phi, double* out);

/* min/max (drainage.rs:180-193), FP64 mass and CFL, rupture count, time. */
int sf_drainage_stats(sf_drainage* ctx, sf_stats* out);
/* Allocation-free statistics of a caller-held device field.  No counterpart
 * in the reference: for callers that cannot allocate device memory (code
 * under HIP graph capture, serving loops, harnesses that forbid hipMalloc).
 * h = h_hi + h_lo (float-float; h_lo may be NULL for an FP32 field), row-major
 * [n_theta][n_phi] on the current device, with the grid convention and the
 * params of sf_drainage_create.  Writes *out — device memory — with what
 * sf_drainage_stats reports for such a field (min, max, FP64 film volume,
 * FP64 CFL step, rupture count, has_critical_region; time, fronts and steps
 * 0), asynchronously on `stream` (NULL: the null stream).  Calls no device
 * allocator: its row tables and partial sums live in one fixed library array
 * per device (about 0.6 MB, held until the process ends), so n_theta is at
 * most SF_NOALLOC_MAX_THETA (larger grids fail with SF_ERR_INVALID), and
 * calls that overlap on one device race on that array. */
#define SF_NOALLOC_MAX_THETA 16384
int sf_field_stats_noalloc(const sf_params* params, int32_t n_theta, int32_t n_phi, const float* h_hi,
                           const float* h_lo, void* stream, sf_stats* out);

/* DrainageSimulator::reset (drainage.rs:575-583) / GPUDrainageSimulator::reset
 * (gpu_drainage.rs:415-432): uniform h0 and gamma0, time 0, fronts cleared. */
int sf_drainage_reset(sf_drainage* ctx, double h0, double gamma0);

/* set_marginal_regeneration (drainage.rs:588-593), set_tfe_thickness_ratio
 * (601-603), active_tfe_count (611-613), current_time (560-562). */
int sf_drainage_set_marginal_regeneration(sf_drainage* ctx, int32_t enabled);
int sf_drainage_set_tfe_thickness_ratio(sf_drainage* ctx, double ratio);
int sf_drainage_active_tfe_count(const sf_drainage* ctx, int32_t* out);
int sf_drainage_current_time(const sf_drainage* ctx, double* out);

/* Per-step FP64 mass and CFL values recorded while SF_FLAG_DIAG is set: the
 * most recent min(n, recorded) entries, oldest first; *n_out = count. */
int sf_drainage_history(sf_drainage* ctx, int32_t n, double* mass, double* dt_cfl,
                        int32_t* n_out);

/* TFE membership of the fronts active after the last step: mask[i*n_phi+j] = 1
 * when a front contains cell (i, j) (TfeFront::contains, drainage.rs:55-71). */
int sf_drainage_tfe_mask(sf_drainage* ctx, uint8_t* mask);
/* Rupture mask of the current state: h < critical_thickness (drainage.rs:331). */
int sf_drainage_rupture_mask(sf_drainage* ctx, uint8_t* mask);

/* Device pointers of the current FP32 fields for GPU consumers (the
 * reference hands its thickness buffer to caustics and branched flow,
 * pipeline.rs:685-694): h = hi + lo, Γ.  Valid until the next call on ctx.
 * The stream is the context's hipStream_t. */
int sf_drainage_device_fields(sf_drainage* ctx, const float** h_hi, const float** h_lo,
                              const float** gamma);
int sf_drainage_stream(sf_drainage* ctx, void** stream);
/* Average duration in ms of the drainage kernel launches recorded with HIP
 * events on the context's stream since the last call (0 when none). */
int sf_drainage_kernel_timing(sf_drainage* ctx, int32_t enable, double* avg_ms,
                              int64_t* launches);
/* Steps advanced per stencil launch: 1, or 2 (temporal blocking: each band
 * is read and written once per two steps, bit-identical to two one-step
 * launches, marginal-regeneration fronts and the fused FP64 diagnostics of
 * both steps included).  Two steps per launch serve flags without
 * SF_FLAG_MARANGONI and averaged Γ poles, on grids with n_phi % 256 == 0,
 * fewer than 2^30 cells and rebase_interval != 1; other contexts keep one
 * step per launch.  A launch of two steps never straddles a rebase (every
 * rebase_interval steps), so an odd step count or a rebase boundary gives
 * one one-step launch.  The default is 2 on grids of about 2 million cells
 * and more (1024 x 2048 and up), 1 below.  sf_drainage_get_steps_per_launch
 * returns what sf_drainage_step will use; sf_drainage_kernel_timing then
 * averages over all its launches, of one or two steps.
 * sf_drainage_step_spectral renders both states of a two-step launch: the
 * launch also writes the state between its steps.  No counterpart in the
 * reference, whose step dispatches one compute pass per step
 * (gpu_drainage.rs:354-369). */
int sf_drainage_set_steps_per_launch(sf_drainage* ctx, int32_t steps);
int sf_drainage_get_steps_per_launch(sf_drainage* ctx, int32_t* steps);

/* ---- interference --------------------------------------------------------- */

/* generate_interference_lut (interference_lut.rs:18-41): 64 angles x 256
 * thicknesses, angle-major RGBA8, 65536 bytes. */
int sf_lut_rgba8(float refractive_index, float intensity, uint8_t* out);
/* The same LUT before quantisation: [64][256][3] floats. */
int sf_lut_rgb_f32(float refractive_index, float intensity, float* out);

/* InterferenceCalculator::generate_2d_color_lookup_table (interference.rs:491-512)
 * with the calculator's three wavelengths: out[t][a][3], cos from 0 to 1;
 * gamma != 0 applies the 1/2.2 power of calculate_interference_color
 * (interference.rs:380-400), 0 gives calculate_interference_color_linear. */
int sf_lut_2d(double refractive_index, const double wavelengths_nm[3], double tmin_nm,
              double tmax_nm, int32_t n_thickness, int32_t n_angle, int32_t gamma, float* out);
/* InterferenceCalculator::generate_color_lookup_table (interference.rs:461-477):
 * out[n][3] at one cos(theta). */
int sf_lut_1d(double refractive_index, const double wavelengths_nm[3], double tmin_nm,
              double tmax_nm, int32_t n, double cos_theta, float* out);
/* calculate_reflected_intensity (interference.rs:312-353) at n points:
 * out[k] = I(d_nm[k], cos_i[k], lambda_nm[k]). */
int sf_reflected_intensity(double refractive_index, int32_t n, const float* d_nm,
                           const float* cos_i, const float* lambda_nm, float* out);
/* calculate_interference_color / _linear (interference.rs:380-426) at n
 * points: out[k][3], gamma != 0 applies the 1/2.2 power. */
int sf_interference_color(double refractive_index, const double wavelengths_nm[3], int32_t n,
                          const float* d_nm, const float* cos_i, int32_t gamma, float* out);
/* The intermediate terms of interference.rs:197-298 at n points, out[k][6] =
 * {cos_t (197-211), optical path difference (226-238), phase (248-254),
 *  R_s, R_p, R_avg (266-298)}. */
int sf_interference_terms(double refractive_index, int32_t n, const float* d_nm,
                          const float* cos_i, const float* lambda_nm, float* out);
/* calculate_optical_path_difference and calculate_phase_difference
 * (interference.rs:226-254) for given transmission cosines:
 * out[k][2] = {OPD, phase}. */
int sf_optical_path(double refractive_index, int32_t n, const float* d_nm, const float* cos_t,
                    const float* lambda_nm, float* out);
/* calculate_phase_difference (interference.rs:248-254): out[k] = 2 pi OPD / lambda. */
int sf_phase_difference(int32_t n, const float* opd_nm, const float* lambda_nm, float* out);

/* Camera of the spectral evaluation: world position, bubble centred at the
 * origin, y up (theta = 0 is the top). */
typedef struct sf_camera {
    double position[3];
} sf_camera;

/* Per-cell spectral interference: 81 wavelengths (380-780 nm, 5 nm step) of
 * the Airy/Fresnel model of interference_lut.rs:54-72 at each cell's view
 * angle |dot(normalize(cam - p), n)| (bubble.wgsl:491-494), contracted with
 * cmf[81][3] on 16-bit MFMA with FP32 accumulation:
 *   rgb[c][cell] = (1/81) sum_k I_k cmf[k][c].
 * Operand format (sf_spectral_set_operand): by default FP16 when its
 * rounding of the reflectance moves no channel by more than 2.5e-4
 * (2^-18 sum_k |cmf[k][c]|) and the curves fit FP16, else a BF16 hi/lo
 * pair that keeps about 16 bits.  The longest wavelengths (780 nm, then
 * 760-775 nm) are left out while their sum_k |cmf[k][c]| / 81 stays within
 * 1e-5 for every c, which bounds the change of each channel by 1e-5.
 * Reads the context's current h; writes 3 planes of n_theta*n_phi floats
 * to host memory (rgb_host) and/or keeps them on the device. */
int sf_spectral_rgb(sf_drainage* ctx, const sf_camera* cam, double refractive_index,
                    const float* cmf, float* rgb_host);
/* n_steps of {drainage step, spectral RGB of the new state} back to back on
 * the context's stream: the per-step workload of a rendered simulation.
 * dt as in sf_drainage_step; cmf may be NULL after the first call.
 * Asynchronous; the RGB planes of the last step stay on the device
 * (sf_spectral_copy_rgb, sf_spectral_device_rgb). */
int sf_drainage_step_spectral(sf_drainage* ctx, int32_t n_steps, double dt, const sf_camera* cam,
                              double refractive_index, const float* cmf);
/* sf_drainage_step_spectral keeping every step's frame: frame s (s = 0 ..
 * n_steps - 1) is written to frames_dev + 3 * n_theta * n_phi * s (device
 * memory the caller owns, planar R, G, B as sf_spectral_rgb), for consumers
 * of every frame (video export, temporal analysis).  No counterpart in the
 * reference, which renders once per displayed frame (pipeline.rs). */
int sf_drainage_step_spectral_frames(sf_drainage* ctx, int32_t n_steps, double dt, const sf_camera* cam,
                                     double refractive_index, const float* cmf, float* frames_dev);
/* Blocks until all work enqueued on ctx has finished. */
int sf_drainage_synchronize(sf_drainage* ctx);
/* The thickness field rendered through the runtime LUT, one pixel per grid
 * point (image width n_phi, height n_theta): the fragment colour of
 * bubble.wgsl:425-439 -- u = clamp(t_nm / 2000, 0, 1), v = clamp(view cos, 0,
 * 1) with the view cosine of bubble.wgsl:491-494, bilinear sampling of the
 * 256 x 64 RGBA8 LUT with clamp-to-edge addressing (pipeline.rs:350-359),
 * times `intensity` (pipeline.rs:75 default 4), clamped to [0, 1] -- written
 * as RGBA8 with alpha 255 into rgba_out[n_theta*n_phi*4], the frame layout
 * export_frame writes to PNG (image_export.rs:42).  lut_rgba8: the renderer's
 * angle-major 64 x 256 RGBA8 table, 65536 bytes are read; NULL: the LUT of
 * sf_lut_rgba8(refractive_index, 1) as the renderer builds it
 * (pipeline.rs:311-314).  cam NULL: (0, 0, 4R). */
int sf_render_lut_rgba8(sf_drainage* ctx, const uint8_t* lut_rgba8, double refractive_index, float intensity,
                        const sf_camera* cam, uint8_t* rgba_out);
/* Device pointer of the last spectral RGB planes of ctx (3*n_theta*n_phi). */
int sf_spectral_device_rgb(sf_drainage* ctx, const float** rgb);
/* Copies the last spectral RGB planes to rgb_host[3*n_theta*n_phi] after the
 * work enqueued on ctx (blocks). */
int sf_spectral_copy_rgb(sf_drainage* ctx, float* rgb_host);
/* Average duration in ms of the spectral kernel launches since the last call
 * (HIP events on the context's stream); enable switches recording on/off. */
int sf_spectral_kernel_timing(sf_drainage* ctx, int32_t enable, double* avg_ms,
                              int64_t* launches);
/* Operand format of the spectral contraction: AUTO (the default, see
 * sf_spectral_rgb), BF16_PAIR, or FP16 (an error at the next spectral call
 * if some |cmf| exceeds 2^14).  No counterpart in the reference, which has
 * no spectral contraction; for accuracy studies and tests. */
#define SF_SPECTRAL_OPERAND_AUTO 0
#define SF_SPECTRAL_OPERAND_BF16_PAIR 1
#define SF_SPECTRAL_OPERAND_FP16 2
int sf_spectral_set_operand(sf_drainage* ctx, int32_t kind);
/* Operand format the next spectral call uses with the curves last given
 * (SF_SPECTRAL_OPERAND_BF16_PAIR or _FP16; AUTO before any curves). */
int sf_spectral_get_operand(sf_drainage* ctx, int32_t* kind);

/* ---- caustic intensity (consumer of the thickness field) ------------------
 * CausticParams of caustics.rs:20-54 without the grid size (taken from the
 * field).  Only focal_length, caustic_intensity, caustic_sharpness and
 * branch_threshold enter the map (caustics_compute.wgsl:89-114); the others
 * are carried for the ground-plane renderer. */
typedef struct sf_caustic_params {
    float refractive_index, film_thickness_scale;
    float light_dir[3];
    float light_intensity, focal_length, caustic_intensity, caustic_sharpness, branch_threshold;
    float ground_y, bubble_radius, time;
} sf_caustic_params;

/* Defaults of CausticParams::default (caustics.rs:56-78). */
int sf_caustic_params_default(sf_caustic_params* params);
/* Caustic-intensity map of the context's current thickness (metres, FP32
 * h = hi + lo), [n_theta][n_phi] (caustics_compute.wgsl:117-126 dispatched by
 * CausticRenderer::compute, caustics.rs:441-457).  map_host may be NULL: the
 * map then stays on the device (sf_caustics_device_map). */
int sf_caustics(sf_drainage* ctx, const sf_caustic_params* params, float* map_host);
int sf_caustics_device_map(sf_drainage* ctx, const float** map);
/* The same on a host thickness array [n_theta][n_phi]. */
int sf_caustics_field(const sf_caustic_params* params, int32_t n_theta, int32_t n_phi, const float* thickness,
                      float* map);

/* ---- batched foam --------------------------------------------------------
 * Many bubbles, each with its own grid, radius, viscosity and time step,
 * advanced together: one kernel launch per code path in use per step (a
 * 16-step HIP graph is replayed for long runs).  The reference keeps a scalar thickness per foam bubble
 * (foam.rs:32-49); a foam context gives every bubble the drainage field of
 * DrainageSimulator::step (drainage.rs:298-392) with the surfactant
 * transport of drainage.wgsl:178-201: h poles are ring means
 * (drainage.rs:480-503), Γ poles stay fixed, no Marangoni term, no fronts.
 * params->flags may hold SF_FLAG_SURFACTANT only; bubble_radius and
 * viscosity of params are ignored (each bubble brings its own).
 *
 * Layout: the fields of all bubbles are concatenated; bubble b occupies
 * [offsets[b], offsets[b] + n_theta*n_phi) of an array of offsets[n]
 * elements (sf_foam_layout), row-major [theta][phi] as for one bubble. */
typedef struct sf_bubble {
    int32_t n_theta;        /* >= 3                                        */
    int32_t n_phi;          /* even, >= 2                                  */
    double bubble_radius;   /* m                                           */
    double viscosity;       /* Pa s                                        */
    double h0;              /* initial uniform thickness, m                */
    double dt;              /* s per step; <= 0: CFL step (cfl_safety times
                               the FP64 stability limit of the initial
                               state, fixed for the bubble's lifetime)     */
} sf_bubble;

typedef struct sf_foam sf_foam;

/* Creates the foam on the current device with uniform h0 and Γ = 0.5. */
int sf_foam_create(const sf_params* params, int32_t n_bubbles, const sf_bubble* bubbles, sf_foam** out);
void sf_foam_destroy(sf_foam* foam);
/* offsets[n_bubbles + 1]: start of each bubble in the concatenated fields;
 * the last entry is the array length. */
int sf_foam_layout(sf_foam* foam, int64_t* offsets);
/* Replaces h (and Γ, unless NULL: Γ = 0.5) from FP32 arrays of the layout
 * length; CFL time steps are re-resolved. */
int sf_foam_set_fields(sf_foam* foam, const float* h, const float* gamma);
/* As sf_foam_set_fields from FP64 (h kept to ~48 bits); NULL keeps a field. */
int sf_foam_set_fields_f64(sf_foam* foam, const double* h, const double* gamma);
/* Copies h (FP64 value of the float-float pair) and Γ out; NULL skips. */
int sf_foam_copy_fields_f64(sf_foam* foam, double* h, double* gamma);
/* Enqueues n_steps steps of every bubble, each with its own dt. */
int sf_foam_step(sf_foam* foam, int32_t n_steps);
int sf_foam_synchronize(sf_foam* foam);
/* Replays 16-step HIP graphs for long step counts (default on). */
int sf_foam_use_graph(sf_foam* foam, int32_t enable);
/* Steps per stencil launch, 1 or 2 (temporal blocking: a launch reads and
 * writes the foam's state once for two steps, bit-identical to two one-step
 * launches).  Two need every bubble at least 4 x 8 cells; they are the
 * default for foams above 2^23 cells.  A step count that is odd, or that
 * starts at an odd step of the 8-step renormalisation cycle, takes one step
 * per launch where needed.  In foams where two steps are possible a column
 * chunk holds 62 lanes' columns in both modes (lanes 0 and 63 form the
 * two-step kernel's halo columns), so both give identical results.  No counterpart
 * in the reference (foam.rs keeps no per-bubble grids). */
int sf_foam_set_steps_per_launch(sf_foam* foam, int32_t steps);
int sf_foam_get_steps_per_launch(sf_foam* foam, int32_t* steps);
/* Tile shapes: a wave streams a band of rows of one column chunk.
 * index:            0   1   2   3   4   5   6   7   8   9
 * columns per lane: 1   1   2   2   2   4   4   4   1   1
 * rows per band:    16  32  16  32  64  16  32  64  4   2
 * Foams of up to 2^19 cells start with shape 9, up to 2^23 with shape 8,
 * larger ones with 1, 3, 6, 6 per bucket.  Times every shape (0-7 only for
 * foams above 2^23 cells) on each bubble-size bucket (n_phi <= 96, <= 192,
 * <= 384, > 384) for trial_steps steps on a copy of the state, keeps the
 * fastest per bucket and restores the state.  chosen[4] receives the shape
 * per bucket, trial_ms[4][10] the average kernel time per step of each trial (-1
 * for buckets without bubbles or shapes not tried); either may be NULL. */
int sf_foam_autotune(sf_foam* foam, int32_t trial_steps, int32_t* chosen, double* trial_ms);
/* Sets the tile shape index (0-9, as above) of each of the 4 buckets. */
int sf_foam_set_tile_shapes(sf_foam* foam, const int32_t* shapes);
/* dt[n_bubbles]: each bubble's time step. */
int sf_foam_bubble_dt(sf_foam* foam, double* dt);
/* Per-bubble FP64 film volume, CFL limit, min and max of h (arrays of
 * n_bubbles; NULL skips). */
int sf_foam_stats(sf_foam* foam, double* mass, double* dt_cfl, double* hmin, double* hmax);
/* Average step-kernel duration in ms since the last call (HIP events,
 * graphs off while enabled); enable switches recording on/off. */
int sf_foam_kernel_timing(sf_foam* foam, int32_t enable, double* avg_ms, int64_t* launches);
/* Total cells, wave work items per step, and the tile shape per bucket. */
int sf_foam_info(sf_foam* foam, int64_t* total_cells, int32_t* n_items, int32_t* variants);
/* Only bubbles with mask[b] != 0 are stepped from now on (NULL: all); the
 * others keep their state and time until they are stepped again. */
int sf_foam_set_active(sf_foam* foam, const uint8_t* mask);
/* Replaces the state from float-float h (hi + lo) and Γ arrays of the layout
 * length, in host or device memory; Γ NULL keeps it.  Time steps are kept. */
int sf_foam_set_state(sf_foam* foam, const float* hi, const float* lo, const float* gamma);
/* Device pointers of the current state (pole rows written); valid until the
 * next call that steps or changes the foam.  NULL outputs are skipped. */
int sf_foam_device_state(sf_foam* foam, const float** hi, const float** lo, const float** gamma);
/* out[n_bubbles]: the bubble descriptions with their resolved time steps. */
int sf_foam_bubbles(sf_foam* foam, sf_bubble* out);
/* The foam's HIP stream (hipStream_t). */
int sf_foam_stream(sf_foam* foam, void** stream);
/* Number of bubbles of the foam. */
int sf_foam_num_bubbles(sf_foam* foam, int32_t* n);

/* ---- foam cluster dynamics ------------------------------------------------
 * FoamSimulator (foam_dynamics.rs:9-202) over a BubbleCluster (foam.rs:208-
 * 399): all-pairs van der Waals / Hertzian repulsion / contact adhesion
 * forces, buoyancy and Stokes drag, semi-implicit Euler with damping and a
 * 0.1 m/s speed clamp, contact connections and coalescence.  FP32 state as
 * the reference (glam Vec3 / f32); pair forces in FP32, their sums in FP64.
 * Bubbles keep the reference's ids and vector order (removal keeps order,
 * additions and merges append with the next id). */
typedef struct sf_cluster_params {
    float surface_tension;          /* N/m                         0.025   */
    float viscosity;                /* Pa s                        0.001   */
    float density;                  /* kg/m^3                      1000    */
    float gravity[3];               /* m/s^2                   (0,-9.81,0) */
    float van_der_waals_strength;   /*                             0.001   */
    float repulsion_stiffness;      /*                             5.0     */
    float damping;                  /* velocity factor per step    0.8     */
    int32_t coalescence_enabled;    /*                             0       */
    float coalescence_threshold;    /* Weber number                0.5     */
} sf_cluster_params;

/* Bubble (foam.rs:31-49) without its neighbour list. */
typedef struct sf_cluster_bubble {
    uint32_t id;
    float position[3];      /* m   */
    float radius;           /* m   */
    float velocity[3];      /* m/s */
    float thickness_nm;
    float aspect_ratio;
    float refractive_index;
} sf_cluster_bubble;

/* BubbleConnection (foam.rs:89-139). */
typedef struct sf_cluster_connection {
    uint32_t bubble_a, bubble_b;
    float contact_point[3];
    float contact_normal[3];        /* from a toward b                     */
    float wall_curvature_radius;    /* +inf for equal radii               */
    float wall_thickness_nm;
} sf_cluster_connection;

typedef struct sf_cluster sf_cluster;

/* FoamSimulator::new defaults (foam_dynamics.rs:37-53). */
int sf_cluster_params_default(sf_cluster_params* p);
/* BubbleCluster from bubbles in vector order; the next id is one past the
 * largest given.  Replaces FoamSimulator::new (foam_dynamics.rs:37). */
int sf_cluster_create(const sf_cluster_params* params, int32_t n, const sf_cluster_bubble* bubbles,
                      sf_cluster** out);
void sf_cluster_destroy(sf_cluster* c);
int sf_cluster_set_params(sf_cluster* c, const sf_cluster_params* params);
/* n_steps of FoamSimulator::step(dt) (foam_dynamics.rs:166-202), enqueued
 * on the cluster's stream (with coalescence on, each step waits for its
 * merge test). */
int sf_cluster_step(sf_cluster* c, float dt, int32_t n_steps);
/* Blocks until the cluster's work has finished. */
int sf_cluster_synchronize(sf_cluster* c);
/* Average duration in ms of the all-pairs force kernel since the last call
 * (HIP events on the cluster's stream); enable switches recording on/off. */
int sf_cluster_kernel_timing(sf_cluster* c, int32_t enable, double* avg_ms, int64_t* launches);
/* compute_forces (foam_dynamics.rs:61-92) of the current state:
 * forces[n][3]. */
int sf_cluster_compute_forces(sf_cluster* c, float* forces);
int sf_cluster_num_bubbles(sf_cluster* c, int32_t* n);
/* Bubbles in vector order (BubbleCluster::bubbles, foam.rs:273-275). */
int sf_cluster_get_bubbles(sf_cluster* c, sf_cluster_bubble* out);
/* Replaces every bubble (state written by a caller, e.g. from a saved run). */
int sf_cluster_set_bubbles(sf_cluster* c, int32_t n, const sf_cluster_bubble* bubbles);
/* Connections of the last update in (i, j) vector order
 * (BubbleCluster::connections, foam.rs:305-307; connection_count,
 * foam_dynamics.rs:236-238). */
int sf_cluster_num_connections(sf_cluster* c, int32_t* n);
int sf_cluster_get_connections(sf_cluster* c, sf_cluster_connection* out);
/* Neighbour id lists (Bubble::neighbors): offsets[n + 1] and ids[offsets[n]]
 * (ids NULL: offsets only). */
int sf_cluster_neighbors(sf_cluster* c, int32_t* offsets, uint32_t* ids);
/* update_connections (foam.rs:321-360) on the current state. */
int sf_cluster_update_connections(sf_cluster* c);
/* add_bubble (foam.rs:235-244), remove_bubble (249-259), merge_bubbles
 * (365-399). */
int sf_cluster_add_bubble(sf_cluster* c, const float position[3], float radius, uint32_t* id);
int sf_cluster_remove_bubble(sf_cluster* c, uint32_t id, int32_t* removed);
int sf_cluster_merge(sf_cluster* c, uint32_t id_a, uint32_t id_b, uint32_t* new_id);
/* Feeds the cluster into batched drainage: one sf_bubble per cluster bubble
 * (vector order) with its radius, film thickness (h0), the cluster's
 * viscosity, CFL dt, and n_theta mapped log-linearly from radius_min..max to
 * n_theta_min..max (n_phi = 2 n_theta).  The reference keeps only a scalar
 * thickness per foam bubble (foam.rs:44). */
int sf_cluster_foam_bubbles(sf_cluster* c, int32_t n_theta_min, int32_t n_theta_max, float radius_min,
                            float radius_max, sf_bubble* out);
/* Sets each bubble's thickness_nm to its foam bubble's mean film thickness
 * (FP64 film volume / 4 pi R^2); the foam must have as many bubbles. */
int sf_cluster_thickness_from_foam(sf_cluster* c, sf_foam* foam);

/* ---- learned surrogate (Fourier neural operator) --------------------------
 * No reference counterpart: a network that advances a bubble's film
 * thickness by `window` drainage steps, trained on trajectories of the plain
 * solver (a foam context) with this library's own HIP forward and backward
 * kernels (BF16 MFMA products, FP32 master weights, Adam), and checked every
 * window by one plain step that acts as judge:
 *   the network predicts h(t+K-1) and h(t+K); h_J = one plain step from
 *   h(t+K-1); a bubble's window is accepted when
 *   ||h_J - h(t+K)||_2 <= tol ||h(t+K)||_2 and its state becomes h_J,
 *   otherwise it falls back to K plain steps from h(t).
 * Input channels per cell: relative thickness perturbation, surfactant,
 * sin(theta), the window's gravity-sink number K dt rho g h_ref^2 / (3 eta),
 * its diffusion number K dt D / (R^2 dtheta^2) (h_ref = the bubble's mean h),
 * the previous window's increment d = h(t) - h(t-K) (zero on the first
 * window of a train or advance call), dt D lap(d), its change under one
 * explicit diffusion step, and one explicit plain step of h,
 * dt (D lap(h) - c h^3 sin theta) (zero on the pole rows); the last three are
 * divided by the bubble's increment scale y = rms(d), or h_ref s_y when
 * there is no previous window.  Outputs: corrections to linear extrapolation
 * of that increment, (h(t+j) - h(t) - (j/K) d) / y for j = K-1, K.
 * Tensors are channel-last [batch][n_theta][n_phi][channels], FP32. */
typedef struct sf_fno_config {
    int32_t n_theta, n_phi;   /* grid of every bubble; n_theta*n_phi % 4096 == 0 */
    int32_t width;            /* channels (divides 256, <= 64)                  */
    int32_t modes_theta;      /* theta modes k in [-modes_theta, modes_theta)   */
    int32_t modes_phi;        /* phi modes m in [0, modes_phi)                  */
    int32_t layers;           /* Fourier layers                                 */
    int32_t window;           /* K, steps per inference                         */
    int32_t max_batch;        /* bubbles per forward/backward pass              */
    uint64_t seed;            /* weight initialisation                          */
} sf_fno_config;

typedef struct sf_fno_stats {
    int64_t windows, bubble_windows, accepted, fallbacks;
    double max_residual;            /* max over bubble-windows of the judge residual   */
    double sum_residual;
    double max_increment_residual;  /* same error relative to ||h(t+K) - h(t)||        */
    double surrogate_ms, judge_ms, fallback_ms;   /* host wall clock per phase (sums) */
} sf_fno_stats;

typedef struct sf_fno sf_fno;

/* 256 x 512 grid, width 32, 16 x 16 modes, 4 layers, K = 16, batch 16. */
int sf_fno_config_default(sf_fno_config* config);
int sf_fno_create(const sf_fno_config* config, sf_fno** out);
void sf_fno_destroy(sf_fno* model);
int sf_fno_num_params(sf_fno* model, int64_t* n);
/* Flat FP32 parameters: lift W[8][width], b[width]; per layer W[width][width],
 * b[width], R[2 modes_theta][modes_phi][2][width][width] (re, im); projection
 * Q[width][2], q[2]. */
int sf_fno_get_params(sf_fno* model, float* params);
int sf_fno_set_params(sf_fno* model, const float* params);
/* scales[2] = {s_x, s_y}: input perturbation and target scales (set from the
 * first training window when unset). */
int sf_fno_get_scales(sf_fno* model, double* scales);
int sf_fno_set_scales(sf_fno* model, const double* scales);
/* Host-array forward pass: in[batch][N][8] -> out[batch][N][2]. */
int sf_fno_forward(sf_fno* model, int32_t batch, const float* in, float* out);
/* Kernel paths of this model's shape: the fused per-row layer kernel (training
 * and inference), the lift folded into the first phi DFT, the fused theta
 * stage and the one-wave-per-row layer kernel (inference only); shapes
 * outside them run the general batched GEMMs. */
#define SF_FNO_PATH_FUSED_ROWS   (1 << 0)
#define SF_FNO_PATH_FUSED_THETA  (1 << 1)
#define SF_FNO_PATH_LIFT_DFT     (1 << 2)
#define SF_FNO_PATH_WAVE_ROWS    (1 << 3)   /* inference layers one wave per grid row */
int sf_fno_kernel_paths(sf_fno* model, int32_t* paths);
/* Loss mean((out - target)^2) and its gradient w.r.t. every parameter. */
int sf_fno_gradient(sf_fno* model, int32_t batch, const float* in, const float* target, double* loss,
                    float* grad);
/* One Adam step (lr, beta 0.9 / 0.999) on a host batch; returns the loss. */
int sf_fno_train_step(sf_fno* model, int32_t batch, const float* in, const float* target, double lr,
                      double* loss);
/* `windows` consecutive training windows generated online by the foam's
 * plain solver (every bubble n_theta x n_phi); losses[windows] (may be NULL).
 * The previous-increment input is carried between the windows of one call
 * only: a call's first window sees zero. */
int sf_fno_train_foam(sf_fno* model, sf_foam* foam, const sf_params* params, int32_t windows, double lr,
                      double* losses);
/* Predictors of sf_fno_advance_foam: the network, or the no-network
 * baselines judged the same way: h(t) itself, or the previous window's
 * increment carried on linearly (identity on a call's first window). */
#define SF_FNO_PREDICT_NETWORK      0
#define SF_FNO_PREDICT_IDENTITY     1
#define SF_FNO_PREDICT_EXTRAPOLATE  2
/* `windows` surrogate windows with judge and fallback on every bubble;
 * stats (may be NULL) are accumulated. */
int sf_fno_advance_foam(sf_fno* model, sf_foam* foam, const sf_params* params, int32_t windows, double tol,
                        int32_t predictor, sf_fno_stats* stats);
/* Input features of bubbles [first, first + count) of the foam's current
 * state into in_host[count][N][8] (previous-increment channels zero). */
int sf_fno_features(sf_fno* model, sf_foam* foam, const sf_params* params, int32_t first, int32_t count,
                    float* in_host);

#ifdef __cplusplus
}
#endif

#endif /* SOAPFILM_H */
