// drainage.hpp -- the sf_drainage context (one bubble on one GPU).
#pragma once

#include <utility>
#include <vector>

#include "drainage_kernels.hpp"

namespace sf {

struct StatsPart;

// Host copy of one marginal-regeneration front (TfeFront, drainage.rs:23-33)
// plus its precomputed column run.
struct HostFront {
    double phi_center, extent, width, velocity;
    int j0, len;
};

// HIP-event timing of one kernel's launches on a stream: an event pair
// brackets every launch, and drain() sums the elapsed times.
struct KernelTimer {
    static constexpr size_t kPool = 4096;
    bool enabled = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pool;
    size_t used = 0;
    double ms = 0.0;
    int64_t launches = 0;

    std::pair<hipEvent_t, hipEvent_t>& next(hipStream_t s);
    void drain(hipStream_t s);
    void destroy();
};

struct Drainage;

// A film state on the device: h = hi + lo, with the pole rows held as ring
// partial sums when pole_virtual is set.
struct FieldView {
    const float* hi;
    const float* lo;
    const float2* hpole;
    bool pole_virtual;
};

// Spectral RGB of the context's current state, or of `view` when given,
// into the context's RGB planes, or into `rgb_out` (device) when given.
void spectral_launch(Drainage* d, const sf_camera* cam, double n_film, const float* cmf, hipStream_t st,
                     const FieldView* view = nullptr, float* rgb_out = nullptr);

struct Drainage {
    static constexpr int64_t kHistCap = 1 << 16;
    static constexpr size_t kStatsBlocks = 1024;

    sf_params p;
    int nt, np;
    size_t n;
    double dtheta, dphi;
    int device = 0;
    hipStream_t stream = nullptr;

    // launch geometry
    bool vec4 = false;   // float4 kernel (n_phi % 4 == 0)
    int wx = 1, block = 64, ntiles = 1, rows_per_block = 16;
    // Two steps per launch (drainage_band2_kernel) where two_step_ok() holds:
    // steps_per_launch 2 asks for it, 1 keeps one step per launch.
    int steps_per_launch = 1;
    int rows_per_block2 = 16;     // band height of the two-step kernel

    // device state (see drainage_kernels.hpp for the layout)
    float* hi[2] = {nullptr, nullptr};
    float* lo[2] = {nullptr, nullptr};
    float* g[2] = {nullptr, nullptr};
    int cur_hi = 0, cur = 0, cur_g = 0;
    int since_rebase = 0;
    float2* hpole[2] = {nullptr, nullptr};
    float2* gpole[2] = {nullptr, nullptr};
    int cur_pole = 0;
    // The state between the two steps of a two-step launch, for consumers
    // of every step: its low word (allocated on first use) and ring partial
    // sums (ping-pong, so a pending diagnosis of one launch's first step
    // survives the next launch).
    float* lo_mid = nullptr;
    float2* hpole_mid[2] = {nullptr, nullptr};
    int cur_mid = 0;
    bool poles_virtual = false;   // pole rows are held as ring partials
    RowF* rows = nullptr;
    RowD* rowd = nullptr;
    bool tables_dirty = true;
    double pole_area_top = 0.0, pole_area_bot = 0.0, pole_mass_fixed = 0.0;

    // FP64 diagnostics: per-block partials of each diagnosed step, reduced
    // into the history by the next launch (or finalize_diag); the steps of
    // the last launch (one or two) are pending until then.
    static constexpr int kDiagBufs = 4;
    struct PendingDiag {
        int buf;                  // diag_part index
        int slot;                 // history slot
        int nblocks;
        bool pole_sum;
        const float2* hpole;      // the step's ring partial sums
    };
    size_t nblocks_max = 0;
    double2* diag_part[kDiagBufs] = {nullptr, nullptr, nullptr, nullptr};
    double2* diag_hist = nullptr;
    int64_t hist_count = 0;
    PendingDiag pending[2] = {{0, -1, 0, false, nullptr}, {0, -1, 0, false, nullptr}};
    int npending = 0;
    int new_buf[2] = {0, 1};      // diag_part buffers of the launch being set up
    StatsPart* stats_part = nullptr;
    StatsPart* stats_out = nullptr;

    // queries
    int64_t* gather_idx = nullptr;
    float2* gather_out = nullptr;
    size_t gather_cap = 0;

    // spectral interference
    float* rgb = nullptr;
    float* caustic = nullptr;       // caustic-intensity map (caustics.hip), allocated on first use
    float* cmf_dev = nullptr;       // colour curves as MFMA B fragments
    // Per operand format (0: BF16 hi/lo pair, 1: FP16): sum over the 81
    // wavelengths of each curve as the MFMA sees it, and whether the curves
    // need their lo fragments.
    float cmf_sum[2][3] = {{0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}};
    int cmf_lo[2] = {0, 0};
    bool cmf_f16_bound = false;     // FP16 r keeps the RGB error within kSpecF16Budget
    bool cmf_f16_range = false;     // the curves fit FP16
    int cmf_skip = 0;               // trailing wavelength slots the spectral pass leaves out
    int spec_operand = 0;           // SF_SPECTRAL_OPERAND_*
    float4* spec_rows = nullptr;
    float2* spec_view = nullptr;    // per-cell (F, cos theta_t) of spec_view_key's camera and index
    float spec_view_key[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    bool spec_view_valid = false;
    float2* spec_cols = nullptr;

    // time and TFE schedule (drainage.rs:237-255)
    double current_time = 0.0;
    int64_t steps_taken = 0;
    bool tfe_enabled = false;
    double next_spawn = 0.5;
    std::vector<HostFront> fronts;
    std::vector<FrontBox> last_boxes;

    // kernel timing with HIP events on `stream`
    KernelTimer step_timer, spec_timer;

    Drainage(const sf_params& prm, int n_theta, int n_phi);
    ~Drainage();
    Drainage(const Drainage&) = delete;
    Drainage& operator=(const Drainage&) = delete;

    void upload_tables();
    void update_fixed_pole_mass();
    void upload(const float* h_hi, const float* h_lo, const float* gam);
    void fill(double h0, double gamma0);
    void front_columns(HostFront& f) const;
    void spawn_front();
    void schedule_fronts(double dt, StepArgs& a);
    void fill_common_args(StepArgs& a, float dt_f) const;
    void materialize_poles();
    void finalize_diag();
    void materialize();
    double resolve_dt(double dt);
    void step(int n_steps, double dt, bool flush = true);
    // Whether drainage_band2_kernel serves these flags and this grid.
    bool two_step_ok() const;
    // Whether the next two of `remaining` steps (remaining >= 2) go into one
    // launch: two steps per launch are on and no rebase falls between them.
    bool two_step_next(int remaining) const;
    // One launch of one step, or of two steps (with_mid: also writing the
    // state between them).  dt is resolved; tables are uploaded.
    void launch_one(double dt);
    void launch_two(double dt, bool with_mid);
    // The state between the steps of the last two-step launch made with
    // with_mid: high word, low word, ring partials.
    const float* mid_hi = nullptr;
    const float2* mid_hpole = nullptr;
    // Pending-diagnosis bookkeeping of a launch's steps.
    void diag_args(StepArgs& a, int nsteps);
    void diag_commit(int nsteps, int nblocks, bool pole_sum, const float2* hpole_first, const float2* hpole_last);
    StatsPart compute_stats();
    void ensure_gather(size_t count);
};

}  // namespace sf
