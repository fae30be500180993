// soapfilm.hpp -- C++ host API over the C ABI of soapfilm.h.
//
// Mirrors the reference's Rust API for the drainage and interference path
// so that C++ callers read like the reference's own code:
//   soapfilm::SimulationConfig      config.rs:283-370 (fields used by the path)
//   soapfilm::DrainageSimulator     drainage.rs:212-614
//   soapfilm::GPUDrainageSimulator  gpu_drainage.rs:81-491
//   soapfilm::InterferenceCalculator interference.rs:127-513
//   soapfilm::generate_interference_lut interference_lut.rs:18-41
// Errors from the library become soapfilm::Error exceptions.
#pragma once

#include <array>
#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "soapfilm.h"

namespace soapfilm {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline void check(int rc) {
    if (rc != SF_OK) throw Error(rc, std::string("soapfilm: ") + sf_last_error());
}

// The subset of SimulationConfig (config.rs:283-316) that the drainage path
// reads, with the same defaults.
struct SimulationConfig {
    double diameter = 0.05;
    double film_thickness_nm = 500.0;
    double critical_thickness_nm = 10.0;
    double refractive_index = 1.33;
    double viscosity = 0.001;
    double surface_tension = 0.025;
    double density = 1000.0;
    double surfactant_diffusion = 1e-9;
    double surfactant_concentration = 0.5;
    double gravity = 9.81;
    double dt = 0.001;
    uint32_t resolution = 128;

    double bubble_radius() const { return diameter / 2.0; }
    double film_thickness_meters() const { return film_thickness_nm * 1e-9; }
    double internal_pressure_difference() const { return 4.0 * surface_tension / bubble_radius(); }

    sf_params to_params() const {
        sf_params p;
        check(sf_params_cpu_default(&p));
        p.density = density;
        p.gravity = gravity;
        p.viscosity = viscosity;
        p.diffusion_coeff = surfactant_diffusion;
        p.critical_thickness = critical_thickness_nm * 1e-9;
        p.bubble_radius = bubble_radius();
        p.dt = dt;
        return p;
    }
};

// Owning handle of one sf_drainage context.
class Context {
public:
    Context(const sf_params& p, int n_theta, int n_phi, double h0, double gamma0)
        : nt_(n_theta), np_(n_phi) {
        check(sf_drainage_create(&p, n_theta, n_phi, h0, gamma0, &h_));
    }
    ~Context() { sf_drainage_destroy(h_); }
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;
    Context(Context&& o) noexcept : h_(o.h_), nt_(o.nt_), np_(o.np_) { o.h_ = nullptr; }

    sf_drainage* get() const { return h_; }
    int num_theta() const { return nt_; }
    int num_phi() const { return np_; }

    void step(int n, double dt) { check(sf_drainage_step(h_, n, dt)); }
    double get_thickness(double theta, double phi) const {
        double v;
        check(sf_drainage_get_thickness(h_, theta, phi, &v));
        return v;
    }
    std::pair<std::vector<double>, std::vector<double>> fields() const {
        std::vector<double> h(static_cast<size_t>(nt_) * np_), g(h.size());
        check(sf_drainage_copy_field_f64(h_, h.data(), g.data()));
        return {std::move(h), std::move(g)};
    }
    void set_fields(const std::vector<double>& h, const std::vector<double>& g) {
        check(sf_drainage_set_fields_f64(h_, h.empty() ? nullptr : h.data(),
                                         g.empty() ? nullptr : g.data()));
    }
    sf_stats stats() const {
        sf_stats s;
        check(sf_drainage_stats(h_, &s));
        return s;
    }
    sf_params params() const {
        sf_params p;
        check(sf_drainage_get_params(h_, &p));
        return p;
    }
    void set_params(const sf_params& p) { check(sf_drainage_set_params(h_, &p)); }

private:
    sf_drainage* h_ = nullptr;
    int nt_, np_;
};

// ThicknessField (drainage.rs:96-204): a host snapshot of the device field.
class ThicknessField {
public:
    ThicknessField(int nt, int np, std::vector<double> data)
        : nt_(nt), np_(np), data_(std::move(data)),
          dtheta_(M_PI / static_cast<double>(nt - 1)), dphi_(2.0 * M_PI / static_cast<double>(np)) {}
    double get(size_t i, size_t j) const { return data_[i * static_cast<size_t>(np_) + j]; }
    double theta_at(size_t i) const { return static_cast<double>(i) * dtheta_; }
    double phi_at(size_t j) const { return static_cast<double>(j) * dphi_; }
    int num_theta() const { return nt_; }
    int num_phi() const { return np_; }
    double delta_theta() const { return dtheta_; }
    double delta_phi() const { return dphi_; }
    double min_thickness() const {
        double m = INFINITY;
        for (double v : data_) m = std::fmin(m, v);
        return m;
    }
    double max_thickness() const {
        double m = -INFINITY;
        for (double v : data_) m = std::fmax(m, v);
        return m;
    }
    const std::vector<double>& data() const { return data_; }

private:
    int nt_, np_;
    std::vector<double> data_;
    double dtheta_, dphi_;
};

// DrainageSimulator (drainage.rs:212-614) on the GPU.
class DrainageSimulator {
public:
    explicit DrainageSimulator(const SimulationConfig& config = SimulationConfig())
        : gamma0_(config.surfactant_concentration),
          ctx_(config.to_params(), static_cast<int>(config.resolution),
               static_cast<int>(config.resolution) * 2, config.film_thickness_meters(),
               config.surfactant_concentration) {}

    void step(double dt) { ctx_.step(1, dt); }
    double get_thickness(double theta, double phi) const { return ctx_.get_thickness(theta, phi); }
    ThicknessField thickness_field() const {
        return ThicknessField(ctx_.num_theta(), ctx_.num_phi(), ctx_.fields().first);
    }
    double current_time() const {
        double t;
        check(sf_drainage_current_time(ctx_.get(), &t));
        return t;
    }
    bool has_critical_region() const { return ctx_.stats().has_critical_region != 0; }
    double critical_thickness() const { return ctx_.params().critical_thickness; }
    void reset(double initial_thickness) { check(sf_drainage_reset(ctx_.get(), initial_thickness, gamma0_)); }
    void set_marginal_regeneration(bool on) { check(sf_drainage_set_marginal_regeneration(ctx_.get(), on)); }
    bool marginal_regeneration_enabled() const { return (ctx_.params().flags & SF_FLAG_TFE) != 0; }
    void set_tfe_thickness_ratio(double r) { check(sf_drainage_set_tfe_thickness_ratio(ctx_.get(), r)); }
    double tfe_thickness_ratio() const { return ctx_.params().tfe_thickness_ratio; }
    size_t active_tfe_count() const {
        int32_t n;
        check(sf_drainage_active_tfe_count(ctx_.get(), &n));
        return static_cast<size_t>(n);
    }
    Context& context() { return ctx_; }

private:
    double gamma0_;
    Context ctx_;
};

// DrainageParams (gpu_drainage.rs:12-68).
struct DrainageParams {
    float dt = 0.001f, gravity = 9.81f, viscosity = 0.001f, density = 1000.0f;
    float diffusion_coeff = 1e-9f, bubble_radius = 0.025f, critical_thickness = 30e-9f;
    uint32_t grid_width = 128, grid_height = 64, marangoni_enabled = 0;
    float gamma_air = 0.072f, gamma_reduction = 0.045f, surfactant_diffusion = 1e-9f;
    float marangoni_coeff = 0.01f;

    sf_params to_sf() const {
        sf_params p;
        check(sf_params_gpu_default(&p));
        p.dt = dt;
        p.gravity = gravity;
        p.viscosity = viscosity;
        p.density = density;
        p.diffusion_coeff = diffusion_coeff;
        p.bubble_radius = bubble_radius;
        p.critical_thickness = critical_thickness;
        p.gamma_air = gamma_air;
        p.gamma_reduction = gamma_reduction;
        p.surfactant_diffusion = surfactant_diffusion;
        p.marangoni_coeff = marangoni_coeff;
        if (marangoni_enabled) p.flags |= SF_FLAG_MARANGONI;
        return p;
    }
};

// GPUDrainageSimulator (gpu_drainage.rs:81-491): the kernel steps with
// params.dt while the frame dt only advances current_time.
class GPUDrainageSimulator {
public:
    GPUDrainageSimulator(float initial_thickness, uint32_t grid_width, uint32_t grid_height)
        : params_(make(grid_width, grid_height)),
          ctx_(params_.to_sf(), static_cast<int>(grid_height), static_cast<int>(grid_width),
               initial_thickness, 0.5) {}

    bool enabled = false;
    float time_scale = 100.0f;
    uint32_t steps_per_frame = 10;
    bool marangoni_enabled = false;

    void step(float dt) {
        if (!enabled) return;
        const float step_dt = dt * time_scale / static_cast<float>(steps_per_frame);
        ctx_.step(static_cast<int>(steps_per_frame), params_.dt);
        for (uint32_t k = 0; k < steps_per_frame; ++k) current_time_ += static_cast<double>(step_dt);
    }
    void update_params(const DrainageParams& p) {
        params_ = p;
        ctx_.set_params(p.to_sf());
    }
    void set_dt(float dt) {
        params_.dt = dt;
        ctx_.set_params(params_.to_sf());
    }
    void set_marangoni_enabled(bool on) {
        marangoni_enabled = on;
        params_.marangoni_enabled = on ? 1 : 0;
        ctx_.set_params(params_.to_sf());
    }
    void reset(float initial_thickness) {
        check(sf_drainage_reset(ctx_.get(), initial_thickness, 0.5));
        current_time_ = 0.0;
    }
    double current_time() const { return current_time_; }
    std::pair<uint32_t, uint32_t> grid_size() const { return {params_.grid_width, params_.grid_height}; }
    const DrainageParams& params() const { return params_; }
    // Device buffers for GPU consumers (current_thickness_buffer, 471-480).
    void current_thickness_buffer(const float** hi, const float** lo) {
        check(sf_drainage_device_fields(ctx_.get(), hi, lo, nullptr));
    }
    Context& context() { return ctx_; }

private:
    static DrainageParams make(uint32_t w, uint32_t h) {
        DrainageParams p;
        p.grid_width = w;
        p.grid_height = h;
        return p;
    }
    DrainageParams params_;
    Context ctx_;
    double current_time_ = 0.0;
};

struct RgbColor {
    double red = 0.0, green = 0.0, blue = 0.0;
    static RgbColor make(double r, double g, double b) {
        auto c = [](double v) { return v < 0.0 ? 0.0 : (v > 1.0 ? 1.0 : v); };
        return RgbColor{c(r), c(g), c(b)};
    }
};

// InterferenceCalculator (interference.rs:127-513); evaluations on the GPU.
class InterferenceCalculator {
public:
    explicit InterferenceCalculator(double n = 1.33, std::array<double, 3> w = {650.0, 532.0, 450.0})
        : n_(n), w_(w) {}
    static InterferenceCalculator with_custom_wavelengths(double n, double r, double g, double b) {
        return InterferenceCalculator(n, {r, g, b});
    }
    double refractive_index_film() const { return n_; }
    double refractive_index_medium() const { return 1.0; }

    double calculate_reflected_intensity(double d_nm, double cos_i, double lambda_nm) const {
        const float d = static_cast<float>(d_nm), c = static_cast<float>(cos_i),
                    l = static_cast<float>(lambda_nm);
        float out;
        check(sf_reflected_intensity(n_, 1, &d, &c, &l, &out));
        return out;
    }
    RgbColor calculate_interference_color(double d_nm, double cos_i, bool gamma = true) const {
        const float d = static_cast<float>(d_nm), c = static_cast<float>(cos_i);
        float out[3];
        check(sf_interference_color(n_, w_.data(), 1, &d, &c, gamma ? 1 : 0, out));
        return RgbColor::make(out[0], out[1], out[2]);
    }
    RgbColor calculate_interference_color_linear(double d_nm, double cos_i) const {
        return calculate_interference_color(d_nm, cos_i, false);
    }
    std::vector<RgbColor> generate_color_lookup_table(double tmin, double tmax, size_t n,
                                                      double cos_theta) const {
        std::vector<float> raw(3 * n);
        check(sf_lut_1d(n_, w_.data(), tmin, tmax, static_cast<int32_t>(n), cos_theta, raw.data()));
        std::vector<RgbColor> out(n);
        for (size_t k = 0; k < n; ++k) out[k] = RgbColor::make(raw[3 * k], raw[3 * k + 1], raw[3 * k + 2]);
        return out;
    }
    std::vector<std::vector<RgbColor>> generate_2d_color_lookup_table(double tmin, double tmax,
                                                                      size_t nt, size_t na) const {
        std::vector<float> raw(3 * nt * na);
        check(sf_lut_2d(n_, w_.data(), tmin, tmax, static_cast<int32_t>(nt), static_cast<int32_t>(na),
                        1, raw.data()));
        std::vector<std::vector<RgbColor>> out(nt, std::vector<RgbColor>(na));
        for (size_t t = 0; t < nt; ++t)
            for (size_t a = 0; a < na; ++a) {
                const float* px = raw.data() + 3 * (t * na + a);
                out[t][a] = RgbColor::make(px[0], px[1], px[2]);
            }
        return out;
    }

private:
    double n_;
    std::array<double, 3> w_;
};

constexpr uint32_t LUT_THICKNESS_SAMPLES = 256;   // interference_lut.rs:9
constexpr uint32_t LUT_ANGLE_SAMPLES = 64;        // interference_lut.rs:10
constexpr float LUT_MAX_THICKNESS_NM = 2000.0f;   // interference_lut.rs:13

inline std::vector<uint8_t> generate_interference_lut(float refractive_index, float intensity) {
    std::vector<uint8_t> out(static_cast<size_t>(LUT_THICKNESS_SAMPLES) * LUT_ANGLE_SAMPLES * 4);
    check(sf_lut_rgba8(refractive_index, intensity, out.data()));
    return out;
}

}  // namespace soapfilm
