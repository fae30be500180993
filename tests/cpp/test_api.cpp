// C++ mirror of the reference's unit tests for the drainage and
// interference path (drainage.rs:616-720, interference.rs:612-792,
// interference_lut.rs:165-205), run against the GPU through soapfilm.hpp.
// Prints one line per test and exits non-zero on the first failure.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <limits>

#include "soapfilm.hpp"

using namespace soapfilm;

static int failures = 0;
#define EXPECT(cond)                                                                 \
    do {                                                                             \
        if (!(cond)) {                                                               \
            std::printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #cond);             \
            ++failures;                                                              \
        }                                                                            \
    } while (0)

static const double kEps = std::numeric_limits<double>::epsilon();

static void test_drainage() {
    SimulationConfig config;
    {
        DrainageSimulator sim(config);                                   // 636-642
        EXPECT(std::fabs(sim.current_time()) < kEps);
        EXPECT(!sim.has_critical_region());
        sim.step(0.001);                                                 // 645-653
        EXPECT(std::fabs(sim.current_time() - 0.001) < kEps);
    }
    {
        DrainageSimulator sim(config);                                   // 656-665
        EXPECT(std::fabs(sim.get_thickness(0.5, 1.0) - sim.get_thickness(1.5, 3.0)) < 1e-15);
        const double top = sim.get_thickness(0.1, 0.0);                  // 668-688
        for (int k = 0; k < 100; ++k) sim.step(config.dt);
        EXPECT(sim.get_thickness(0.1, 0.0) < top);
    }
    {
        DrainageSimulator sim(config);                                   // 691-702
        EXPECT(std::fabs(sim.get_thickness(1.0, 0.0) - sim.get_thickness(1.0, 2.0 * M_PI)) < 1e-12);
        for (int k = 0; k < 10; ++k) sim.step(config.dt);                // 705-719
        sim.reset(600e-9);
        EXPECT(std::fabs(sim.current_time()) < kEps);
        EXPECT(std::fabs(sim.get_thickness(0.5, 0.5) - 600e-9) < 1e-15);
        ThicknessField f = sim.thickness_field();
        EXPECT(f.num_theta() == 128 && f.num_phi() == 256);
        EXPECT(std::fabs(f.get(10, 20) - 600e-9) < kEps);
    }
    {
        GPUDrainageSimulator g(500e-9f, 128, 64);
        g.enabled = true;
        g.step(1.0f / 60.0f);
        EXPECT(g.current_time() > 0.0);
        EXPECT(g.grid_size().first == 128 && g.grid_size().second == 64);
    }
    std::printf("drainage: done\n");
}

static void test_interference() {
    InterferenceCalculator c(1.33);
    EXPECT(std::fabs(c.refractive_index_film() - 1.33) < 1e-6);
    const double r0 = std::pow((1.0 - 1.33) / 2.33, 2);
    EXPECT(std::fabs(c.calculate_reflected_intensity(0.0, 1.0, 550.0)) < 1e-6);        // 682-686
    for (double t : {100.0, 200.0, 300.0, 400.0, 500.0}) {                              // 689-700
        RgbColor col = c.calculate_interference_color(t, 1.0);
        EXPECT(std::fmax(col.red, std::fmax(col.green, col.blue)) > 0.01);
    }
    auto table = c.generate_color_lookup_table(0.0, 1000.0, 100, 1.0);                  // 734-745
    EXPECT(table.size() == 100);
    auto t2 = c.generate_2d_color_lookup_table(0.0, 500.0, 10, 5);                      // 748-755
    EXPECT(t2.size() == 10 && t2[0].size() == 5);
    (void)r0;
    auto lut = generate_interference_lut(1.33f, 1.0f);                                  // lut 170-174
    EXPECT(lut.size() == LUT_THICKNESS_SAMPLES * LUT_ANGLE_SAMPLES * 4);
    std::printf("interference: done\n");
}

int main() {
    if (!sf_device_available()) {
        std::printf("no GPU\n");
        return 2;
    }
    try {
        test_drainage();
        test_interference();
    } catch (const Error& e) {
        std::printf("FAIL exception: %s\n", e.what());
        return 1;
    }
    std::printf(failures ? "FAILED %d\n" : "OK\n", failures);
    return failures ? 1 : 0;
}
