// A C++ caller of rocketgpu_amd::BatchSim (include/rocketgpu_amd/batch_sim.hpp), written as a user of the
// reference's BatchSim would write it.  Prints one line per check; the last line is the FNV-1a digest of the car and
// ball records after the run, which tests/test_cpp_client.py compares with the same run through the Python binding.
//
//   batch_sim_client <arenas> <ticks> <seed>
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <vector>

#include "rocketgpu_amd/batch_sim.hpp"

namespace {

constexpr uint64_t kFnvBasis = 0xcbf29ce484222325ull;  // FNV-1a 64 offset basis

uint64_t Fnv1a(uint64_t h, const void* data, size_t n) {
    const unsigned char* p = static_cast<const unsigned char*>(data);
    for (size_t i = 0; i < n; ++i) {
        h ^= p[i];
        h *= 1099511628211ull;
    }
    return h;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc != 4 && argc != 5) {
        std::fprintf(stderr, "usage: batch_sim_client <arenas> <ticks> <seed> [kickoff dump file]\n");
        return 2;
    }
    const int n = std::atoi(argv[1]), ticks = std::atoi(argv[2]);
    const uint64_t seed = std::strtoull(argv[3], nullptr, 10);
    try {
        rocketgpu_amd::SimConfig cfg;
        cfg.env_count = n;
        cfg.max_cars_per_env = 4;
        rocketgpu_amd::BatchSim sim(cfg, 1.0, 256.0);
        const size_t cap = sim.car_capacity();
        std::vector<int32_t> counts(n);
        std::vector<rocketgpu_amd::Team> teams(cap);
        std::vector<rocketgpu_amd::CarConfig> configs(cap);
        for (int e = 0; e < n; ++e) counts[e] = 2 + 2 * (e % 2);
        for (size_t i = 0; i < cap; ++i) {
            teams[i] = (i % 4) >= static_cast<size_t>(counts[i / 4] / 2) ? rocketgpu_amd::Team::Orange
                                                                          : rocketgpu_amd::Team::Blue;
            if (rg_car_preset_config(static_cast<int32_t>(i % 6), &configs[i]) != RG_OK) return 1;
        }
        sim.set_env_car_count(counts.data(), counts.size());
        sim.set_teams(teams.data(), teams.size());
        sim.set_car_configs(configs.data(), configs.size());
        rocketgpu_amd::MutatorConfig mut;
        mut.gravity_z = -600.0f;
        sim.set_mutator_config(mut);
        std::printf("mutator gravity_z %.1f\n", static_cast<double>(sim.mutator_config().gravity_z));
        bool refused = false;
        try {
            rocketgpu_amd::MutatorConfig bad;
            bad.sticky_force = 1.0f;
            sim.set_mutator_config(bad);
        } catch (const std::invalid_argument& e) {
            refused = true;
            std::printf("refused: %s\n", e.what());
        }
        if (!refused) return 1;
        sim.reset_to_kickoff(seed);
        const auto stage_digest = [&](const char* what) {
            std::vector<rocketgpu_amd::CarState> c(cap);
            std::vector<rocketgpu_amd::BallState> b(n);
            sim.download_car_states(c.data(), cap);
            sim.download_ball_states(b.data(), n);
            std::printf("stage %s %016llx\n", what,
                        static_cast<unsigned long long>(Fnv1a(Fnv1a(kFnvBasis, c.data(),
                                                                    cap * sizeof(c[0])),
                                                              b.data(), n * sizeof(b[0]))));
        };
        stage_digest("kickoff");
        if (argc > 4) {
            // Dump the kickoff records (car states, then ball states) for a byte comparison.
            std::vector<rocketgpu_amd::CarState> c(cap);
            std::vector<rocketgpu_amd::BallState> b(n);
            sim.download_car_states(c.data(), cap);
            sim.download_ball_states(b.data(), n);
            if (FILE* f = std::fopen(argv[4], "wb")) {
                std::fwrite(c.data(), sizeof(c[0]), cap, f);
                std::fwrite(b.data(), sizeof(b[0]), n, f);
                std::fclose(f);
            }
        }

        std::vector<rocketgpu_amd::CarControls> controls(static_cast<size_t>(ticks) * cap);
        if (rg_workload_controls(seed, 0, n, 4, ticks, 0.3f, controls.data()) != RG_OK) return 1;
        rocketgpu_amd::CarControls* d_controls = nullptr;
        if (hipMalloc(reinterpret_cast<void**>(&d_controls), controls.size() * sizeof(controls[0])) != hipSuccess) {
            return 1;
        }
        if (hipMemcpy(d_controls, controls.data(), controls.size() * sizeof(controls[0]), hipMemcpyHostToDevice) !=
            hipSuccess) {
            return 1;
        }
        hipStream_t stream = nullptr;
        if (hipStreamCreate(&stream) != hipSuccess) return 1;
        for (int t = 0; t < ticks; t += 8) {
            const int k = ticks - t < 8 ? ticks - t : 8;
            sim.step_with_device_controls(d_controls + static_cast<size_t>(t) * cap, k, static_cast<int32_t>(cap),
                                          stream);
            sim.synchronize(stream);
            stage_digest("step");
        }
        sim.synchronize(stream);

        rocketgpu_amd::BatchSim copy(cfg, 1.0, 256.0);
        copy.set_mutator_config(mut);
        copy.copy_state_from(sim, stream);
        std::vector<rocketgpu_amd::CarState> cars(cap), cars_copy(cap);
        std::vector<rocketgpu_amd::BallState> balls(n);
        sim.download_car_states_async(cars.data(), cap, stream);
        sim.download_ball_states_async(balls.data(), n, stream);
        copy.download_car_states(cars_copy.data(), cap, stream);
        sim.synchronize(stream);
        const uint64_t h = Fnv1a(Fnv1a(kFnvBasis, cars.data(), cap * sizeof(cars[0])), balls.data(),
                                 n * sizeof(balls[0]));
        const uint64_t hc = Fnv1a(kFnvBasis, cars_copy.data(), cap * sizeof(cars[0]));
        const uint64_t hs = Fnv1a(kFnvBasis, cars.data(), cap * sizeof(cars[0]));
        std::printf("copy_state_from %s\n", hc == hs ? "equal" : "DIFFERENT");
        (void)hipFree(d_controls);
        (void)hipStreamDestroy(stream);
        std::printf("digest %016llx\n", static_cast<unsigned long long>(h));
        return hc == hs ? 0 : 1;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
}
