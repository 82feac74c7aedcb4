// Function-level oracle: single phases of the reference CPU step called on records the caller supplies, so that a
// phase of this build can be checked on its own before the whole tick it belongs to exists.  Test infrastructure
// only (compiled into oracle/_ref/librgref.so with the reference's own sources; the product never links it).
#include <cstring>
#include <exception>
#include <string>
#include <vector>

#include "rocketgpu/rocketgpu.h"
#include "rocketgpu/detail/step.hpp"

using namespace rocketgpu;

namespace {
thread_local std::string g_phase_error;
}

extern "C" {

const char* rgref_phase_last_error() { return g_phase_error.c_str(); }

long rgref_wheel_state_size() { return static_cast<long>(sizeof(detail::WheelRuntimeState)); }

// Default-constructed WheelRuntimeState bytes.
void rgref_default_wheel_state(void* out) {
    const detail::WheelRuntimeState w{};
    std::memcpy(out, &w, sizeof(w));
}

// The MutatorConfig defaults the car pre-tick reads, in this order: tick_dt (SimConfig), boost_pad_frame_frac,
// car_mass, car_max_speed, car_max_ang_speed, supersonic_start_speed, supersonic_maintain_min_speed,
// supersonic_maintain_max_time, jump_immediate_speed, jump_min_time, jump_reset_time_pad, jump_max_time,
// doublejump_max_delay, flip_torque_time, flip_initial_vel_scale, flip_forward_impulse_max_speed_scale,
// flip_side_impulse_max_speed_scale, flip_backward_impulse_max_speed_scale, flip_backward_impulse_scale_x,
// boost_spawn_amount, unlimited_flips, unlimited_double_jumps.  Returns the count written.
int rgref_car_params(float* out) {
    const SimConfig cfg{};
    const MutatorConfig& m = cfg.mutator;
    const float v[] = {cfg.tick_dt, m.boost_pad_frame_frac, m.car_mass, m.car_max_speed, m.car_max_ang_speed,
                       m.supersonic_start_speed, m.supersonic_maintain_min_speed, m.supersonic_maintain_max_time,
                       m.jump_immediate_speed, m.jump_min_time, m.jump_reset_time_pad, m.jump_max_time,
                       m.doublejump_max_delay, m.flip_torque_time, m.flip_initial_vel_scale,
                       m.flip_forward_impulse_max_speed_scale, m.flip_side_impulse_max_speed_scale,
                       m.flip_backward_impulse_max_speed_scale, m.flip_backward_impulse_scale_x,
                       m.boost_spawn_amount, static_cast<float>(m.unlimited_flips),
                       static_cast<float>(m.unlimited_double_jumps)};
    std::memcpy(out, v, sizeof(v));
    return static_cast<int>(sizeof(v) / sizeof(float));
}

// QuantizeAxisInput (raycast_wheels_and_contact_prep.hpp:2977) over n inputs.
void rgref_quantize_axis(const float* x, float* out, long n) {
    for (long i = 0; i < n; ++i) out[i] = detail::QuantizeAxisInput(x[i]);
}

}  // extern "C"

extern "C" {

// StepEnvPostCar (env_step.hpp:6902-6926) on n independent cars without a world-side finish (no solver scratch),
// reporting fields when report is nonzero.  Car k is slot 0 of env k of a default config; wheels may be null.
int rgref_car_posttick(long n, void* cars, const void* controls, void* wheels, int report) {
    try {
        SimConfig cfg{};
        cfg.env_count = static_cast<int32_t>(n);
        cfg.max_cars_per_env = 1;
        std::vector<CarConfig> configs(static_cast<size_t>(n));
        CarState* c = static_cast<CarState*>(cars);
        const CarControls* ctl = static_cast<const CarControls*>(controls);
        detail::WheelRuntimeState* w = static_cast<detail::WheelRuntimeState*>(wheels);
        for (long k = 0; k < n; ++k) {
            detail::StepEnvPostCar(static_cast<int>(k), 0, cfg, c, w, nullptr, detail::TickCars{}, configs.data(),
                                   ctl[k], false, report != 0);
        }
        return 0;
    } catch (const std::exception& e) {
        g_phase_error = e.what();
        return 1;
    }
}

}  // extern "C"
