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
// boost_spawn_amount, unlimited_flips, unlimited_double_jumps, jump_accel, boost_max, boost_used_per_second,
// boost_min_time, boost_accel_ground, recharge_boost_per_second, recharge_boost_delay, recharge_boost_enabled,
// car_air_throttle_accel, flip_pitchlock_extra_time, car_air_control_pitch, _yaw, _roll, car_air_damp_pitch, _yaw,
// _roll, flip_torque_x, flip_torque_y.  Returns the count written.
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
                       static_cast<float>(m.unlimited_double_jumps), m.jump_accel, m.boost_max,
                       m.boost_used_per_second, m.boost_min_time, m.boost_accel_ground, m.recharge_boost_per_second,
                       m.recharge_boost_delay, static_cast<float>(m.recharge_boost_enabled),
                       m.car_air_throttle_accel, m.flip_pitchlock_extra_time, m.car_air_control_pitch,
                       m.car_air_control_yaw, m.car_air_control_roll, m.car_air_damp_pitch, m.car_air_damp_yaw,
                       m.car_air_damp_roll, m.flip_torque_x, m.flip_torque_y};
    std::memcpy(out, v, sizeof(v));
    return static_cast<int>(sizeof(v) / sizeof(float));
}

// QuantizeAxisInput (raycast_wheels_and_contact_prep.hpp:2977) over n inputs.
void rgref_quantize_axis(const float* x, float* out, long n) {
    for (long i = 0; i < n; ++i) out[i] = detail::QuantizeAxisInput(x[i]);
}

}  // extern "C"

extern "C" {

// A car's force queue as the caller hands it over: the frame's queued sums, jump force and counts, then the first four
// forces and torques (host_validation_and_contact_debug.hpp:184-212), as plain floats and ints.
struct ForceQueueRecord {
    float force[3], ang_accel[3], torque_bt[3], jump_force[3];
    int32_t has_jump_force, n_force, n_torque;
    float force_items[4][3], torque_items[4][3];
};
static_assert(sizeof(ForceQueueRecord) == 156, "force queue record");

static Vec3 V(const float* f) { return Vec3(f[0], f[1], f[2]); }
static void Put(float* f, const Vec3& v) {
    f[0] = v.x;
    f[1] = v.y;
    f[2] = v.z;
}

// UpdateJumpState (phase 0, raycast_wheels_and_contact_prep.hpp:3629-3652), ApplyBoost (phase 1, :3252-3300),
// ApplyAirControl (phase 2, :3048-3241) or ApplyFlipTorque (phase 3, :3432-3484; flags bit 0 passes the control's
// pitch as the raw pitch, else 2) on n independent cars, as the controls step calls them (env_step.hpp:
// 3189-3200, :3309-3331, :3418): with the car's frame (MakeCarFrame), and with the wheel runtime state (so the frame
// queues the forces) when queues is not null, else without it (the force changes the velocity).  Air control reads
// configs (one CarConfig per car) and flags (bit 0: update air control, bit 1: air throttle), with the tick's
// pitch-lock restart time and damping angular velocity taken from the car.  params: tick_dt, car_mass,
// jump_min_time, jump_reset_time_pad, jump_max_time, jump_accel, boost_max, boost_used_per_second, boost_min_time,
// boost_accel_ground, recharge_boost_per_second, recharge_boost_delay, recharge_boost_enabled,
// car_air_throttle_accel, flip_torque_time, flip_pitchlock_extra_time, car_air_control_pitch, _yaw, _roll,
// car_air_damp_pitch, _yaw, _roll, flip_torque_x, flip_torque_y.
int rgref_car_force_phase(long n, void* cars, const void* controls, const void* configs, void* queues, int phase,
                          int flags, const float* params) {
    if (phase < 0 || phase > 3 || (phase >= 2 && configs == nullptr)) return 1;
    MutatorConfig m = SimConfig{}.mutator;
    const float dt = params[0];
    m.car_mass = params[1];
    m.jump_min_time = params[2];
    m.jump_reset_time_pad = params[3];
    m.jump_max_time = params[4];
    m.jump_accel = params[5];
    m.boost_max = params[6];
    m.boost_used_per_second = params[7];
    m.boost_min_time = params[8];
    m.boost_accel_ground = params[9];
    m.recharge_boost_per_second = params[10];
    m.recharge_boost_delay = params[11];
    m.recharge_boost_enabled = params[12] != 0.0f ? 1 : 0;
    m.car_air_throttle_accel = params[13];
    m.flip_torque_time = params[14];
    m.flip_pitchlock_extra_time = params[15];
    m.car_air_control_pitch = params[16];
    m.car_air_control_yaw = params[17];
    m.car_air_control_roll = params[18];
    m.car_air_damp_pitch = params[19];
    m.car_air_damp_yaw = params[20];
    m.car_air_damp_roll = params[21];
    m.flip_torque_x = params[22];
    m.flip_torque_y = params[23];
    const CarConfig* cfg = static_cast<const CarConfig*>(configs);
    CarState* c = static_cast<CarState*>(cars);
    const CarControls* ctl = static_cast<const CarControls*>(controls);
    ForceQueueRecord* q = static_cast<ForceQueueRecord*>(queues);
    for (long k = 0; k < n; ++k) {
        detail::CarFrame frame = phase >= 2 ? detail::MakeCarFrame(c[k], cfg[k], m.car_mass) : detail::CarFrame{};
        detail::WheelRuntimeState ws{};
        if (q) {
            const ForceQueueRecord& r = q[k];
            frame.queue_forces = true;
            frame.queued_force = V(r.force);
            frame.queued_ang_accel = V(r.ang_accel);
            frame.queued_torque_bt = V(r.torque_bt);
            frame.queued_jump_force = V(r.jump_force);
            frame.has_queued_jump_force = r.has_jump_force != 0;
            frame.n_force_items = r.n_force;
            frame.n_torque_items = r.n_torque;
            for (int i = 0; i < 4; ++i) {
                frame.force_items[i] = V(r.force_items[i]);
                frame.torque_items[i] = V(r.torque_items[i]);
            }
        }
        detail::WheelRuntimeState* wsp = q ? &ws : nullptr;
        if (phase == 0) {
            detail::UpdateJumpState(c[k], ctl[k], m, dt, wsp, &frame);
        } else if (phase == 1) {
            detail::ApplyBoost(c[k], ctl[k], m, dt, wsp, &frame);
        } else if (phase == 2) {
            detail::ApplyAirControl(c[k], ctl[k], cfg[k], m, dt, (flags & 1) != 0, wsp, (flags & 2) != 0,
                                    c[k].cal_pre_pitchlock_restart_time, &c[k].cal_tick_start_ang_vel, &frame);
        } else {
            // The step defers the queued flip torque past air control (env_step.hpp:3311-3320); queued here at once.
            Vec3 deferred(0.0f, 0.0f, 0.0f);
            detail::ApplyFlipTorque(c[k], ctl[k], cfg[k], m, dt, wsp, (flags & 1) != 0 ? ctl[k].pitch : 2.0f, &frame,
                                    &deferred);
            if (deferred.x != 0.0f || deferred.y != 0.0f || deferred.z != 0.0f) {
                detail::QueueOfficialTorqueBt(wsp, deferred, &frame);
            }
        }
        if (q) {
            ForceQueueRecord& r = q[k];
            Put(r.force, frame.queued_force);
            Put(r.ang_accel, frame.queued_ang_accel);
            Put(r.torque_bt, frame.queued_torque_bt);
            Put(r.jump_force, frame.queued_jump_force);
            r.has_jump_force = frame.has_queued_jump_force ? 1 : 0;
            r.n_force = frame.n_force_items;
            r.n_torque = frame.n_torque_items;
            for (int i = 0; i < 4; ++i) {
                Put(r.force_items[i], frame.force_items[i]);
                Put(r.torque_items[i], frame.torque_items[i]);
            }
        }
    }
    return 0;
}

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
