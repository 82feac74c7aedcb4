// MutatorConfig at the boundary (rocketgpu.h:708-840): the flat rg_mutator_config of rg_capi.h, its defaults (the
// reference's), and how a batch takes one (set_mutator_config, rocketgpu.h:1003).  Fields the step reads from its
// parameters are applied ("car": the car parameters, "ball": the ball parameters, "both", "frame": the render-frame
// fraction of both, "settled": the settled-kickoff switch).  Fields marked "fixed" are constants of this build's
// step (their physics is not built, or is built for the default value only): a config that changes one is refused
// with the field's name, so a batch never steps with a mutator it does not honour.
#pragma once

#include <cstring>
#include <stdexcept>
#include <string>

#include "core/rg_car_phases.hpp"
#include "rocketgpu_amd/rg_capi.h"

namespace rg {

// (type, field, reference default, where the step reads it)
#define RG_MUTATOR_FIELDS(X) \
    X(float, gravity_z, -650.0f, both) \
    X(float, ball_radius, 91.24999f, fixed) \
    X(float, ball_mass, 30.000002f, fixed) \
    X(float, car_mass, 180.0f, car) \
    X(float, ball_drag, 0.03f, ball) \
    X(float, ball_friction, 0.35f, ball) \
    X(float, ball_restitution, 0.60f, ball) \
    X(float, ball_max_speed, 6000.0f, ball) \
    X(float, ball_max_ang_speed, 6.0f, ball) \
    X(float, car_max_speed, 2300.0f, car) \
    X(float, car_max_ang_speed, 5.5f, car) \
    X(float, car_air_throttle_accel, 200.0f / 3.0f, car) \
    X(float, boost_max, 100.0f, car) \
    X(float, boost_used_per_second, 33.3f, car) \
    X(float, boost_min_time, 0.10f, car) \
    X(float, boost_accel_ground, 2975.0f / 3.0f, car) \
    X(float, boost_accel_air, 3175.0f / 3.0f, fixed) \
    X(float, boost_spawn_amount, 33.3f, car) \
    X(float, recharge_boost_per_second, 10.0f, car) \
    X(float, recharge_boost_delay, 0.25f, car) \
    X(float, jump_accel, 4375.0f / 3.0f, car) \
    X(float, jump_immediate_speed, 875.0f / 3.0f, car) \
    X(float, jump_min_time, 0.0249f, car) \
    X(float, jump_reset_time_pad, 1.0f / 40.0f, car) \
    X(float, jump_max_time, 0.20f, car) \
    X(float, doublejump_max_delay, 1.25f, car) \
    X(float, flip_torque_time, 0.65f, car) \
    X(float, flip_torque_min_time, 0.41f, fixed) \
    X(float, flip_pitchlock_time, 1.0f, fixed) \
    X(float, flip_pitchlock_extra_time, 0.30f, car) \
    X(float, flip_initial_vel_scale, 500.0f, car) \
    X(float, flip_torque_x, 260.0f, car) \
    X(float, flip_torque_y, 224.0f, car) \
    X(float, flip_forward_impulse_max_speed_scale, 1.0f, car) \
    X(float, flip_side_impulse_max_speed_scale, 1.9f, car) \
    X(float, flip_backward_impulse_max_speed_scale, 2.5f, car) \
    X(float, flip_backward_impulse_scale_x, 16.0f / 15.0f, car) \
    X(float, supersonic_start_speed, 2200.0f, car) \
    X(float, supersonic_maintain_min_speed, 2100.0f, car) \
    X(float, supersonic_maintain_max_time, 1.0f, car) \
    X(float, powerslide_rise_rate, 5.0f, fixed) \
    X(float, powerslide_fall_rate, 2.0f, fixed) \
    X(float, arena_friction, 0.30f, fixed) \
    X(float, arena_restitution, 0.30f, fixed) \
    X(float, sticky_force, 325.0f, fixed) \
    X(float, ball_car_extra_impulse_z_scale, 0.35f, fixed) \
    X(float, ball_car_extra_impulse_forward_scale, 0.65f, fixed) \
    X(float, ball_car_extra_impulse_max_delta_vel, 4600.0f, fixed) \
    X(float, ball_hit_extra_force_scale, 1.0f, fixed) \
    X(float, bump_force_scale, 1.0f, fixed) \
    X(float, goal_base_threshold_y, 5124.2373046875f, fixed) \
    X(float, bump_cooldown_time, 0.25f, fixed) \
    X(float, bump_min_forward_dist, 64.5f, fixed) \
    X(float, demo_respawn_time, 3.0f, fixed) \
    X(float, boost_pad_cooldown_small, 4.0f, car) \
    X(float, boost_pad_cooldown_big, 10.0f, car) \
    X(float, boost_pad_frame_frac, 0.5f, frame) \
    X(float, frame_start_dt, -1.0f, fixed) \
    X(float, car_air_control_pitch, 130.0f, car) \
    X(float, car_air_control_yaw, 95.0f, car) \
    X(float, car_air_control_roll, 400.0f, car) \
    X(float, car_air_damp_pitch, 30.0f, car) \
    X(float, car_air_damp_yaw, 20.0f, car) \
    X(float, car_air_damp_roll, 50.0f, car) \
    X(uint8_t, recharge_boost_enabled, 0, car) \
    X(uint8_t, unlimited_flips, 0, car) \
    X(uint8_t, unlimited_double_jumps, 0, car) \
    X(uint8_t, enable_team_demos, 0, fixed) \
    X(uint8_t, countdown_on_frame_start, 0, fixed) \
    X(uint8_t, game_mode, 0, fixed) \
    X(uint8_t, ball_no_rot, 1, fixed) \
    X(uint8_t, kickoff_settled, 1, settled)

inline rg_mutator_config DefaultMutatorConfig() {
    rg_mutator_config m;
    std::memset(&m, 0, sizeof(m));
#define RG_MUT_DEFAULT(type, name, def, target) m.name = static_cast<type>(def);
    RG_MUTATOR_FIELDS(RG_MUT_DEFAULT)
#undef RG_MUT_DEFAULT
    return m;
}

// Throws std::invalid_argument naming the first fixed field that differs from its default.
inline void CheckMutatorConfig(const rg_mutator_config& m) {
    const rg_mutator_config d = DefaultMutatorConfig();
    auto fixed = [](const char* target) { return std::strcmp(target, "fixed") == 0; };
#define RG_MUT_CHECK(type, name, def, target)                                                                   \
    if (fixed(#target) && std::memcmp(&m.name, &d.name, sizeof(type)) != 0) {                                 \
        throw std::invalid_argument("mutator field " #name " is fixed at its default (" #def ") in this build"); \
    }
    RG_MUTATOR_FIELDS(RG_MUT_CHECK)
#undef RG_MUT_CHECK
    if (!(m.ball_drag >= 0.0f && m.ball_drag < 1.0f)) throw std::invalid_argument("ball_drag must be in [0, 1)");
}

// The applied fields into the step's parameter blocks.
inline void ApplyMutatorConfig(const rg_mutator_config& m, core::StepParams& p, core::CarParams& c) {
    p.gravity_z = c.gravity_z = m.gravity_z;
    p.ball_drag = m.ball_drag;
    p.ball_damping_scale = m.ball_drag <= 0.0f ? 1.0f : core::UcrtPowf(core::Max(0.0f, 1.0f - m.ball_drag), p.tick_dt);
    p.ball_friction = m.ball_friction;
    p.ball_restitution = m.ball_restitution;
    p.ball_max_speed = m.ball_max_speed;
    p.ball_max_ang_speed = m.ball_max_ang_speed;
    p.frame_frac = c.frame_frac = m.boost_pad_frame_frac;
#define RG_MUT_CAR(type, name, def, target) RG_MUT_CAR_##target(name)
#define RG_MUT_CAR_car(name) c.name = m.name;
#define RG_MUT_CAR_both(name)
#define RG_MUT_CAR_ball(name)
#define RG_MUT_CAR_frame(name)
#define RG_MUT_CAR_settled(name)
#define RG_MUT_CAR_fixed(name)
    RG_MUTATOR_FIELDS(RG_MUT_CAR)
#undef RG_MUT_CAR
#undef RG_MUT_CAR_car
#undef RG_MUT_CAR_both
#undef RG_MUT_CAR_ball
#undef RG_MUT_CAR_frame
#undef RG_MUT_CAR_settled
#undef RG_MUT_CAR_fixed
}

}  // namespace rg
