// Oracle shim: a C-ABI around the reference CPU BatchSim (include/rocketgpu/rocketgpu.h:988-1089), compiled against
// the reference's own headers by oracle/Makefile into oracle/_ref/librgref.so.  TEST INFRASTRUCTURE ONLY: tests,
// the fixture generator (tests/golden/make_fixtures.py) and bench.py's CPU baseline load it as the checker; the
// product never links it.
//
// The shim builds the procedural arena (include/rocketgpu_amd/workload.h) as a custom ArenaLayout, the way a
// reference user would (ArenaLayout::meshes, rocketgpu.h:943-959), and exposes the BatchSim calls the parity tests
// need plus a seeded multi-tick driver that records per-tick state digests.
#include <cstdio>
#include <cstring>
#include <exception>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "rocketgpu/rocketgpu.h"
#include "rocketgpu/detail/arena/mesh_bvh.hpp"
#include "rocketgpu/detail/step.hpp"
#include "rocketgpu_amd/workload.h"

using namespace rocketgpu;

namespace {

struct RefSim {
    std::unique_ptr<BatchSim> sim;
    int env_count = 0;
    int max_cars = 0;
    double scale = 1.0;
};

thread_local std::string g_error;

int Fail(const std::exception& e) {
    g_error = e.what();
    return -1;
}

CarConfig PresetConfig(int id) {
    switch (id) {
    case 1: return DominusConfig();
    case 2: return PlankConfig();
    case 3: return BreakoutConfig();
    case 4: return HybridConfig();
    case 5: return MercConfig();
    default: return OctaneConfig();
    }
}

ArenaLayout ProceduralLayout(double scale, double fillet_radius) {
    rgw::ArenaParams p;
    p.scale = scale;
    p.fillet_radius = fillet_radius;
    const rgw::ArenaGeometry g = rgw::BuildProceduralArena(p);
    const rgw::ArenaRulesSpec rules = rgw::ProceduralArenaRules(p);
    ArenaLayout layout;
    ArenaMesh mesh;
    mesh.vertices.reserve(g.vertices.size());
    for (const rgw::V3& v : g.vertices) mesh.vertices.push_back(Vec3(v.x, v.y, v.z));
    mesh.indices = g.indices;
    layout.meshes.push_back(std::move(mesh));
    layout.boost_pads.clear();
    for (int k = 0; k < rules.pad_count; ++k) {
        BoostPadConfig pad;
        pad.pos = Vec3(rules.pad_pos[k].x, rules.pad_pos[k].y, rules.pad_pos[k].z);
        pad.is_big = rules.pad_big[k];
        layout.boost_pads.push_back(pad);
    }
    const auto v3 = [](const rgw::V3& v) { return Vec3(v.x, v.y, v.z); };
    for (int t = 0; t < 2; ++t) {
        layout.goals[t].point = v3(rules.goals[t].point);
        layout.goals[t].normal = v3(rules.goals[t].normal);
        layout.goals[t].min = v3(rules.goals[t].min);
        layout.goals[t].max = v3(rules.goals[t].max);
        for (int k = 0; k < 5; ++k) {
            const rgw::SpawnSpot& s = rules.kickoff[t][k];
            layout.kickoff_spawns[t][k] = ArenaSpawn{Vec3(s.x, s.y, s.z), s.yaw};
        }
        for (int k = 0; k < 4; ++k) {
            const rgw::SpawnSpot& s = rules.respawn[t][k];
            layout.respawn_spawns[t][k] = ArenaSpawn{Vec3(s.x, s.y, s.z), s.yaw};
        }
    }
    layout.ball_spawn = v3(rules.ball_spawn);
    return layout;
}

uint64_t EnvDigest(const BallState& ball, const CarState* cars, int count) {
    uint64_t h = rgw::kFnvOffset;
    const float brot[4] = {ball.rot.w, ball.rot.x, ball.rot.y, ball.rot.z};
    h = rgw::BodyDigest(h, &ball.pos.x, &ball.vel.x, &ball.ang_vel.x, brot);
    for (int i = 0; i < count; ++i) {
        const CarState& c = cars[i];
        const float rot[4] = {c.rot.w, c.rot.x, c.rot.y, c.rot.z};
        h = rgw::BodyDigest(h, &c.pos.x, &c.vel.x, &c.ang_vel.x, rot);
    }
    return h;
}

}  // namespace

extern "C" {

const char* rgref_last_error() { return g_error.c_str(); }

// Struct sizes of the reference build: 0 CarState, 1 BallState, 2 EnvState, 3 BoostPadState, 4 CarControls,
// 5 CarConfig, 6 EnvCommand, 7 PhysicsEvent, 8 PhysicsEventWindow.
long rgref_sizeof(int which) {
    switch (which) {
    case 0: return sizeof(CarState);
    case 1: return sizeof(BallState);
    case 2: return sizeof(EnvState);
    case 3: return sizeof(BoostPadState);
    case 4: return sizeof(CarControls);
    case 5: return sizeof(CarConfig);
    case 6: return sizeof(EnvCommand);
    case 7: return sizeof(PhysicsEvent);
    case 8: return sizeof(PhysicsEventWindow);
    default: return -1;
    }
}

// Field offsets used by the layout-equality test: which 0 CarState, 1 BallState, 2 EnvState, 3 BoostPadState,
// 4 CarControls, 5 CarConfig.  Dotted names address a member of a nested record.
long rgref_offsetof(int which, const char* field) {
    const std::string f(field);
#define RGREF_FIELD(T, name) if (f == #name) return static_cast<long>(offsetof(T, name));
#define RGREF_SUB(T, m, S, name) if (f == #m "." #name) return static_cast<long>(offsetof(T, m) + offsetof(T::S, name));
    if (which == 1) {
        RGREF_SUB(BallState, hs_info, HeatseekerInfo, y_target_dir)
        RGREF_SUB(BallState, hs_info, HeatseekerInfo, cur_target_speed)
        RGREF_SUB(BallState, hs_info, HeatseekerInfo, time_since_hit)
        RGREF_SUB(BallState, ds_info, DropshotInfo, charge_level)
        RGREF_SUB(BallState, ds_info, DropshotInfo, accumulated_hit_force)
        RGREF_SUB(BallState, ds_info, DropshotInfo, y_target_dir)
        RGREF_SUB(BallState, ds_info, DropshotInfo, has_damaged)
        RGREF_SUB(BallState, ds_info, DropshotInfo, last_damage_tick)
        RGREF_FIELD(BallState, vel_impulse_cache) RGREF_FIELD(BallState, tick_count_when_extra_impulse_applied)
        RGREF_FIELD(BallState, basis_right) RGREF_FIELD(BallState, basis_up) RGREF_FIELD(BallState, basis_authoritative)
        RGREF_FIELD(BallState, vel_impulse_cache_bt)
    } else if (which == 2) {
        RGREF_FIELD(EnvState, orange_score) RGREF_FIELD(EnvState, last_scored_team)
        RGREF_FIELD(EnvState, car_visit_present_mask) RGREF_FIELD(EnvState, kickoff_resets)
        RGREF_FIELD(EnvState, countdown_phase) RGREF_FIELD(EnvState, countdown_fires) RGREF_FIELD(EnvState, frame_ticks)
        RGREF_FIELD(EnvState, countdown_count) RGREF_FIELD(EnvState, kickoff_match_order)
        RGREF_FIELD(EnvState, kickoff_match_index) RGREF_FIELD(EnvState, goal_latched)
    } else if (which == 3) {
        RGREF_FIELD(BoostPadState, is_big) RGREF_FIELD(BoostPadState, prev_locked_car_id)
    } else if (which == 4) {
        RGREF_FIELD(CarControls, throttle) RGREF_FIELD(CarControls, steer) RGREF_FIELD(CarControls, pitch)
        RGREF_FIELD(CarControls, yaw) RGREF_FIELD(CarControls, roll) RGREF_FIELD(CarControls, jump)
        RGREF_FIELD(CarControls, boost) RGREF_FIELD(CarControls, handbrake)
    } else if (which == 5) {
        RGREF_FIELD(CarConfig, hitbox_size) RGREF_FIELD(CarConfig, hitbox_offset)
        RGREF_FIELD(CarConfig, front_wheels_offset) RGREF_FIELD(CarConfig, back_wheels_offset)
        RGREF_FIELD(CarConfig, front_wheel_radius) RGREF_FIELD(CarConfig, back_wheel_radius)
        RGREF_FIELD(CarConfig, front_suspension_rest) RGREF_FIELD(CarConfig, back_suspension_rest)
        RGREF_FIELD(CarConfig, dodge_deadzone) RGREF_FIELD(CarConfig, three_wheels)
        RGREF_FIELD(CarConfig, hitbox_published) RGREF_FIELD(CarConfig, body)
        RGREF_FIELD(CarConfig, steer_angle_curve_deg) RGREF_FIELD(CarConfig, powerslide_steer_angle_curve_deg)
        RGREF_FIELD(CarConfig, spawn_z)
    }
    if (which == 0) {
        RGREF_FIELD(CarState, pos)
        RGREF_FIELD(CarState, rot)
        RGREF_FIELD(CarState, vel)
        RGREF_FIELD(CarState, ang_vel)
        RGREF_FIELD(CarState, flip_rel_torque)
        RGREF_FIELD(CarState, jump_time)
        RGREF_FIELD(CarState, flip_time)
        RGREF_FIELD(CarState, air_time)
        RGREF_FIELD(CarState, air_time_since_jump)
        RGREF_FIELD(CarState, boost)
        RGREF_FIELD(CarState, time_since_boosted)
        RGREF_FIELD(CarState, boosting_time)
        RGREF_FIELD(CarState, supersonic_time)
        RGREF_FIELD(CarState, handbrake_val)
        RGREF_FIELD(CarState, auto_flip_timer)
        RGREF_FIELD(CarState, boost_amount)
        RGREF_FIELD(CarState, demo_respawn_timer)
        RGREF_FIELD(CarState, tick_count_since_update)
        RGREF_FIELD(CarState, ball_contact_last_tick)
        RGREF_FIELD(CarState, ball_contact_this_tick)
        RGREF_FIELD(CarState, is_on_ground)
        RGREF_FIELD(CarState, wheels_with_contact)
        RGREF_FIELD(CarState, has_jumped)
        RGREF_FIELD(CarState, has_double_jumped)
        RGREF_FIELD(CarState, has_flipped)
        RGREF_FIELD(CarState, is_flipping)
        RGREF_FIELD(CarState, is_jumping)
        RGREF_FIELD(CarState, is_boosting)
        RGREF_FIELD(CarState, is_supersonic)
        RGREF_FIELD(CarState, is_auto_flipping)
        RGREF_FIELD(CarState, is_demoed)
        RGREF_FIELD(CarState, just_respawned)
        RGREF_FIELD(CarState, demo_no_contact_response)
        RGREF_FIELD(CarState, respawn_slot_override_plus1)
        RGREF_FIELD(CarState, rlbotv5_boost_hold)
        if (f == "world_contact.has_contact") return static_cast<long>(offsetof(CarState, world_contact) + offsetof(decltype(CarState::world_contact), has_contact));
        if (f == "world_contact.wheel_ball_mask") return static_cast<long>(offsetof(CarState, world_contact) + offsetof(decltype(CarState::world_contact), wheel_ball_mask));
        if (f == "world_contact.reserved1") return static_cast<long>(offsetof(CarState, world_contact) + offsetof(decltype(CarState::world_contact), reserved1));
        if (f == "world_contact.reserved2") return static_cast<long>(offsetof(CarState, world_contact) + offsetof(decltype(CarState::world_contact), reserved2));
        if (f == "world_contact.contact_normal") return static_cast<long>(offsetof(CarState, world_contact) + offsetof(decltype(CarState::world_contact), contact_normal));
        if (f == "car_contact.other_car_id") return static_cast<long>(offsetof(CarState, car_contact) + offsetof(decltype(CarState::car_contact), other_car_id));
        if (f == "car_contact.cooldown_timer") return static_cast<long>(offsetof(CarState, car_contact) + offsetof(decltype(CarState::car_contact), cooldown_timer));
        RGREF_FIELD(CarState, vel_impulse_cache)
        if (f == "ball_hit_info.is_valid") return static_cast<long>(offsetof(CarState, ball_hit_info) + offsetof(decltype(CarState::ball_hit_info), is_valid));
        if (f == "ball_hit_info.extra_pending") return static_cast<long>(offsetof(CarState, ball_hit_info) + offsetof(decltype(CarState::ball_hit_info), extra_pending));
        if (f == "ball_hit_info.extra_tick_is_previous") return static_cast<long>(offsetof(CarState, ball_hit_info) + offsetof(decltype(CarState::ball_hit_info), extra_tick_is_previous));
        if (f == "ball_hit_info.tick_count_when_hit") return static_cast<long>(offsetof(CarState, ball_hit_info) + offsetof(decltype(CarState::ball_hit_info), tick_count_when_hit));
        if (f == "ball_hit_info.tick_count_when_extra_impulse_applied") return static_cast<long>(offsetof(CarState, ball_hit_info) + offsetof(decltype(CarState::ball_hit_info), tick_count_when_extra_impulse_applied));
        if (f == "ball_hit_info.hit_car_id") return static_cast<long>(offsetof(CarState, ball_hit_info) + offsetof(decltype(CarState::ball_hit_info), hit_car_id));
        if (f == "ball_hit_info.relative_pos_on_ball") return static_cast<long>(offsetof(CarState, ball_hit_info) + offsetof(decltype(CarState::ball_hit_info), relative_pos_on_ball));
        if (f == "ball_hit_info.ball_pos") return static_cast<long>(offsetof(CarState, ball_hit_info) + offsetof(decltype(CarState::ball_hit_info), ball_pos));
        if (f == "ball_hit_info.extra_hit_vel") return static_cast<long>(offsetof(CarState, ball_hit_info) + offsetof(decltype(CarState::ball_hit_info), extra_hit_vel));
        if (f == "last_controls.throttle") return static_cast<long>(offsetof(CarState, last_controls) + offsetof(decltype(CarState::last_controls), throttle));
        if (f == "last_controls.steer") return static_cast<long>(offsetof(CarState, last_controls) + offsetof(decltype(CarState::last_controls), steer));
        if (f == "last_controls.pitch") return static_cast<long>(offsetof(CarState, last_controls) + offsetof(decltype(CarState::last_controls), pitch));
        if (f == "last_controls.yaw") return static_cast<long>(offsetof(CarState, last_controls) + offsetof(decltype(CarState::last_controls), yaw));
        if (f == "last_controls.roll") return static_cast<long>(offsetof(CarState, last_controls) + offsetof(decltype(CarState::last_controls), roll));
        if (f == "last_controls.jump") return static_cast<long>(offsetof(CarState, last_controls) + offsetof(decltype(CarState::last_controls), jump));
        if (f == "last_controls.boost") return static_cast<long>(offsetof(CarState, last_controls) + offsetof(decltype(CarState::last_controls), boost));
        if (f == "last_controls.handbrake") return static_cast<long>(offsetof(CarState, last_controls) + offsetof(decltype(CarState::last_controls), handbrake));
        RGREF_FIELD(CarState, pos_bt)
        RGREF_FIELD(CarState, vel_bt)
        RGREF_FIELD(CarState, tick_start_pos_bt)
        RGREF_FIELD(CarState, pad_frame_pos_bt)
        RGREF_FIELD(CarState, vel_touched_solve_stage)
        RGREF_FIELD(CarState, basis_authoritative)
        RGREF_FIELD(CarState, ext_force_impulse_bt_valid)
        RGREF_FIELD(CarState, pad_frame_pos_valid)
        RGREF_FIELD(CarState, basis_forward)
        RGREF_FIELD(CarState, basis_right)
        RGREF_FIELD(CarState, basis_up)
        RGREF_FIELD(CarState, ext_force_impulse_bt)
        RGREF_FIELD(CarState, ext_torque_impulse)
        RGREF_FIELD(CarState, solve_delta_linear_bt)
        RGREF_FIELD(CarState, vel_impulse_cache_bt)
        RGREF_FIELD(CarState, solve_delta_angular)
        RGREF_FIELD(CarState, bump_event_bumps)
        RGREF_FIELD(CarState, bump_event_demos)
        RGREF_FIELD(CarState, boost_pad_touching)
        RGREF_FIELD(CarState, bump_event_reserved1)
        RGREF_FIELD(CarState, previous_ball_contact_start)
        RGREF_FIELD(CarState, last_ball_contact_start)
        RGREF_FIELD(CarState, last_world_surface_contact)
        RGREF_FIELD(CarState, first_world_after_previous_ball)
        RGREF_FIELD(CarState, first_world_after_last_ball)
        RGREF_FIELD(CarState, last_flip_or_jump_start)
        RGREF_FIELD(CarState, last_directional_flip_start)
        RGREF_FIELD(CarState, last_ball_started_with_world)
        RGREF_FIELD(CarState, flip_or_jump_last_tick)
        RGREF_FIELD(CarState, internal_jump_time)
        RGREF_FIELD(CarState, internal_flip_time)
        RGREF_FIELD(CarState, flip_timer_live)
        RGREF_FIELD(CarState, air_state)
        RGREF_FIELD(CarState, driving_disabled)
        RGREF_FIELD(CarState, dodge_timeout)
        RGREF_FIELD(CarState, dodge_elapsed)
        RGREF_FIELD(CarState, dodge_dir_x)
        RGREF_FIELD(CarState, dodge_dir_y)
        RGREF_FIELD(CarState, latest_touch_tick)
        RGREF_FIELD(CarState, latest_touch_location)
        RGREF_FIELD(CarState, latest_touch_normal)
        RGREF_FIELD(CarState, internal_vel)
        RGREF_FIELD(CarState, internal_ang_vel)
        RGREF_FIELD(CarState, reported_vel)
        RGREF_FIELD(CarState, reported_ang_vel)
        RGREF_FIELD(CarState, internal_has_jumped)
        RGREF_FIELD(CarState, internal_has_double_jumped)
        RGREF_FIELD(CarState, internal_has_flipped)
        RGREF_FIELD(CarState, internal_is_supersonic)
        RGREF_FIELD(CarState, reported_flags)
        RGREF_FIELD(CarState, game_has_jumped)
        RGREF_FIELD(CarState, double_jump_ticks)
        RGREF_FIELD(CarState, jump_hold_broken)
        RGREF_FIELD(CarState, supersonic_set_v3)
        RGREF_FIELD(CarState, rlbotv5_dodged)
        RGREF_FIELD(CarState, set_wheel_normals)
        RGREF_FIELD(CarState, set_wheel_normals_mask)
        RGREF_FIELD(CarState, cal_pre_flags)
        RGREF_FIELD(CarState, cal_pre_forward_speed_uu)
        RGREF_FIELD(CarState, cal_pre_pitchlock_restart_time)
        RGREF_FIELD(CarState, cal_tick_start_ang_vel)
        RGREF_FIELD(CarState, tick_start_basis_forward)
        RGREF_FIELD(CarState, tick_start_basis_right)
        RGREF_FIELD(CarState, tick_start_basis_up)
        RGREF_FIELD(CarState, restored_inv_inertia_bt)
        RGREF_FIELD(CarState, restored_inv_inertia_valid)
        RGREF_FIELD(CarState, pending_bump_bt)
        RGREF_FIELD(CarState, pending_bump_hitter)
        RGREF_FIELD(CarState, pending_bump_count)
        RGREF_FIELD(CarState, auto_flip_torque_scale)
    } else if (which == 1) {
        RGREF_FIELD(BallState, pos) RGREF_FIELD(BallState, rot) RGREF_FIELD(BallState, vel)
        RGREF_FIELD(BallState, ang_vel) RGREF_FIELD(BallState, pos_bt) RGREF_FIELD(BallState, vel_bt)
        RGREF_FIELD(BallState, basis_forward) RGREF_FIELD(BallState, sleeping)
        RGREF_FIELD(BallState, deactivation_time) RGREF_FIELD(BallState, constrained_2d)
        RGREF_FIELD(BallState, tick_count_since_update) RGREF_FIELD(BallState, restored_inv_inertia_valid)
    } else if (which == 2) {
        RGREF_FIELD(EnvState, tick) RGREF_FIELD(EnvState, blue_score) RGREF_FIELD(EnvState, goal_scored_this_tick)
        RGREF_FIELD(EnvState, car_visit_order_mode) RGREF_FIELD(EnvState, car_visit_order_packed)
        RGREF_FIELD(EnvState, frame_frac_prev) RGREF_FIELD(EnvState, respawn_match_order)
    } else if (which == 3) {
        RGREF_FIELD(BoostPadState, cooldown) RGREF_FIELD(BoostPadState, is_active)
        RGREF_FIELD(BoostPadState, pickup_frac_q) RGREF_FIELD(BoostPadState, cur_locked_car_id)
    } else if (which == 7) {
        RGREF_FIELD(PhysicsEvent, tick) RGREF_FIELD(PhysicsEvent, type) RGREF_FIELD(PhysicsEvent, car_id)
        RGREF_FIELD(PhysicsEvent, other_car_id) RGREF_FIELD(PhysicsEvent, team) RGREF_FIELD(PhysicsEvent, pad_index)
        RGREF_FIELD(PhysicsEvent, car_on_ground) RGREF_FIELD(PhysicsEvent, is_big_pad)
        RGREF_FIELD(PhysicsEvent, is_demo) RGREF_FIELD(PhysicsEvent, position) RGREF_FIELD(PhysicsEvent, normal)
        RGREF_FIELD(PhysicsEvent, car_velocity) RGREF_FIELD(PhysicsEvent, ball_position)
        RGREF_FIELD(PhysicsEvent, ball_velocity) RGREF_FIELD(PhysicsEvent, boost_before)
        RGREF_FIELD(PhysicsEvent, boost_after)
    } else if (which == 8) {
        RGREF_FIELD(PhysicsEventWindow, tick) RGREF_FIELD(PhysicsEventWindow, count)
        RGREF_FIELD(PhysicsEventWindow, overflow)
    }
#undef RGREF_SUB
#undef RGREF_FIELD
    return -1;
}

// flags: bit 0 boost pads, bit 1 car-car collisions, bit 2 demos, bit 3 scoring (rgref_create sets all four).
void* rgref_create_flags(int env_count, int max_cars, double scale, double fillet_radius, uint64_t respawn_seed,
                         int flags) {
    try {
        auto h = std::make_unique<RefSim>();
        SimConfig cfg;
        cfg.env_count = env_count;
        cfg.max_cars_per_env = max_cars;
        cfg.respawn_rng_seed = respawn_seed;
        cfg.enable_boost_pads = (flags & 1) ? 1 : 0;
        cfg.enable_car_car_collisions = (flags & 2) ? 1 : 0;
        cfg.enable_demos = (flags & 4) ? 1 : 0;
        cfg.enable_scoring = (flags & 8) ? 1 : 0;
        h->sim = std::make_unique<BatchSim>(cfg, ProceduralLayout(scale, fillet_radius));
        h->env_count = env_count;
        h->max_cars = max_cars;
        h->scale = scale;
        return h.release();
    } catch (const std::exception& e) {
        Fail(e);
        return nullptr;
    }
}

void* rgref_create(int env_count, int max_cars, double scale, double fillet_radius, uint64_t respawn_seed) {
    return rgref_create_flags(env_count, max_cars, scale, fillet_radius, respawn_seed, 15);
}

void rgref_destroy(void* p) { delete static_cast<RefSim*>(p); }

int rgref_set_env_car_count(void* p, const int32_t* counts, long n) {
    try { static_cast<RefSim*>(p)->sim->set_env_car_count(counts, static_cast<size_t>(n)); return 0; }
    catch (const std::exception& e) { return Fail(e); }
}

int rgref_set_teams(void* p, const uint8_t* teams, long n) {
    try { static_cast<RefSim*>(p)->sim->set_teams(reinterpret_cast<const Team*>(teams), static_cast<size_t>(n)); return 0; }
    catch (const std::exception& e) { return Fail(e); }
}

int rgref_set_car_presets(void* p, const uint8_t* presets, long n) {
    try {
        std::vector<CarConfig> cfgs(static_cast<size_t>(n));
        for (long i = 0; i < n; ++i) cfgs[static_cast<size_t>(i)] = PresetConfig(presets[i]);
        static_cast<RefSim*>(p)->sim->set_car_configs(cfgs.data(), cfgs.size());
        return 0;
    } catch (const std::exception& e) { return Fail(e); }
}

// set_mutator_config with the batch's current MutatorConfig and n float fields replaced by name (gravity_z,
// ball_drag, ball_friction, ball_restitution, car_mass, jump_accel, boost_accel_ground, car_max_speed,
// flip_torque_x, car_air_control_pitch).  Returns -1 for an unknown name.
int rgref_set_mutator_floats(void* p, const char* const* names, const float* values, int n) {
    try {
        RefSim& h = *static_cast<RefSim*>(p);
        MutatorConfig m = h.sim->config().mutator;
        for (int i = 0; i < n; ++i) {
            const std::string k = names[i];
            float* f = k == "gravity_z" ? &m.gravity_z : k == "ball_drag" ? &m.ball_drag
                     : k == "ball_friction" ? &m.ball_friction : k == "ball_restitution" ? &m.ball_restitution
                     : k == "car_mass" ? &m.car_mass : k == "jump_accel" ? &m.jump_accel
                     : k == "boost_accel_ground" ? &m.boost_accel_ground : k == "car_max_speed" ? &m.car_max_speed
                     : k == "flip_torque_x" ? &m.flip_torque_x : k == "car_air_control_pitch" ? &m.car_air_control_pitch
                     : nullptr;
            if (f == nullptr) throw std::invalid_argument("unknown mutator field " + k);
            *f = values[i];
        }
        h.sim->set_mutator_config(m);
        return 0;
    } catch (const std::exception& e) { return Fail(e); }
}

int rgref_reset_to_kickoff(void* p, uint64_t seed) {
    try { static_cast<RefSim*>(p)->sim->reset_to_kickoff(seed); return 0; }
    catch (const std::exception& e) { return Fail(e); }
}

// which: 0 CarState, 1 BallState, 2 EnvState, 3 BoostPadState, 4 CarControls (upload only).
int rgref_upload(void* p, int which, const void* src, long n) {
    try {
        BatchSim& s = *static_cast<RefSim*>(p)->sim;
        const size_t c = static_cast<size_t>(n);
        switch (which) {
        case 0: s.upload_car_states(static_cast<const CarState*>(src), c); break;
        case 1: s.upload_ball_states(static_cast<const BallState*>(src), c); break;
        case 2: s.upload_env_states(static_cast<const EnvState*>(src), c); break;
        case 3: s.upload_boost_pad_states(static_cast<const BoostPadState*>(src), c); break;
        case 4: s.upload_controls(static_cast<const CarControls*>(src), c); break;
        default: throw std::invalid_argument("unknown state array");
        }
        return 0;
    } catch (const std::exception& e) { return Fail(e); }
}

int rgref_download(void* p, int which, void* dst, long n) {
    try {
        const BatchSim& s = *static_cast<RefSim*>(p)->sim;
        const size_t c = static_cast<size_t>(n);
        switch (which) {
        case 0: s.download_car_states(static_cast<CarState*>(dst), c); break;
        case 1: s.download_ball_states(static_cast<BallState*>(dst), c); break;
        case 2: s.download_env_states(static_cast<EnvState*>(dst), c); break;
        case 3: s.download_boost_pad_states(static_cast<BoostPadState*>(dst), c); break;
        default: throw std::invalid_argument("unknown state array");
        }
        return 0;
    } catch (const std::exception& e) { return Fail(e); }
}

int rgref_apply_env_commands(void* p, const void* cmds, long n) {
    try { static_cast<RefSim*>(p)->sim->apply_env_commands(static_cast<const EnvCommand*>(cmds), static_cast<size_t>(n)); return 0; }
    catch (const std::exception& e) { return Fail(e); }
}

// BatchSim::set_physics_event_stream with host buffers (the CPU build's step writes host memory): capacity records
// per env and one window per env; all-null disables.
int rgref_set_event_stream(void* p, void* records, void* windows, uint32_t capacity, int goal_only) {
    try {
        RefSim& r = *static_cast<RefSim*>(p);
        PhysicsEventStream st;
        if (records || windows || capacity) {
            st.records = static_cast<PhysicsEvent*>(records);
            st.windows = static_cast<PhysicsEventWindow*>(windows);
            st.capacity_per_env = capacity;
            st.env_count = static_cast<uint32_t>(r.env_count);
            st.cars_per_env = static_cast<uint32_t>(r.max_cars);
            st.goal_only = goal_only != 0;
        }
        r.sim->set_physics_event_stream(st);
        return 0;
    } catch (const std::exception& e) { return Fail(e); }
}

int rgref_step(void* p, int ticks) {
    try { static_cast<RefSim*>(p)->sim->step(ticks); return 0; }
    catch (const std::exception& e) { return Fail(e); }
}

// BatchSim::ball_path (host pointers in the CPU build): out[env][sample][6], result[env][2].
int rgref_ball_path(void* p, int samples, int ticks, float goal_y, float* out, int32_t* result) {
    try { static_cast<RefSim*>(p)->sim->ball_path(samples, ticks, goal_y, out, result); return 0; }
    catch (const std::exception& e) { return Fail(e); }
}

// BatchSim::download_wheel_runtime_states / upload_wheel_runtime_states (rocketgpu.h:1037-1039): n records of the
// reference's detail::WheelRuntimeState (292 bytes each).
int rgref_download_wheels(void* p, void* dst, long n) {
    try {
        static_cast<RefSim*>(p)->sim->download_wheel_runtime_states(static_cast<detail::WheelRuntimeState*>(dst),
                                                                   static_cast<size_t>(n));
        return 0;
    } catch (const std::exception& e) { return Fail(e); }
}

// StepEnvPreTickCar (env_step.hpp:6854-6873) on every car slot of the batch's arrays as given (host copies, batch
// layout: slot k of env e at e * max_cars + k), with this batch's config (arena rules included) and a respawn seed.
int rgref_car_pretick_batch(void* p, uint64_t seed, void* cars, const void* controls, const void* configs,
                            void* wheels, const void* envs, const uint8_t* teams) {
    try {
        const RefSim& r = *static_cast<RefSim*>(p);
        const SimConfig& cfg = r.sim->config();
        std::vector<Team> team(static_cast<size_t>(r.env_count) * r.max_cars);
        for (size_t i = 0; i < team.size(); ++i) team[i] = teams[i] ? Team::Orange : Team::Blue;
        CarState* c = static_cast<CarState*>(cars);
        const CarControls* ctl = static_cast<const CarControls*>(controls);
        const CarConfig* cc = static_cast<const CarConfig*>(configs);
        detail::WheelRuntimeState* w = static_cast<detail::WheelRuntimeState*>(wheels);
        const EnvState* ev = static_cast<const EnvState*>(envs);
        for (int e = 0; e < r.env_count; ++e) {
            for (int k = 0; k < r.max_cars; ++k) {
                const size_t i = static_cast<size_t>(e) * r.max_cars + k;
                detail::StepEnvPreTickCar(e, k, cfg, team.data(), c, detail::TickCars{}, ev[e], nullptr, seed, ctl[i],
                                          cc[i], w ? &w[i] : nullptr);
            }
        }
        return 0;
    } catch (const std::exception& e) { return Fail(e); }
}

int rgref_download_car_configs(void* p, void* dst, long n) {
    try {
        const DeviceView v = static_cast<RefSim*>(p)->sim->device_view();
        std::memcpy(dst, v.car_configs, static_cast<size_t>(n) * sizeof(CarConfig));
        return 0;
    } catch (const std::exception& e) { return Fail(e); }
}

int rgref_step_with_controls(void* p, const void* controls_by_tick, int ticks, int stride) {
    try {
        static_cast<RefSim*>(p)->sim->step_with_device_controls(static_cast<const CarControls*>(controls_by_tick),
                                                                 ticks, stride);
        return 0;
    } catch (const std::exception& e) { return Fail(e); }
}

// BatchSim::step_with_device_commands (host pointers in the CPU build): controls [tick][control_stride] and
// commands [tick][command_stride].
int rgref_step_with_commands(void* p, const void* controls_by_tick, int control_stride, const void* commands_by_tick,
                             int command_stride, int ticks) {
    try {
        static_cast<RefSim*>(p)->sim->step_with_device_commands(static_cast<const CarControls*>(controls_by_tick),
                                                                 control_stride,
                                                                 static_cast<const EnvCommand*>(commands_by_tick),
                                                                 command_stride, ticks);
        return 0;
    } catch (const std::exception& e) { return Fail(e); }
}

int rgref_apply_env_commands_device(void* p, const void* cmds, long n) {
    try {
        static_cast<RefSim*>(p)->sim->apply_env_commands_device(static_cast<const EnvCommand*>(cmds),
                                                                 static_cast<size_t>(n));
        return 0;
    } catch (const std::exception& e) { return Fail(e); }
}

// Per-env body digests (ball, then cars in slot order up to the env's car count) of the current state.
int rgref_env_digests(void* p, const int32_t* car_counts, uint64_t* out) {
    try {
        RefSim& h = *static_cast<RefSim*>(p);
        std::vector<BallState> balls(static_cast<size_t>(h.env_count));
        std::vector<CarState> cars(static_cast<size_t>(h.env_count) * static_cast<size_t>(h.max_cars));
        h.sim->download_ball_states(balls.data(), balls.size());
        if (!cars.empty()) h.sim->download_car_states(cars.data(), cars.size());
        for (int e = 0; e < h.env_count; ++e) {
            const int n = car_counts ? car_counts[e] : 0;
            out[e] = EnvDigest(balls[static_cast<size_t>(e)], cars.data() + static_cast<size_t>(e) * h.max_cars, n);
        }
        return 0;
    } catch (const std::exception& e) { return Fail(e); }
}

// Mesh tables the reference built for this sim's layout (host memory in the CPU build): counts into counts[0..2]
// (triangles, BVH nodes, meshes).  With non-null outputs: per triangle 16 floats (a, b, c, normal, the three edge
// angles, edge flags as an integer value); per node 4 int32 (qmin[a] | qmax[a] << 16 for a = 0..2, then the
// escape/triangle index); per mesh 15 floats (aabb min/max, quantization min/max/scale) and 4 ints (triangle
// start/count, node offset/count).
int rgref_mesh_tables(void* p, long* counts, float* tris, int32_t* nodes, float* meshes, int32_t* mesh_ints) {
    try {
        const SimConfig& c = static_cast<RefSim*>(p)->sim->config();
        counts[0] = c.arena_mesh_tri_count;
        counts[1] = c.arena_mesh_bvh_node_count;
        counts[2] = c.arena_mesh_bvh_mesh_count;
        if (tris) {
            for (int i = 0; i < c.arena_mesh_tri_count; ++i) {
                const detail::ArenaMeshTriangle& t = c.arena_mesh_tris[i];
                float* o = tris + 16 * static_cast<size_t>(i);
                const Vec3 v[4] = {t.a, t.b, t.c, t.normal};
                for (int k = 0; k < 4; ++k) {
                    o[3 * k] = v[k].x;
                    o[3 * k + 1] = v[k].y;
                    o[3 * k + 2] = v[k].z;
                }
                o[12] = t.edge_v0v1_angle;
                o[13] = t.edge_v1v2_angle;
                o[14] = t.edge_v2v0_angle;
                o[15] = static_cast<float>(t.edge_flags);
            }
        }
        if (nodes) {
            for (int i = 0; i < c.arena_mesh_bvh_node_count; ++i) {
                const detail::ArenaMeshBvhNode& n = c.arena_mesh_bvh_nodes[i];
                int32_t* o = nodes + 4 * static_cast<size_t>(i);
                o[0] = n.quantized_aabb_min[0] | (n.quantized_aabb_max[0] << 16);
                o[1] = n.quantized_aabb_min[1] | (n.quantized_aabb_max[1] << 16);
                o[2] = n.quantized_aabb_min[2] | (n.quantized_aabb_max[2] << 16);
                o[3] = n.escape_index_or_triangle_index;
            }
        }
        if (meshes) {
            for (int i = 0; i < c.arena_mesh_bvh_mesh_count; ++i) {
                const detail::ArenaMeshBvhMeshInfo& m = c.arena_mesh_bvh_meshes[i];
                const detail::ArenaMeshBvhQuantization& q = c.arena_mesh_bvh_quantization[m.mesh_index];
                float* o = meshes + 15 * static_cast<size_t>(i);
                const Vec3 v[5] = {m.aabb_min, m.aabb_max, q.bvh_aabb_min, q.bvh_aabb_max, q.bvh_quantization};
                for (int k = 0; k < 5; ++k) {
                    o[3 * k] = v[k].x;
                    o[3 * k + 1] = v[k].y;
                    o[3 * k + 2] = v[k].z;
                }
                int32_t* oi = mesh_ints + 4 * static_cast<size_t>(i);
                oi[0] = m.triangle_start;
                oi[1] = m.triangle_count;
                oi[2] = m.node_offset;
                oi[3] = m.node_count;
            }
        }
        return 0;
    } catch (const std::exception& e) { return Fail(e); }
}

// Default-constructed record bytes of the reference's structs (same ids as rgref_sizeof).
// Default-constructed records with every padding byte zero (the records are built in zeroed storage), so that
// fixtures written from them are deterministic.
extern "C++" {
template <typename T>
static void DefaultInto(void* out) {
    alignas(T) unsigned char buf[sizeof(T)];
    std::memset(buf, 0, sizeof(T));
    T* v = new (buf) T;
    std::memcpy(out, v, sizeof(T));
    v->~T();
}
}

int rgref_default_state(int which, void* out) {
    switch (which) {
    case 0: DefaultInto<CarState>(out); return 0;
    case 1: DefaultInto<BallState>(out); return 0;
    case 2: DefaultInto<EnvState>(out); return 0;
    case 3: DefaultInto<BoostPadState>(out); return 0;
    case 4: DefaultInto<CarControls>(out); return 0;
    default: return -1;
    }
}

// Ball-launch workload: default BallState records with pos / vel / ang_vel from rgw::SeededBallLaunch.
int rgref_workload_ball_states(uint64_t seed, long first_arena, long count, double scale, void* out) {
    BallState* b = static_cast<BallState*>(out);
    for (long i = 0; i < count; ++i) {
        const rgw::BallLaunch l = rgw::SeededBallLaunch(seed, static_cast<uint64_t>(first_arena + i), scale);
        DefaultInto<BallState>(&b[i]);
        b[i].pos = Vec3(l.pos[0], l.pos[1], l.pos[2]);
        b[i].vel = Vec3(l.vel[0], l.vel[1], l.vel[2]);
        b[i].ang_vel = Vec3(l.ang_vel[0], l.ang_vel[1], l.ang_vel[2]);
    }
    return 0;
}

// The reference's float kernels (include/rocketgpu/detail/math.hpp:36-492, and its sinf/cosf wrappers that use the
// host libm beyond the UCRT range): fn 0 sin, 1 cos, 2 atan2(x, y), 3 pow(x, y), 4 acos, 5 SafeAtan2(x, y) (the
// double atan2 rounded to float, math.hpp:1546-1552).
int rgref_math(int fn, const float* x, const float* y, float* out, long n) {
    for (long i = 0; i < n; ++i) {
        switch (fn) {
        case 0: out[i] = RocketGpuSinf(x[i]); break;
        case 1: out[i] = RocketGpuCosf(x[i]); break;
        case 2: out[i] = RocketGpuUcrtAtan2f(x[i], y[i]); break;
        case 3: out[i] = RocketGpuUcrtPowf(x[i], y[i]); break;
        case 4: out[i] = RocketGpuUcrtAcosf(x[i]); break;
        case 5: out[i] = rocketgpu::detail::SafeAtan2(x[i], y[i]); break;
        default: return -1;
        }
    }
    return 0;
}

// Seeded driver: steps `ticks` ticks one at a time with controls drawn from per-car ControlStreams (seeded by
// ctrl_seed, arena, slot; boost probability p_boost), and after each tick records the FNV-1a of the per-env
// digests into batch_digest[t].  Every `checkpoint_every` ticks (if > 0) the per-env digests are appended to
// env_digests (env_count values per checkpoint).  Returns the number of checkpoints written, or -1.
long rgref_run(void* p, const int32_t* car_counts, uint64_t ctrl_seed, float p_boost, int ticks,
               uint64_t* batch_digest, int checkpoint_every, uint64_t* env_digests) {
    try {
        RefSim& h = *static_cast<RefSim*>(p);
        const size_t cap = static_cast<size_t>(h.env_count) * static_cast<size_t>(h.max_cars);
        std::vector<rgw::ControlStream> streams(cap);
        for (int e = 0; e < h.env_count; ++e)
            for (int s = 0; s < h.max_cars; ++s)
                streams[static_cast<size_t>(e) * h.max_cars + s].seed(ctrl_seed, e, s, p_boost);
        std::vector<CarControls> ctrl(cap);
        std::vector<uint64_t> digests(static_cast<size_t>(h.env_count));
        long checkpoints = 0;
        for (int t = 0; t < ticks; ++t) {
            for (size_t i = 0; i < cap; ++i) {
                const rgw::Controls c = streams[i].next();
                std::memcpy(&ctrl[i], &c, sizeof(CarControls));
            }
            if (cap > 0) {
                h.sim->step_with_device_controls(ctrl.data(), 1, static_cast<int>(cap));
            } else {
                h.sim->step(1);
            }
            if (rgref_env_digests(p, car_counts, digests.data()) != 0) return -1;
            batch_digest[t] = rgw::Fnv1a(rgw::kFnvOffset, digests.data(), digests.size() * sizeof(uint64_t));
            if (checkpoint_every > 0 && (t + 1) % checkpoint_every == 0) {
                std::memcpy(env_digests + static_cast<size_t>(checkpoints) * h.env_count, digests.data(),
                            digests.size() * sizeof(uint64_t));
                ++checkpoints;
            }
        }
        return checkpoints;
    } catch (const std::exception& e) {
        Fail(e);
        return -1;
    }
}

}  // extern "C"
