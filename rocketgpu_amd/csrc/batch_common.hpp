// Host-side pieces shared by the HIP BatchSim and the host-compiled test twin: step parameters from the reference's
// default MutatorConfig, the procedural layout's rules, and upload canonicalization.
#pragma once

#include <cmath>
#include <vector>

#include "core/rg_ball_step.hpp"
#include "core/rg_car_phases.hpp"
#include "mesh_build.hpp"
#include "rocketgpu_amd/workload.h"

namespace rg {

// MutatorConfig defaults (include/rocketgpu/rocketgpu.h:708-840) used by the step, with the ball damping scale
// cached the way BatchSim caches it (src/rocketgpu.cpp:259-270).
inline core::StepParams DefaultStepParams(const rgw::ArenaRulesSpec& rules) {
    core::StepParams p{};
    p.tick_dt = 1.0f / 120.0f;
    p.gravity_z = -650.0f;
    p.ball_radius = 91.24999f;
    p.ball_mass = 30.000002f;
    p.ball_drag = 0.03f;
    p.ball_damping_scale = p.ball_drag <= 0.0f ? 1.0f : core::UcrtPowf(core::Max(0.0f, 1.0f - p.ball_drag), p.tick_dt);
    p.ball_friction = 0.35f;
    p.ball_restitution = 0.60f;
    p.ball_max_speed = 6000.0f;
    p.ball_max_ang_speed = 6.0f;
    p.frame_frac = 0.5f;
    p.enable_scoring = 1;
    for (int t = 0; t < 2; ++t) {
        const rgw::GoalSpec& g = rules.goals[t];
        const float len = std::sqrt(g.normal.x * g.normal.x + g.normal.y * g.normal.y + g.normal.z * g.normal.z);
        p.goal_point[t] = Vec3(g.point.x, g.point.y, g.point.z);
        p.goal_normal[t] = Vec3(g.normal.x / len, g.normal.y / len, g.normal.z / len);
        p.goal_min[t] = Vec3(g.min.x, g.min.y, g.min.z);
        p.goal_max[t] = Vec3(g.max.x, g.max.y, g.max.z);
    }
    p.ball_spawn = Vec3(rules.ball_spawn.x, rules.ball_spawn.y, rules.ball_spawn.z);
    p.pad_count = rules.pad_count;
    p.pad_big_mask = 0;
    for (int k = 0; k < rules.pad_count; ++k) p.pad_big_mask |= static_cast<uint64_t>(rules.pad_big[k] != 0) << k;
    return p;
}

// The layout's spawn spots as the step reads them (MakeArenaRules, src/rocketgpu.cpp:858-868).
inline core::CarSpawnRules SpawnRules(const rgw::ArenaRulesSpec& rules) {
    core::CarSpawnRules r{};
    for (int t = 0; t < 2; ++t) {
        for (int k = 0; k < 5; ++k) {
            const rgw::SpawnSpot& s = rules.kickoff[t][k];
            r.kickoff_pos[t][k] = Vec3(s.x, s.y, s.z);
            r.kickoff_yaw[t][k] = s.yaw;
        }
        for (int k = 0; k < 4; ++k) {
            const rgw::SpawnSpot& s = rules.respawn[t][k];
            r.respawn_pos[t][k] = Vec3(s.x, s.y, s.z);
            r.respawn_yaw[t][k] = s.yaw;
        }
    }
    return r;
}

// The number of floats in a force-phase parameter list.
constexpr int kForcePhaseParamCount = 24;

// Car parameters from a force-phase list, in this order: tick_dt, car_mass, jump_min_time, jump_reset_time_pad,
// jump_max_time, jump_accel, boost_max, boost_used_per_second, boost_min_time, boost_accel_ground,
// recharge_boost_per_second, recharge_boost_delay, recharge_boost_enabled (nonzero: on), car_air_throttle_accel,
// flip_torque_time, flip_pitchlock_extra_time, car_air_control_pitch, _yaw, _roll, car_air_damp_pitch, _yaw, _roll,
// flip_torque_x, flip_torque_y.  The other fields keep their defaults.
inline core::CarParams ForcePhaseParams(const float* p) {
    core::CarParams m{};
    m.tick_dt = p[0];
    m.car_mass = p[1];
    m.jump_min_time = p[2];
    m.jump_reset_time_pad = p[3];
    m.jump_max_time = p[4];
    m.jump_accel = p[5];
    m.boost_max = p[6];
    m.boost_used_per_second = p[7];
    m.boost_min_time = p[8];
    m.boost_accel_ground = p[9];
    m.recharge_boost_per_second = p[10];
    m.recharge_boost_delay = p[11];
    m.recharge_boost_enabled = p[12] != 0.0f ? 1 : 0;
    m.car_air_throttle_accel = p[13];
    m.flip_torque_time = p[14];
    m.flip_pitchlock_extra_time = p[15];
    m.car_air_control_pitch = p[16];
    m.car_air_control_yaw = p[17];
    m.car_air_control_roll = p[18];
    m.car_air_damp_pitch = p[19];
    m.car_air_damp_yaw = p[20];
    m.car_air_damp_roll = p[21];
    m.flip_torque_x = p[22];
    m.flip_torque_y = p[23];
    return m;
}

inline ArenaMeshHost BuildProceduralArenaMesh(const rgw::ArenaParams& params) {
    const rgw::ArenaGeometry g = rgw::BuildProceduralArena(params);
    std::vector<std::vector<Vec3>> verts(1);
    for (const rgw::V3& v : g.vertices) verts[0].push_back(Vec3(v.x, v.y, v.z));
    std::vector<std::vector<int32_t>> idx(1, g.indices);
    return BuildArenaMesh(GroupLayoutMeshes(verts, idx));
}

// uu -> bt for an uploaded linear quantity, preferring the bt value whose x50 reproduces the uu value exactly
// (src/rocketgpu.cpp:321-338).
inline float ImportedScalarBt(float value) {
    const float bt = value * (1.0f / 50.0f);
    if (bt * 50.0f == value) return bt;
    const float up = std::nextafter(bt, INFINITY);
    const float down = std::nextafter(bt, -INFINITY);
    if (up * 50.0f == value) return up;
    if (down * 50.0f == value) return down;
    return bt;
}

// Public ball state upload (CanonicalizeImportedBallStatesInPlace, src/rocketgpu.cpp:373-383).
inline void CanonicalizeBall(BallState& b) {
    b.pos_bt = Vec3(ImportedScalarBt(b.pos.x), ImportedScalarBt(b.pos.y), ImportedScalarBt(b.pos.z));
    b.vel_bt = Vec3(ImportedScalarBt(b.vel.x), ImportedScalarBt(b.vel.y), ImportedScalarBt(b.vel.z));
    b.pos = Vec3(b.pos_bt.x * 50.0f, b.pos_bt.y * 50.0f, b.pos_bt.z * 50.0f);
    b.vel = Vec3(b.vel_bt.x * 50.0f, b.vel_bt.y * 50.0f, b.vel_bt.z * 50.0f);
    b.sleeping = 2;
    b.deactivation_time = 0.0f;
}

// Bullet's btMatrix3x3::getRotation (math.hpp:988-1028): Shepperd's method on the largest diagonal term.
inline Quat BulletBasisToQuat(const Vec3& f, const Vec3& r, const Vec3& u) {
    const float m00 = f.x, m01 = r.x, m02 = u.x;
    const float m10 = f.y, m11 = r.y, m12 = u.y;
    const float m20 = f.z, m21 = r.z, m22 = u.z;
    const float trace = m00 + m11 + m22;
    float q[4], x;  // x y z w before scaling
    if (trace > 0.0f) {
        x = trace + 1.0f;
        q[0] = m21 - m12, q[1] = m02 - m20, q[2] = m10 - m01, q[3] = x;
    } else {
        const int i = m00 < m11 ? (m11 < m22 ? 2 : 1) : (m00 < m22 ? 2 : 0);
        if (i == 0) {
            x = m00 - m11 - m22 + 1.0f;
            q[0] = x, q[1] = m10 + m01, q[2] = m20 + m02, q[3] = m21 - m12;
        } else if (i == 1) {
            x = m11 - m22 - m00 + 1.0f;
            q[0] = m01 + m10, q[1] = x, q[2] = m21 + m12, q[3] = m02 - m20;
        } else {
            x = m22 - m00 - m11 + 1.0f;
            q[0] = m02 + m20, q[1] = m12 + m21, q[2] = x, q[3] = m10 - m01;
        }
    }
    const float scale = 0.5f / std::sqrt(x);
    return Quat(q[3] * scale, q[0] * scale, q[1] * scale, q[2] * scale);
}

// Public car state upload (CanonicalizeImportedCarStatesInPlace, src/rocketgpu.cpp:355-371): a car that has been
// stepped gets its working fields back from the last report (and reports them again afterwards), its contact history
// restarts, position and velocity go through bt, and the orientation becomes an authoritative basis with the Bullet
// quaternion of that basis.
inline void CanonicalizeCar(CarState& c) {
    c.pad_frame_pos_valid = 0;
    constexpr uint8_t kVel = 1, kAngVel = 2, kJumped = 4, kDoubleJumped = 8, kFlipped = 16, kSupersonic = 32,
                      kValid = 128;
    uint8_t kept = 0;
    if (c.tick_count_since_update != 0 && (c.reported_flags & kValid)) {
        kept = kValid;
        if (c.vel.x == c.reported_vel.x && c.vel.y == c.reported_vel.y && c.vel.z == c.reported_vel.z) {
            c.vel = c.internal_vel;
            kept |= kVel;
        }
        if (c.ang_vel.x == c.reported_ang_vel.x && c.ang_vel.y == c.reported_ang_vel.y &&
            c.ang_vel.z == c.reported_ang_vel.z) {
            c.ang_vel = c.internal_ang_vel;
            kept |= kAngVel;
        }
        const uint8_t rf = c.reported_flags;
        core::RestoreWorkingFlag(c.has_jumped, c.internal_has_jumped, kJumped, rf, kept);
        core::RestoreWorkingFlag(c.has_double_jumped, c.internal_has_double_jumped, kDoubleJumped, rf, kept);
        core::RestoreWorkingFlag(c.has_flipped, c.internal_has_flipped, kFlipped, rf, kept);
        core::RestoreWorkingFlag(c.is_supersonic, c.internal_is_supersonic, kSupersonic, rf, kept);
    }
    c.restored_inv_inertia_valid = 0;
    c.world_contact.wheel_ball_mask = 0;
    c.previous_ball_contact_start = c.last_ball_contact_start = c.last_world_surface_contact = ~0ULL;
    c.first_world_after_previous_ball = c.first_world_after_last_ball = ~0ULL;
    c.last_flip_or_jump_start = c.last_directional_flip_start = ~0ULL;
    c.ball_contact_last_tick = c.ball_contact_this_tick = c.last_ball_started_with_world = 0;
    c.flip_or_jump_last_tick = core::HasFlipOrJump(c) ? 1 : 0;
    c.pos_bt = Vec3(ImportedScalarBt(c.pos.x), ImportedScalarBt(c.pos.y), ImportedScalarBt(c.pos.z));
    c.vel_bt = Vec3(ImportedScalarBt(c.vel.x), ImportedScalarBt(c.vel.y), ImportedScalarBt(c.vel.z));
    c.pos = Vec3(c.pos_bt.x * 50.0f, c.pos_bt.y * 50.0f, c.pos_bt.z * 50.0f);
    c.vel = Vec3(c.vel_bt.x * 50.0f, c.vel_bt.y * 50.0f, c.vel_bt.z * 50.0f);
    if (!c.basis_authoritative) {
        core::QuatColumns(c.rot, c.basis_forward, c.basis_right, c.basis_up);
        c.basis_authoritative = 1;
    }
    c.rot = BulletBasisToQuat(c.basis_forward, c.basis_right, c.basis_up);
    if (kept & kValid) {
        c.internal_vel = c.vel;
        c.internal_ang_vel = c.ang_vel;
        c.internal_has_jumped = c.has_jumped;
        c.internal_has_double_jumped = c.has_double_jumped;
        c.internal_has_flipped = c.has_flipped;
        c.internal_is_supersonic = c.is_supersonic;
        if (kept & kVel) c.vel = c.reported_vel; else c.reported_vel = c.vel;
        if (kept & kAngVel) c.ang_vel = c.reported_ang_vel; else c.reported_ang_vel = c.ang_vel;
        uint8_t flags = kValid;
        const auto report = [&](uint8_t& field, uint8_t bit) {
            if (kept & bit) field = (c.reported_flags & bit) ? 1 : 0;
            if (field) flags |= bit;
        };
        report(c.has_jumped, kJumped);
        report(c.has_double_jumped, kDoubleJumped);
        report(c.has_flipped, kFlipped);
        report(c.is_supersonic, kSupersonic);
        c.reported_flags = flags;
    }
}

}  // namespace rg
