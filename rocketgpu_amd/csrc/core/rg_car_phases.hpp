// The per-car phases around the physics of a tick: the pre-tick (before the vehicle update and the world step) and
// the post-tick (after the world step).
//
// Pre-tick, the reference's StepEnvPreTickCar (include/rocketgpu/detail/step/env_step.hpp:6751-6873):
//   1. start: restore the working fields the last report replaced, derive the internal jump/flip timers of a car
//      whose state was just set, count a demoed car's respawn timer down              env_step.hpp:6751-6816
//   2. tick-start speed caps and the game's replicated-state quantization         raycast_wheels...:3993-4008
//   3. finish: the tick-start rigid-body snapshot, the mid-flip z damping, the jump / double jump / dodge / auto-flip
//      state machine with its impulses, and the body velocity clamp               env_step.hpp:6818-6852, 2981-3120
//
// Every expression follows the reference's float operation order (no contraction, correctly rounded division and
// square root); the arithmetic contract is the one rg_math.hpp states.  Three-wheel bodies are outside this file:
// CarPreTickSupported() says which cars it handles, and the caller rejects the rest.
//
// Post-tick (CarPostTick): cooldowns, last controls, pending bumps and velocity caches, the tick counter, and the
// fields the game reports.
//
// Force phases of the controls step (CarUpdateJumpState, CarApplyBoost, CarApplyAirControl, CarApplyFlipTorque):
// the jump hold, boost, air control and flip torque, queueing their forces and torques for the solver.  The wheel
// rays, suspension, friction and auto-roll of that step are not here.
#pragma once

#include "rg_env_commands.hpp"
#include "rg_math.hpp"

namespace rg {
namespace core {

// The MutatorConfig fields the car phases read (rocketgpu.h:708-815), with the reference's defaults.
struct CarParams {
    float tick_dt = 1.0f / 120.0f;
    float frame_frac = 0.5f;  // boost_pad_frame_frac: negative on ticks inside a render frame
    float car_mass = 180.0f;
    float car_max_speed = 2300.0f;
    float car_max_ang_speed = 5.5f;
    float supersonic_start_speed = 2200.0f;
    float supersonic_maintain_min_speed = 2100.0f;
    float supersonic_maintain_max_time = 1.0f;
    float jump_immediate_speed = 875.0f / 3.0f;
    float jump_min_time = 0.0249f;
    float jump_reset_time_pad = 1.0f / 40.0f;
    float jump_max_time = 0.20f;
    float doublejump_max_delay = 1.25f;
    float flip_torque_time = 0.65f;
    float flip_initial_vel_scale = 500.0f;
    float flip_forward_impulse_max_speed_scale = 1.0f;
    float flip_side_impulse_max_speed_scale = 1.9f;
    float flip_backward_impulse_max_speed_scale = 2.5f;
    float flip_backward_impulse_scale_x = 16.0f / 15.0f;
    float boost_spawn_amount = 33.3f;
    uint8_t unlimited_flips = 0;
    uint8_t unlimited_double_jumps = 0;
    // Read by the vehicle update and the world step.
    float gravity_z = -650.0f;
    float ball_mass = 30.0f;
    float boost_pad_cooldown_small = 4.0f;
    float boost_pad_cooldown_big = 10.0f;
    // Read by the force phases (jump hold and boost).
    float jump_accel = 4375.0f / 3.0f;
    float boost_max = 100.0f;
    float boost_used_per_second = 33.3f;
    float boost_min_time = 0.10f;
    float boost_accel_ground = 2975.0f / 3.0f;
    float recharge_boost_per_second = 10.0f;
    float recharge_boost_delay = 0.25f;
    uint8_t recharge_boost_enabled = 0;
    // Read by air control.
    float car_air_throttle_accel = 200.0f / 3.0f;
    float flip_pitchlock_extra_time = 0.30f;
    float car_air_control_pitch = 130.0f;
    float car_air_control_yaw = 95.0f;
    float car_air_control_roll = 400.0f;
    float car_air_damp_pitch = 30.0f;
    float car_air_damp_yaw = 20.0f;
    float car_air_damp_roll = 50.0f;
    // Read by the flip torque.
    float flip_torque_x = 260.0f;
    float flip_torque_y = 224.0f;
    // Read by the car's contact rows against the arena mesh.
    float arena_friction = 0.30f;
    float arena_restitution = 0.30f;
    // Read by car-car contacts and bumps.
    uint8_t enable_car_car = 1;
    uint8_t enable_demos = 1;
    uint8_t enable_team_demos = 0;
    float bump_force_scale = 1.0f;
    float bump_cooldown_time = 0.25f;
};

// ------------------------------------------------------------------------------------------------------------
// Car body helpers (arena_geometry.hpp:1799-1818).

RG_HD void SetCarPosBt(CarState& car, const Vec3& p) {
    car.pos_bt = p;
    car.pos = p * 50.0f;
}
RG_HD void SetCarVelBt(CarState& car, const Vec3& v) {
    car.vel_bt = v;
    car.vel = v * 50.0f;
}

// Rotation matrix columns of a quaternion (math.hpp:901-927): |q|^2 summed x, y, z, then w.
RG_HD void QuatBasis(const Quat& q, Vec3& fwd, Vec3& right, Vec3& up) {
    const float len_sq = ((q.x * q.x + q.y * q.y) + q.z * q.z) + q.w * q.w;
    if (len_sq <= 1e-12f) {
        fwd = Vec3(1.0f, 0.0f, 0.0f);
        right = Vec3(0.0f, 1.0f, 0.0f);
        up = Vec3(0.0f, 0.0f, 1.0f);
        return;
    }
    const float s = 2.0f / len_sq;
    const float xs = q.x * s, ys = q.y * s, zs = q.z * s;
    const float wx = q.w * xs, wy = q.w * ys, wz = q.w * zs;
    const float xx = q.x * xs, xy = q.x * ys, xz = q.x * zs;
    const float yy = q.y * ys, yz = q.y * zs, zz = q.z * zs;
    fwd = Vec3(1.0f - (yy + zz), xy + wz, xz - wy);
    right = Vec3(xy - wz, 1.0f - (xx + zz), yz + wx);
    up = Vec3(xz + wy, yz - wx, 1.0f - (xx + yy));
}

// The car's orientation columns: the authoritative basis when it has one, else from its quaternion
// (math.hpp:1392-1400).
RG_HD void CarBasis(const CarState& car, Vec3& fwd, Vec3& right, Vec3& up) {
    if (car.basis_authoritative) {
        fwd = car.basis_forward;
        right = car.basis_right;
        up = car.basis_up;
        return;
    }
    QuatBasis(car.rot, fwd, right, up);
}

// The game's axes of an FQuat (math.hpp:1133-1139), used for the jump direction and the auto-flip.
RG_HD void UeQuatAxes(const Quat& q, Vec3& fwd, Vec3& right, Vec3& up) {
    const float z2 = q.z + q.z, y2 = q.y + q.y;
    const float x2x = (q.x + q.x) * q.x, x2w = (q.x + q.x) * q.w;
    fwd = Vec3(1.0f - (z2 * q.z + y2 * q.y), z2 * q.w + y2 * q.x, z2 * q.x - y2 * q.w);
    right = Vec3(y2 * q.x - z2 * q.w, 1.0f - (z2 * q.z + x2x), x2w + z2 * q.y);
    up = Vec3(y2 * q.w + z2 * q.x, z2 * q.y - x2w, 1.0f - (y2 * q.y + x2x));
}

// v rotated by a unit quaternion in the game's form (w^2 - |q|^2) v + 2 (q.v) q + 2w (q x v), summed per component
// in the game's order (math.hpp:1141-1148).  Zero components of v are multiplied, not skipped, so signed zeros come
// out as the game's.
RG_HD Vec3 UeQuatRotate(const Quat& q, const Vec3& v) {
    const float qq = (q.y * q.y + q.x * q.x) + q.z * q.z;
    const float d = q.w * q.w - qq;
    const float w2 = q.w + q.w;
    const float dot = (v.x * q.x + v.y * q.y) + v.z * q.z;
    const float t = dot + dot;
    const float cx = v.z * q.y - v.y * q.z, cy = v.x * q.z - v.z * q.x, cz = v.y * q.x - v.x * q.y;
    return Vec3((cx * w2 + d * v.x) + t * q.x, (cy * w2 + d * v.y) + t * q.y, (cz * w2 + d * v.z) + t * q.z);
}

// The game's integer rotators (UE3 FRotator, math.hpp:1442-1533 of the reference): 65,536 units per turn, angles
// read through a 16,384-entry sine table (entry j = sin(2 pi j / 16384), indexed by units / 4), radians rounded to
// units half up.
struct Ue3Rotator {
    int pitch, yaw, roll;
};

RG_HD float Ue3SinTab(int j) {
    const float turns = (static_cast<float>(j) + static_cast<float>(j)) * 3.14159274f;
    return UcrtSinf(turns * 6.103515625e-05f);  // 6.1e-5 = 1 / 16384
}
RG_HD float Ue3RotatorSin(int units) { return Ue3SinTab((units >> 2) & 16383); }
RG_HD float Ue3RotatorCos(int units) { return Ue3RotatorSin(units + 16384); }

// Round to the nearest unit, halves up: round(2 v + 0.5) / 2 in integers.
RG_HD int Ue3Round(float v) { return static_cast<int>(rintf(v + v + 0.5f)) >> 1; }
RG_HD int RadiansToRotator(float radians) { return Ue3Round(radians * 32768.0f / 3.14159274f); }

// The rotator of a quaternion: yaw and pitch from the forward axis, roll from the up and right axes seen in the
// frame turned by the yaw alone (its right axis is (-sin yaw, cos yaw, -0)).
RG_HD Ue3Rotator Ue3QuatToRotator(const Quat& q) {
    Vec3 fwd, right, up;
    UeQuatAxes(q, fwd, right, up);
    Ue3Rotator r;
    r.yaw = RadiansToRotator(UcrtAtan2f(fwd.y, fwd.x));
    const float ground = static_cast<float>(sqrt(static_cast<double>(fwd.y * fwd.y + fwd.x * fwd.x)));
    r.pitch = RadiansToRotator(UcrtAtan2f(fwd.z, ground));
    const Vec3 yawed_right(-Ue3RotatorSin(r.yaw), Ue3RotatorCos(r.yaw), -0.0f);
    r.roll = RadiansToRotator(UcrtAtan2f(Dot(up, yawed_right), Dot(right, yawed_right)));
    return r;
}

// The up axis of a rotator: the third column of its rotation matrix.
RG_HD Vec3 Ue3RotatorUpAxis(const Ue3Rotator& r) {
    const float sp = Ue3RotatorSin(r.pitch), cp = Ue3RotatorCos(r.pitch);
    const float sy = Ue3RotatorSin(r.yaw), cy = Ue3RotatorCos(r.yaw);
    const float sr = Ue3RotatorSin(r.roll), cr = Ue3RotatorCos(r.roll);
    const float cr_sp = cr * sp;
    return Vec3(-(cr_sp * cy + sy * sr), cy * sr - cr_sp * sy, cp * cr);
}

// ------------------------------------------------------------------------------------------------------------
// Inertia (host_validation_and_contact_debug.hpp:156-174, 232-256, 274-289, 528-558).

// Diagonal inertia of the hitbox as Bullet's btBoxShape computes it, in bt.
RG_HD Vec3 BoxInertiaDiagBt(const CarConfig& cfg, float mass) {
    const Vec3 sized = cfg.hitbox_published + Vec3(2.5f, 2.5f, 2.5f);
    const bool has_published = sized.x == cfg.hitbox_size.x && sized.y == cfg.hitbox_size.y && sized.z == cfg.hitbox_size.z;
    const Vec3 published = has_published ? cfg.hitbox_published : cfg.hitbox_size - Vec3(2.5f, 2.5f, 2.5f);
    const Vec3 setup_half = (published * 0.02f) * 0.5f + Vec3(0.025f, 0.025f, 0.025f);
    const Vec3 box_half = (setup_half * 50.0f) * 0.02f;
    const float margin = Min(0.04f, 0.1f * Min(Min(box_half.x, box_half.y), box_half.z));
    const Vec3 margin_v(margin, margin, margin);
    const Vec3 half = (box_half - margin_v) + margin_v;
    const float x = half.x * 2.0f, y = half.y * 2.0f, z = half.z * 2.0f;
    const float x2 = x * x, y2 = y * y, z2 = z * z;
    const float scale = mass / 12.0f;
    return Vec3(scale * (y2 + z2), scale * (x2 + z2), scale * (x2 + y2));
}

RG_HD Vec3 InvInertiaLocalBt(const CarConfig& cfg, float mass) {
    const Vec3 i = BoxInertiaDiagBt(cfg, mass);
    return Vec3(i.x != 0.0f ? 1.0f / i.x : 0.0f, i.y != 0.0f ? 1.0f / i.y : 0.0f, i.z != 0.0f ? 1.0f / i.z : 0.0f);
}

// B diag(inv) B^T from the basis columns, row by row.
RG_HD void InvInertiaWorldRows(const Vec3& f, const Vec3& r, const Vec3& u, const Vec3& inv, Vec3& row0, Vec3& row1,
                               Vec3& row2) {
    const float b00 = f.x, b01 = r.x, b02 = u.x;
    const float b10 = f.y, b11 = r.y, b12 = u.y;
    const float b20 = f.z, b21 = r.z, b22 = u.z;
    const float s00 = b00 * inv.x, s01 = b01 * inv.y, s02 = b02 * inv.z;
    const float s10 = b10 * inv.x, s11 = b11 * inv.y, s12 = b12 * inv.z;
    const float s20 = b20 * inv.x, s21 = b21 * inv.y, s22 = b22 * inv.z;
    row0 = Vec3(s00 * b00 + s01 * b01 + s02 * b02, s00 * b10 + s01 * b11 + s02 * b12, s00 * b20 + s01 * b21 + s02 * b22);
    row1 = Vec3(s10 * b00 + s11 * b01 + s12 * b02, s10 * b10 + s11 * b11 + s12 * b12, s10 * b20 + s11 * b21 + s12 * b22);
    row2 = Vec3(s20 * b00 + s21 * b01 + s22 * b02, s20 * b10 + s21 * b11 + s22 * b12, s20 * b20 + s21 * b21 + s22 * b22);
}

// Bullet's btMatrix3x3::inverse (cofactors over the determinant; 0 for a singular matrix).
RG_HD void InverseRows(const Vec3& a, const Vec3& b, const Vec3& c, Vec3& o0, Vec3& o1, Vec3& o2) {
    const float co_x = b.y * c.z - b.z * c.y;
    const float co_y = b.z * c.x - b.x * c.z;
    const float co_z = b.x * c.y - b.y * c.x;
    const float det = (a.x * co_x + a.y * co_y) + a.z * co_z;
    const float s = det != 0.0f ? 1.0f / det : 0.0f;
    o0 = Vec3(co_x * s, (a.z * c.y - a.y * c.z) * s, (a.y * b.z - a.z * b.y) * s);
    o1 = Vec3(co_y * s, (a.x * c.z - a.z * c.x) * s, (a.z * b.x - a.x * b.z) * s);
    o2 = Vec3(co_z * s, (a.y * c.x - a.x * c.y) * s, (a.x * b.y - a.y * b.x) * s);
}

RG_HD Vec3 MulRows(const Vec3& r0, const Vec3& r1, const Vec3& r2, const Vec3& v) {
    return Vec3(Dot(r0, v), Dot(r1, v), Dot(r2, v));
}

// Inverse diagonal inertia in game units: the hitbox's Bullet half extents with the collision margin taken off and
// put back, scaled to uu (host_validation_and_contact_debug.hpp:43-54, :60-71, :81-87; math.hpp:1554-1560).
RG_HD Vec3 InvInertiaLocalUu(const CarConfig& cfg, float mass) {
    const Vec3 sized = cfg.hitbox_published + Vec3(2.5f, 2.5f, 2.5f);
    const bool has_published = sized.x == cfg.hitbox_size.x && sized.y == cfg.hitbox_size.y && sized.z == cfg.hitbox_size.z;
    const Vec3 published = has_published ? cfg.hitbox_published : cfg.hitbox_size - Vec3(2.5f, 2.5f, 2.5f);
    const Vec3 size = (((published * 0.02f) * 0.5f + Vec3(0.025f, 0.025f, 0.025f)) * 50.0f) * 2.0f;
    const Vec3 half_bt = (size * (1.0f / 50.0f)) / 2.0f;
    const float safe = 0.1f * Min(half_bt.x, Min(half_bt.y, half_bt.z));
    const float margin = safe < 0.04f ? safe : 0.04f;
    const Vec3 half = ((half_bt - Vec3(margin, margin, margin)) + Vec3(margin, margin, margin)) * 50.0f;
    const float x = half.x * 2.0f, y = half.y * 2.0f, z = half.z * 2.0f;
    const float x2 = x * x, y2 = y * y, z2 = z * z;
    const float scale = mass / 12.0f;
    const Vec3 i(scale * (y2 + z2), scale * (x2 + z2), scale * (x2 + y2));
    return Vec3(i.x != 0.0f ? 1.0f / i.x : 0.0f, i.y != 0.0f ? 1.0f / i.y : 0.0f, i.z != 0.0f ? 1.0f / i.z : 0.0f);
}

// World inverse inertia in game units from the tick-start basis, or the restored Bullet rows scaled by 1/50^2
// (host_validation_and_contact_debug.hpp:89-131).
RG_HD void InvInertiaWorldRowsUu(const CarState& car, const CarConfig& cfg, float mass, Vec3& r0, Vec3& r1, Vec3& r2) {
    if (car.restored_inv_inertia_valid) {
        r0 = car.restored_inv_inertia_bt[0] * 0.0004f;
        r1 = car.restored_inv_inertia_bt[1] * 0.0004f;
        r2 = car.restored_inv_inertia_bt[2] * 0.0004f;
        return;
    }
    InvInertiaWorldRows(car.tick_start_basis_forward, car.tick_start_basis_right, car.tick_start_basis_up,
                        InvInertiaLocalUu(cfg, mass), r0, r1, r2);
}

// The per-tick car frame the reference builds when the jump logic needs it (MakeCarFrame,
// host_validation_and_contact_debug.hpp:267-296): the orientation columns and the world inverse inertia of the
// tick-start basis (or the restored one).
struct CarFrame {
    Vec3 fwd, right, up;
    Vec3 inv_row0, inv_row1, inv_row2;
};

RG_HD CarFrame MakeCarFrame(const CarState& car, const CarConfig& cfg, float mass) {
    CarFrame f;
    CarBasis(car, f.fwd, f.right, f.up);
    if (car.restored_inv_inertia_valid) {
        f.inv_row0 = car.restored_inv_inertia_bt[0];
        f.inv_row1 = car.restored_inv_inertia_bt[1];
        f.inv_row2 = car.restored_inv_inertia_bt[2];
    } else {
        InvInertiaWorldRows(car.tick_start_basis_forward, car.tick_start_basis_right, car.tick_start_basis_up,
                            InvInertiaLocalBt(cfg, mass), f.inv_row0, f.inv_row1, f.inv_row2);
    }
    return f;
}

// ------------------------------------------------------------------------------------------------------------
// Impulses (host_validation_and_contact_debug.hpp:488-509).

RG_HD void ApplyCentralImpulseBt(CarState& car, const Vec3& impulse_bt, float mass) {
    const float inv_mass = 1.0f / mass;
    const bool was_touched = car.vel_touched_solve_stage != 0;
    SetCarVelBt(car, car.vel_bt + impulse_bt * inv_mass);
    if (was_touched || car.is_on_ground) car.vel_touched_solve_stage = 1;
}

// An impulse given as a velocity change in uu, taken through bt (mass multiplied in, then divided out).
RG_HD void ApplyCentralImpulseUuDeltaV(CarState& car, const Vec3& dv_uu, float mass) {
    ApplyCentralImpulseBt(car, (dv_uu * (1.0f / 50.0f)) * mass, mass);
}

// The jump impulse or force along the car's up axis, scale in uu, result in bt.
RG_HD Vec3 JumpUpScaledBt(const CarState& car, float scale_uu) {
    Vec3 f, r, u;
    UeQuatAxes(car.rot, f, r, u);
    return Vec3((scale_uu * u.x) * 0.02f, (scale_uu * u.y) * 0.02f, (scale_uu * u.z) * 0.02f);
}

// ------------------------------------------------------------------------------------------------------------
// Controls (raycast_wheels_and_contact_prep.hpp:2977-3030; env_step.hpp:2961-2978).

// An axis input as the game replicates it: one byte, 128 steps below zero and 127 above.
RG_HD float QuantizeAxisInput(float x) {
    const float c = Clamp(x, -1.0f, 1.0f);
    const float y = c < 0.0f ? Max(c * 128.0f, -128.0f) : Min(c * 127.0f, 127.0f);
    const float w = ((y + 128.0f) + (y + 128.0f)) + 0.5f;
    const int byte = (static_cast<int>(rintf(w)) >> 1) & 0xFF;
    const float s = static_cast<float>(byte) - 128.0f;
    // s / 127 as one product and one fused correction step (the reference's table, QuantizeAxisByteValue).
    constexpr float kInv127 = 1.0f / 127.0f;
    const float q = s * kInv127;
    return byte < 0x80 ? s * (1.0f / 128.0f) : fmaf(fmaf(-127.0f, q, s), kInv127, q);
}

// Clamped, quantized axes; `raw_pitch` receives the clamped pitch before quantization.
RG_HD CarControls QuantizeControls(CarControls c, float* raw_pitch) {
    c.roll = Clamp(c.roll, -1.0f, 1.0f);
    c.jump = c.jump ? 1 : 0;
    c.boost = c.boost ? 1 : 0;
    c.handbrake = c.handbrake ? 1 : 0;
    if (raw_pitch != nullptr) *raw_pitch = Clamp(c.pitch, -1.0f, 1.0f);
    c.throttle = QuantizeAxisInput(c.throttle);
    c.steer = QuantizeAxisInput(c.steer);
    c.pitch = QuantizeAxisInput(c.pitch);
    c.yaw = QuantizeAxisInput(c.yaw);
    c.roll = QuantizeAxisInput(c.roll);
    return c;
}

// The controls the jump logic sees: quantized, and only throttle while driving is disabled.
RG_HD CarControls PreTickControls(const CarState& car, CarControls c, float& raw_pitch) {
    c = QuantizeControls(c, &raw_pitch);
    if (car.driving_disabled) {
        const float throttle = c.throttle;
        c = CarControls{};
        c.throttle = throttle;
    }
    return c;
}

// ------------------------------------------------------------------------------------------------------------
// Timers (raycast_wheels_and_contact_prep.hpp:3653-3699).

// n ticks of dt summed one at a time (the game accumulates its timers tick by tick).
RG_HD float AccumulatedTicks(int n, float dt) {
    if (n > 1024) return static_cast<float>(n) * dt;
    float t = 0.0f;
    for (int i = 0; i < n; ++i) t += dt;
    return t;
}

RG_HD float FlipRestartTimeOnGround(const CarState& car, const CarParams& m, float dt) {
    if (car.is_flipping) return 0.0f;
    if (!car.has_flipped) return car.internal_flip_time;
    if (car.internal_flip_time < m.flip_torque_time) return 0.0f;
    int k = static_cast<int>(rintf(car.internal_flip_time / dt)) - static_cast<int>(rintf(m.flip_torque_time / dt));
    if (k < 1) k = 1;
    if (k > 64) return static_cast<float>(k) * dt;
    float t = 0.0f;
    for (int i = 0; i < k; ++i) t += dt;
    return t;
}

// Internal jump/flip timers of a car whose state was just set (tick_count_since_update == 0).
RG_HD void InternalTimersFromSetState(CarState& car, const CarParams& m, float dt) {
    car.internal_jump_time = (car.is_jumping || car.has_jumped) ? car.jump_time : 0.0f;
    const float flip_ticks = rintf(car.flip_time / dt);
    const int n = flip_ticks >= 0.0f ? static_cast<int>(Min(flip_ticks, 1.0e6f)) : 0;
    if (car.is_flipping) {
        car.internal_flip_time = AccumulatedTicks(n > 1 ? n - 1 : 0, dt);
    } else if (car.has_flipped) {
        car.internal_flip_time = AccumulatedTicks(n + static_cast<int>(rintf(m.flip_torque_time / dt)), dt);
    } else {
        car.internal_flip_time = car.flip_time;
    }
}

// The supersonic flag and its maintain timer (arena_ball.hpp:3168-3181).
RG_HD void UpdateSupersonic(CarState& car, const Vec3& vel_uu, const CarParams& m, float dt) {
    const float speed_sq = LengthSq(vel_uu);
    if (speed_sq >= (m.supersonic_start_speed * m.supersonic_start_speed)) {
        car.is_supersonic = 1;
        car.supersonic_time = 0.0f;
    } else if (car.is_supersonic) {
        car.supersonic_time += dt;
        if (car.supersonic_time >= m.supersonic_maintain_max_time ||
            speed_sq < (m.supersonic_maintain_min_speed * m.supersonic_maintain_min_speed)) {
            car.is_supersonic = 0;
            car.supersonic_time = 0.0f;
        }
    }
}

// ------------------------------------------------------------------------------------------------------------
// Jump, double jump, dodge and auto-flip (raycast_wheels_and_contact_prep.hpp:3388-3397, 3485-3750).

RG_HD void StartJump(CarState& car, const CarParams& m) {
    car.is_jumping = 1;
    car.internal_jump_time = 0.0f;
    ApplyCentralImpulseBt(car, JumpUpScaledBt(car, m.jump_immediate_speed * m.car_mass), m.car_mass);
}

// The wheels-down auto-flip can start: touching a surface that faces up while the car's roof faces it.
RG_HD bool AutoFlipGate(const CarState& car) {
    return car.world_contact.has_contact && car.world_contact.contact_normal.z > 0.707f &&
           Dot(car.world_contact.contact_normal, Ue3RotatorUpAxis(Ue3QuatToRotator(car.rot))) < -0.966f;
}

RG_HD void ApplyAutoFlip(CarState& car, const CarParams& m, float dt, bool starts, const CarFrame* frame) {
    if (starts) {
        Vec3 f(1.0f, 0.0f, 0.0f), r(0.0f, 1.0f, 0.0f), u(0.0f, 0.0f, 1.0f);
        if (frame) {
            f = frame->fwd, r = frame->right, u = frame->up;
        } else {
            CarBasis(car, f, r, u);
        }
        const float roll = -UcrtAtan2f(r.z, u.z);
        const float abs_roll = Abs(roll);
        constexpr float kFlipTime = 0.4f;
        car.auto_flip_timer = static_cast<float>(static_cast<double>(kFlipTime) *
                                                 (static_cast<double>(abs_roll) / 3.14159265358979323846));
        car.auto_flip_torque_scale = roll > 0.0f ? 1.0f : -1.0f;
        car.is_auto_flipping = 1;
        Vec3 af, ar, au;
        UeQuatAxes(car.rot, af, ar, au);
        ApplyCentralImpulseUuDeltaV(car, (-au) * 200.0f, m.car_mass);
    }
    if (car.is_auto_flipping) {
        if (car.auto_flip_timer <= 0.0f) {
            car.is_auto_flipping = 0;
            car.auto_flip_timer = 0.0f;
        } else {
            Vec3 af, ar, au;
            UeQuatAxes(car.rot, af, ar, au);
            const Vec3 t_flip = (af * 1000.0f) * car.auto_flip_torque_scale;
            // Inverse inertia rows, and their inverse (the inertia); the frame always exists on this path.
            Vec3 i0, i1, i2;
            InverseRows(frame->inv_row0, frame->inv_row1, frame->inv_row2, i0, i1, i2);
            car.ang_vel += MulRows(frame->inv_row0, frame->inv_row1, frame->inv_row2, MulRows(i0, i1, i2, t_flip));
            car.auto_flip_timer -= dt;
        }
    }
}

// A double jump or a dodge; returns whether it was a dodge.
RG_HD bool TryDoubleJumpOrFlip(CarState& car, const CarControls& controls_in, const CarConfig& cfg,
                               const CarParams& m, float forward_speed_uu, float raw_pitch) {
    CarControls c = controls_in;
    c.pitch = -QuantizeAxisInput(-(raw_pitch <= 1.0f ? raw_pitch : c.pitch));
    const float input_magnitude = Abs(c.yaw) + Abs(c.pitch) + Abs(c.roll);
    const bool is_flip_input = input_magnitude >= cfg.dodge_deadzone;
    const bool fresh = !car.has_double_jumped && !car.has_flipped;
    const bool can_use = (fresh || (is_flip_input ? m.unlimited_flips : m.unlimited_double_jumps)) &&
                         !car.is_auto_flipping;
    if (!can_use) return false;
    if (!is_flip_input) {
        ApplyCentralImpulseBt(car, JumpUpScaledBt(car, m.jump_immediate_speed * m.car_mass), m.car_mass);
        car.has_double_jumped = 1;
        return false;
    }
    car.internal_flip_time = 0.0f;
    car.flip_timer_live = 1;
    car.last_directional_flip_start = car.tick_count_since_update;
    car.has_flipped = 1;
    car.is_flipping = 1;
    Vec3 dir(-c.pitch, c.yaw + c.roll, 0.0f);
    if (Abs(c.yaw + c.roll) < 0.1f && Abs(c.pitch) < 0.1f) {
        dir = Vec3(0.0f, 0.0f, 0.0f);
    } else {
        dir = Normalize(dir, Vec3(0.0f, 0.0f, 0.0f));
    }
    car.flip_rel_torque = Vec3(-dir.y, dir.x, 0.0f);
    if (Abs(dir.x) < 0.1f) dir.x = 0.0f;
    if (Abs(dir.y) < 0.1f) dir.y = 0.0f;
    if (LengthSq(dir) > 0.0f) {
        const float speed_ratio = Abs(forward_speed_uu) / m.car_max_speed;
        const bool backwards = Abs(forward_speed_uu) < 100.0f ? dir.x < 0.0f
                                                               : ((dir.x >= 0.0f) != (forward_speed_uu >= 0.0f));
        const float imp = m.flip_initial_vel_scale * m.car_mass;
        const float imp_x = backwards ? rintf(imp * m.flip_backward_impulse_scale_x) : imp;
        Vec3 v0(dir.x * imp_x, dir.y * imp, 0.0f);
        const float max_scale_x = backwards ? m.flip_backward_impulse_max_speed_scale
                                            : m.flip_forward_impulse_max_speed_scale;
        v0.x *= ((max_scale_x - 1.0f) * speed_ratio) + 1.0f;
        v0.y *= ((m.flip_side_impulse_max_speed_scale - 1.0f) * speed_ratio) + 1.0f;
        const int yaw = Ue3QuatToRotator(car.rot).yaw;
        const int yaw_idx = (yaw >> 2) & 16383;
        const Vec3 fwd2(Ue3SinTab((yaw_idx + 4096) & 16383), Ue3SinTab(yaw_idx), 0.0f);
        const Vec3 right2(-fwd2.y, fwd2.x, 0.0f);
        const Vec3 dv = fwd2 * v0.x + right2 * v0.y;
        ApplyCentralImpulseBt(car, dv * 0.02f, m.car_mass);
    }
    return true;
}

RG_HD void UpdateDoubleJumpOrFlipState(CarState& car, const CarControls& controls, const CarConfig& cfg,
                                       const CarParams& m, bool jump_pressed, float forward_speed_uu, float dt,
                                       bool press_as_air, float raw_pitch, bool* flip_started) {
    const auto reset_on_ground = [&]() {
        car.has_double_jumped = 0;
        car.internal_flip_time = FlipRestartTimeOnGround(car, m, dt);
        car.has_flipped = 0;
        car.air_time = 0.0f;
        car.air_time_since_jump = 0.0f;
    };
    if (car.is_on_ground && !press_as_air) {
        reset_on_ground();
    } else {
        if (car.is_on_ground) reset_on_ground();
        car.air_time += dt;
        if (car.has_jumped && !car.is_jumping) {
            car.air_time_since_jump += dt;
        } else {
            car.air_time_since_jump = 0.0f;
        }
        if (jump_pressed && car.air_time_since_jump < m.doublejump_max_delay) {
            if (TryDoubleJumpOrFlip(car, controls, cfg, m, forward_speed_uu, raw_pitch) && flip_started) {
                *flip_started = true;
            }
            if (car.is_flipping && car.internal_flip_time == 0.0f) car.internal_flip_time = -dt;
        }
    }
    if (car.is_flipping || car.has_flipped || car.flip_timer_live || car.internal_flip_time != 0.0f) {
        car.internal_flip_time += dt;
    }
}

// ------------------------------------------------------------------------------------------------------------
// Tick-start caps and quantization (raycast_wheels_and_contact_prep.hpp:3752-3755, 3917-3960, 3993-4008).

RG_HD void QuantizeCarStateV3(CarState& car, const Vec3& vel_uu) {
    const uint8_t touched = car.vel_touched_solve_stage;
    const Vec3 p = car.pos_bt * 50.0f;
    const Vec3 qp(QuantRoundPos(p.x, 100.0f), QuantRoundPos(p.y, 100.0f), QuantRoundPos(p.z, 100.0f));
    SetCarPosBt(car, qp * (1.0f / 50.0f));
    car.pos = qp;
    const Vec3 q(QuantTruncOffset(vel_uu.x, 100.0f), QuantTruncOffset(vel_uu.y, 100.0f),
                 QuantTruncOffset(vel_uu.z, 100.0f));
    SetCarVelBt(car, q * (1.0f / 50.0f));
    car.vel = q;
    car.vel_touched_solve_stage = touched;
    car.ang_vel = Vec3(QuantTruncOffset(car.ang_vel.x, 10000.0f), QuantTruncOffset(car.ang_vel.y, 10000.0f),
                       QuantTruncOffset(car.ang_vel.z, 10000.0f));
    Vec3 f, r, u;
    CarBasis(car, f, r, u);
    // The tick's inertia basis is the unquantized one.
    car.tick_start_basis_forward = f;
    car.tick_start_basis_right = r;
    car.tick_start_basis_up = u;
    car.rot = GameQuantizeColumns(f, r, u);
    car.basis_forward = f;
    car.basis_right = r;
    car.basis_up = u;
    car.basis_authoritative = 1;
}

RG_HD void TickStartLimitAndQuantize(CarState& car, const CarParams& m) {
    if (car.is_demoed) return;
    Vec3 vel_uu = car.vel_bt * 50.0f;
    const float len_sq = LengthSq(vel_uu);
    if (len_sq > m.car_max_speed * m.car_max_speed) {
        vel_uu = (vel_uu * Rsqrt(len_sq)) * m.car_max_speed;
        SetCarVelBt(car, vel_uu * (1.0f / 50.0f));
    }
    car.ang_vel = ClampLength(car.ang_vel, m.car_max_ang_speed);
    QuantizeCarStateV3(car, vel_uu);
}

// Part of the dodge's vertical damping applied before the tick (raycast_wheels_and_contact_prep.hpp:4010-4027).
RG_HD void FlipZDampPrepass(CarState& car, const CarParams& m, float dt) {
    if (car.is_demoed || car.is_on_ground || !car.is_flipping) return;
    const float flip_time = car.internal_flip_time + dt;
    if (!(flip_time >= 0.15f && flip_time <= m.flip_torque_time)) return;
    if (car.vel_bt.z < 0.0f || flip_time < 0.21f) {
        const float dv_uu = (-car.vel.z) * 0.35f;
        if (Abs(dv_uu) >= 1.0e-4f) {
            const float inv_mass = 1.0f / m.car_mass;
            const float impulse = (dv_uu * 0.02f) * (1.0f / inv_mass);
            const Vec3 v = car.vel_bt;
            SetCarVelBt(car, Vec3(v.x, v.y, v.z + impulse * inv_mass));
        }
    }
}

// Bullet's per-body velocity clamp (env_step.hpp:2916-2939); bit 0: linear clamped, bit 1: angular clamped.
RG_HD int BodyDampingClampCar(CarState& car, const CarParams& m) {
    int clamped = 0;
    const Vec3 v = car.vel_bt;
    const float max_lin = m.car_max_speed * 0.02f;
    const float vsq = (v.x * v.x + v.y * v.y) + v.z * v.z;
    if (0.0f < max_lin && max_lin * max_lin < vsq) {
        const float s = 1.0f / sqrtf(vsq);
        car.vel_bt = Vec3((v.x * s) * max_lin, (v.y * s) * max_lin, (v.z * s) * max_lin);
        clamped |= 1;
    }
    const Vec3 w = car.ang_vel;
    const float max_ang = m.car_max_ang_speed;
    const float lsq = (w.x * w.x + w.y * w.y) + w.z * w.z;
    if (0.0f < max_ang && max_ang * max_ang < lsq) {
        const float s = 1.0f / sqrtf(lsq);
        car.ang_vel = Vec3((w.x * s) * max_ang, (w.y * s) * max_ang, (w.z * s) * max_ang);
        clamped |= 2;
    }
    return clamped;
}

// ------------------------------------------------------------------------------------------------------------
// The pre-physics jump state machine (PreTickCarPrePhysics, env_step.hpp:2981-3120).

// Bits of CarState::cal_pre_flags, the pre-tick's record for the vehicle update: a ground jump started, a flip
// started, the flip's impulse changed the velocity, the car had used its air jump before, supersonic came from a
// set state, the car was flipping / jumping when the tick began.
constexpr uint8_t kPreGroundJumpStarted = 1, kPreFlipStarted = 2, kPreFlipMovedCar = 4, kPreHadAirJump = 8,
                  kPreSupersonicFromSet = 16, kPreFlippingAtStart = 32, kPreWasJumping = 64;

// Forward speed of the tick-start rotation applied to a velocity (SnapshotForwardSpeed,
// raycast_wheels_and_contact_prep.hpp:989-1015): the quaternion's x axis rotated the game's way.
RG_HD float ForwardSpeed(const Quat& q, const Vec3& v) {
    const float w2 = q.w + q.w;
    const float s = q.w * q.w - ((q.y * q.y + q.x * q.x) + q.z * q.z);
    float d = (1.0f * q.x + 0.0f * q.y) + 0.0f * q.z;
    d = d + d;
    const Vec3 fwd(((0.0f * q.y - 0.0f * q.z) * w2 + s * 1.0f) + d * q.x,
                   ((1.0f * q.z - 0.0f * q.x) * w2 + s * 0.0f) + d * q.y,
                   ((0.0f * q.x - 1.0f * q.y) * w2 + s * 0.0f) + d * q.z);
    return (fwd.x * v.x + fwd.y * v.y) + fwd.z * v.z;
}

// The car's jump-related state as the tick found it: every decision below compares against these values, not
// against fields the tick has already changed.
struct JumpTickStart {
    bool flipping, flipped, jumped, double_jumped, jumping, on_ground;
    bool jump_held;        // last tick's jump control
    bool fresh;            // first tick since the state was set
    float flip_time, jump_time;
    int wheels_on_ground;
};

RG_HD JumpTickStart CaptureJumpTickStart(const CarState& car) {
    JumpTickStart s;
    s.flipping = car.is_flipping != 0;
    s.flipped = car.has_flipped != 0;
    s.jumped = car.has_jumped != 0;
    s.double_jumped = car.has_double_jumped != 0;
    s.jumping = car.is_jumping != 0;
    s.on_ground = car.is_on_ground != 0;
    s.jump_held = car.last_controls.jump != 0;
    s.fresh = car.tick_count_since_update == 0;
    s.flip_time = car.internal_flip_time;
    s.jump_time = car.internal_jump_time;
    s.wheels_on_ground = 0;
    for (int w = 0; w < 4; ++w) s.wheels_on_ground += car.wheels_with_contact[w] != 0 ? 1 : 0;
    return s;
}

// The game's own jump bookkeeping, kept in the wheel runtime state (its cal_* fields): whether the game counts the
// car as having jumped and double-jumped, the ground flag it read at the last frame boundary, and whether a render
// frame is open.  `grounded`: three or more wheels on the ground at the tick's start.  Returns true when the game
// refuses a ground jump this tick (it still counts the last jump).
RG_HD bool UpdateGameJumpWindow(WheelRuntimeState& w, const JumpTickStart& s, bool grounded, float frame_frac) {
    if (s.fresh && !w.cal_restored) {
        // A car whose state was just set: the game's view starts from the set flags.
        w.cal_game_jumped = s.jumped ? 1 : 0;
        w.cal_game_double_jumped = (s.double_jumped || s.flipped) && !grounded ? 1 : 0;
        w.cal_cached_ground = grounded ? 1 : 0;
        w.cal_jump_inactive_time = s.jumping ? 0.0f : s.jump_time;
        w.cal_frame_open = 0;
    }
    w.cal_restored = 0;
    if (w.cal_frame_open == 0) {
        // Frame boundary: the game re-reads the ground, and a landing clears both jumps.
        if (grounded && w.cal_cached_ground == 0) w.cal_game_jumped = w.cal_game_double_jumped = 0;
        w.cal_cached_ground = grounded ? 1 : 0;
    }
    const bool refuse = w.cal_game_jumped != 0;
    // On the ground long enough after a jump ended: the game resets its jump.
    const bool reset = w.cal_game_jumped != 0 && grounded && !s.jumping && w.cal_jump_inactive_time >= 0.125f;
    if (reset) {
        w.cal_game_jumped = w.cal_game_double_jumped = 0;
        w.cal_cached_ground = 1;
    }
    w.cal_frame_open = frame_frac < 0.0f ? 1 : 0;
    return refuse;
}

// The game has closed the car's air-jump window: it double-jumped or flipped already, or its jump ended longer ago
// than the double-jump delay.
RG_HD bool GameAirWindowClosed(const WheelRuntimeState& w, const CarParams& m) {
    return w.cal_game_double_jumped != 0 || (w.cal_game_jumped != 0 && w.cal_jump_inactive_time > m.doublejump_max_delay);
}

// The jump state machine before physics (the reference's PreTickCarPrePhysics, env_step.hpp:2981-3120): supersonic
// tracking, what a new jump press becomes (an auto-flip when the car lies on its roof, a ground jump, or a double
// jump / flip in the air, each gated by the game's own jump bookkeeping), and the per-tick values the later phases
// read (cal_pre_flags, the forward speed at the tick's start, the pitch-lock restart time).
RG_HD void PreTickCarPrePhysics(CarState& car, CarControls controls, const CarConfig& cfg, const CarParams& m,
                                float frame_frac, float dt, WheelRuntimeState* ws, const Quat& start_rot,
                                const Vec3& start_vel_uu) {
    car.cal_pre_flags = 0;
    const JumpTickStart s = CaptureJumpTickStart(car);
    car.demo_no_contact_response = car.is_demoed ? 1 : 0;
    if (car.is_demoed) return;
    if (car.driving_disabled) controls.jump = 0;
    // A set state's supersonic flag stands in for this tick's supersonic update.
    const bool supersonic_from_set = car.supersonic_set_v3 && s.fresh;
    if (car.supersonic_set_v3) {
        car.supersonic_set_v3 = 0;
    } else {
        UpdateSupersonic(car, start_vel_uu, m, dt);
    }
    car.vel_touched_solve_stage = 0;

    // A press outside a flip goes to the auto-flip first: one that starts now or one already running.
    const bool press = controls.jump && !s.jump_held;
    const bool press_outside_flip = press && !s.flipping;
    const bool starts_auto_flip = press_outside_flip && !car.is_auto_flipping && AutoFlipGate(car);
    const bool press_to_auto_flip = press_outside_flip && (car.is_auto_flipping || starts_auto_flip);
    CarFrame frame{};
    const bool has_frame = car.is_auto_flipping != 0 ||
                           (press && car.world_contact.has_contact && car.world_contact.contact_normal.z > 0.707f);
    if (has_frame) frame = MakeCarFrame(car, cfg, m.car_mass);

    if (s.fresh && ws != nullptr && ws->cal_restored) car.flip_timer_live = 1;
    const bool flip_clock_runs = !s.flipped && !s.flipping && (s.flip_time > 0.0f || car.flip_timer_live);
    const float pitchlock_restart_time = flip_clock_runs ? s.flip_time : -1.0f;

    const bool grounded = ws != nullptr && s.wheels_on_ground >= 3;
    const bool ground_jump_refused = ws != nullptr && UpdateGameJumpWindow(*ws, s, grounded, frame_frac);
    const float start_forward_speed = ForwardSpeed(start_rot, start_vel_uu);
    const bool ground_jump = press && s.on_ground && !s.jumping && !ground_jump_refused && !press_to_auto_flip;
    if (ground_jump) StartJump(car, m);

    // The press in the air: a double jump or a flip, unless the game's window is closed.
    const bool window_closed = ws != nullptr && GameAirWindowClosed(*ws, m);
    const bool air_press = press && !grounded && !window_closed && !press_to_auto_flip;
    const bool had_air_jump = car.has_double_jumped || car.has_flipped;
    const bool try_air = air_press && !ground_jump;
    float raw_pitch = 2.0f;
    if (try_air) controls = PreTickControls(car, controls, raw_pitch);
    const float dodge_forward_speed = try_air ? ForwardSpeed(car.rot, car.vel) : 0.0f;
    const Vec3 vel_before = car.vel_bt;
    bool flip_started = false;
    UpdateDoubleJumpOrFlipState(car, controls, cfg, m, try_air, dodge_forward_speed, dt, ws != nullptr && try_air,
                                raw_pitch, &flip_started);
    const bool flip_moved_car = car.vel_bt.x != vel_before.x || car.vel_bt.y != vel_before.y ||
                                car.vel_bt.z != vel_before.z;
    ApplyAutoFlip(car, m, dt, starts_auto_flip, has_frame ? &frame : nullptr);

    uint8_t flags = 0;
    flags |= ground_jump ? kPreGroundJumpStarted : 0;
    flags |= flip_started ? kPreFlipStarted : 0;
    flags |= flip_moved_car ? kPreFlipMovedCar : 0;
    flags |= had_air_jump ? kPreHadAirJump : 0;
    flags |= supersonic_from_set ? kPreSupersonicFromSet : 0;
    flags |= s.flipping ? kPreFlippingAtStart : 0;
    flags |= s.jumping ? kPreWasJumping : 0;
    car.cal_pre_flags = flags;
    car.cal_pre_forward_speed_uu = start_forward_speed;
    car.cal_pre_pitchlock_restart_time = pitchlock_restart_time;
}

// ------------------------------------------------------------------------------------------------------------
// Respawns of demoed cars (arena_ball.hpp:2478-2576, 3092-3166; math.hpp:838-863, 1454-1518): the spot, the spawn
// transform, and the state reset.

// The arena's spawn spots (ArenaRules, constants.hpp:57-67): position (uu, the car's spawn height added on top) and
// yaw, per team.
struct CarSpawnRules {
    Vec3 kickoff_pos[2][5];
    float kickoff_yaw[2][5];
    Vec3 respawn_pos[2][4];
    float respawn_yaw[2][4];
};

// Where this tick's respawns come from: the batch seed, the env, its tick, the car's id (env * max_cars + slot + 1)
// and team.
struct CarRespawnContext {
    const CarSpawnRules* rules;
    uint64_t seed;
    int env_idx;
    uint32_t env_tick;
    uint32_t car_id;
    int team;  // 0 Blue, 1 Orange
};

RG_HD float WrapAngle(float a) {
    while (a > Pi()) a -= 2.0f * Pi();
    while (a < -Pi()) a += 2.0f * Pi();
    return a;
}

// The quaternion of a yaw about z (math.hpp:838-863 of the reference): within 1e-6 of an odd multiple of an eighth
// turn or of a quarter turn, the half angle's exact sine and cosine; otherwise UCRT cos and sin of the half angle.
RG_HD Quat FromYaw(float yaw) {
    constexpr float kSin8 = 0.38268343236508977173f, kCos8 = 0.92387953251128675613f;  // sin, cos of pi / 8
    constexpr float kHalf = 0.7071067811865475244f;                                       // sin, cos of pi / 4
    struct Exact {
        float angle, w, z;
    };
    const Exact exact[6] = {{Pi() * 0.25f, kCos8, kSin8},   {-(Pi() * 0.25f), kCos8, -kSin8},
                            {Pi() * 0.75f, kSin8, kCos8},   {-(Pi() * 0.75f), kSin8, -kCos8},
                            {HalfPi(), kHalf, kHalf},        {-HalfPi(), kHalf, -kHalf}};
    for (const Exact& e : exact) {
        if (Abs(yaw - e.angle) <= 1e-6f) return Quat(e.w, 0.0f, 0.0f, e.z);
    }
    const float half = 0.5f * yaw;
    return Quat(GameCosf(half), 0.0f, 0.0f, GameSinf(half));
}

// The game's quaternion of a rotation matrix m (rows: the rotated x, y, z axes), UE3's FQuat(FMatrix): the trace
// branch, else the branch of the largest diagonal entry; 1 / sqrt is taken first and the scalar part is 0.5 / that.
RG_HD Quat UeQuatFromMatrix(const float m[3][3]) {
    const float trace = (m[1][1] + m[0][0]) + m[2][2];
    if (trace > 0.0f) {
        const float root = 1.0f / sqrtf(trace + 1.0f);
        const float h = root * 0.5f;
        return Quat(0.5f / root, (m[1][2] - m[2][1]) * h, (m[2][0] - m[0][2]) * h, (m[0][1] - m[1][0]) * h);
    }
    int a = m[0][0] < m[1][1] ? 1 : 0;
    if (m[a][a] <= m[2][2] && m[2][2] != m[a][a]) a = 2;
    const int b = (a + 1) % 3, c = (a + 2) % 3;
    const float root = 1.0f / sqrtf(((m[a][a] - m[b][b]) - m[c][c]) + 1.0f);
    const float h = root * 0.5f;
    float v[3];
    v[a] = 0.5f / root;
    v[b] = (m[a][b] + m[b][a]) * h;
    v[c] = (m[a][c] + m[c][a]) * h;
    return Quat((m[b][c] - m[c][b]) * h, v[0], v[1], v[2]);
}

// The rotation matrix of a pure yaw (pitch and roll 0, their table cosine kept as the game computes it), and its
// quaternion.
RG_HD Quat Ue3QuatFromYawRotator(int yaw) {
    const float sy = Ue3RotatorSin(yaw), cy = Ue3RotatorCos(yaw);
    const float cp = Ue3RotatorCos(0);
    const float m[3][3] = {{cy * cp, cp * sy, 0.0f}, {-(cp * sy), cy * cp, -0.0f}, {-0.0f, 0.0f, cp * cp}};
    return UeQuatFromMatrix(m);
}

// UE3's FQuat product a * b, its terms in the game's order.
RG_HD Quat UeQuatMul(const Quat& a, const Quat& b) {
    const float w = ((a.w * b.w - a.x * b.x) - a.y * b.y) - a.z * b.z;
    const float x = ((a.w * b.x + a.x * b.w) + a.y * b.z) - a.z * b.y;
    const float y = ((a.w * b.y + a.y * b.w) + a.z * b.x) - a.x * b.z;
    const float z = ((a.w * b.z + a.z * b.w) + a.x * b.y) - a.y * b.x;
    return Quat(w, x, y, z);
}

RG_HD int SpawnYawRotatorUnits(float yaw) { return static_cast<int>(yaw * (32768.0f / 3.14159274f) + 0.5f); }

// The spawned car's authoritative basis from its yaw in rotator units (half a turn more for Orange).
RG_HD void SeedCarBasisFromSpawnRotator(CarState& car, int blue_yaw, bool orange) {
    Quat q = Ue3QuatFromYawRotator(blue_yaw);
    if (orange) q = UeQuatMul(Ue3QuatFromYawRotator(32768), q);
    GameQuatToColumns(q, car.basis_forward, car.basis_right, car.basis_up);
    car.basis_authoritative = 1;
}

// Everything a spawn resets (ResetCarCommonState).
RG_HD void ResetCarCommonState(CarState& car, float spawn_boost) {
    car.restored_inv_inertia_valid = 0;
    SetCarVelBt(car, Vec3(0.0f, 0.0f, 0.0f));
    car.ang_vel = Vec3(0.0f, 0.0f, 0.0f);
    car.world_contact.has_contact = 0;
    car.world_contact.wheel_ball_mask = 0;
    car.world_contact.reserved1 = 0;
    car.world_contact.reserved2 = 0;
    car.world_contact.contact_normal = Vec3(0.0f, 0.0f, 0.0f);
    car.flip_rel_torque = Vec3(0.0f, 0.0f, 0.0f);
    car.boost = spawn_boost;
    car.boost_amount = spawn_boost / 100.0f;
    car.jump_time = 0.0f;
    car.flip_time = 0.0f;
    car.internal_jump_time = 0.0f;
    car.internal_flip_time = 0.0f;
    car.flip_timer_live = 1;
    car.air_time = 0.0f;
    car.air_time_since_jump = 0.0f;
    car.boosting_time = 0.0f;
    car.time_since_boosted = 0.0f;
    car.supersonic_time = 0.0f;
    car.handbrake_val = 0.0f;
    car.auto_flip_timer = 0.0f;
    car.auto_flip_torque_scale = 0.0f;
    car.demo_respawn_timer = 0.0f;
    car.tick_count_since_update = 0;
    car.is_on_ground = 1;
    car.air_state = 0;
    car.rlbotv5_dodged = 0;
    car.rlbotv5_boost_hold = 0;
    car.dodge_timeout = -1.0f;
    car.dodge_elapsed = 0.0f;
    car.dodge_dir_x = 0.0f;
    car.dodge_dir_y = 0.0f;
    for (int w = 0; w < 4; ++w) car.wheels_with_contact[w] = 0;
    car.has_jumped = 0;
    car.has_double_jumped = 0;
    car.has_flipped = 0;
    car.is_flipping = 0;
    car.is_jumping = 0;
    car.jump_hold_broken = 0;
    car.is_boosting = 0;
    car.is_supersonic = 0;
    car.is_auto_flipping = 0;
    car.is_demoed = 0;
    car.just_respawned = 0;
    car.car_contact.other_car_id = 0;
    car.car_contact.cooldown_timer = 0.0f;
    car.vel_impulse_cache = Vec3(0.0f, 0.0f, 0.0f);
    car.vel_impulse_cache_bt = Vec3(0.0f, 0.0f, 0.0f);
    car.pending_bump_count = 0;
    car.ball_hit_info = BallHitInfo{};
    car.last_controls = CarControls{};
    car.boost_pad_touching = 0;
    car.driving_disabled = 0;
    car.respawn_slot_override_plus1 = 0;
    car.pad_frame_pos_valid = 0;
}

RG_HD void PlaceCarAtWorldSpawn(CarState& car, const Vec3& pos, float yaw, float z, float boost) {
    SetCarPosBt(car, Vec3(pos.x, pos.y, pos.z + z) * (1.0f / 50.0f));
    car.rot = FromYaw(WrapAngle(yaw));
    SeedCarBasisFromSpawnRotator(car, SpawnYawRotatorUnits(yaw), false);
    ResetCarCommonState(car, boost);
}

// Respawn spot `spot` of the team, moved 192 uu in direction dir (0-7, eighths of a turn from the spot's facing)
// when dir >= 0 (RespawnPlace).
RG_HD void RespawnPlace(const CarSpawnRules& rules, int team, int spot, int dir, Vec3& ground, float& yaw) {
    ground = rules.respawn_pos[team][spot & 3];
    yaw = rules.respawn_yaw[team][spot & 3];
    if (dir < 0) return;
    // The spot's facing (cos yaw, sin yaw) from its yaw quaternion, turned by dir eighths of a turn.
    constexpr float h = 0.70710677f;
    constexpr float kEighthCos[8] = {1.0f, h, 0.0f, -h, -1.0f, -h, 0.0f, h};
    constexpr float kEighthSin[8] = {0.0f, h, 1.0f, h, 0.0f, -h, -1.0f, -h};
    const Quat q = FromYaw(WrapAngle(yaw));
    const float face_x = q.w * q.w - q.z * q.z;
    const float face_y = 2.0f * q.w * q.z;
    const float c = kEighthCos[dir & 7], sn = kEighthSin[dir & 7];
    ground.x += 192.0f * (face_x * c - face_y * sn);
    ground.y += 192.0f * (face_x * sn + face_y * c);
}

RG_HD void RespawnCar(CarState& car, const CarSpawnRules& rules, int team, int slot, float spawn_z, float boost,
                      int offset_dir) {
    Vec3 ground;
    float yaw;
    RespawnPlace(rules, team, slot % 4, offset_dir, ground, yaw);
    PlaceCarAtWorldSpawn(car, ground, yaw, spawn_z, boost);
    car.vel_impulse_cache = Vec3(0.0f, 0.0f, 0.0f);
    car.is_on_ground = 0;
}

RG_HD void KickoffSpawnCar(CarState& car, const CarSpawnRules& rules, int team, int spot, float spawn_z, float boost) {
    const uint16_t driving_disabled = car.driving_disabled;
    const int s = spot >= 0 && spot < 4 ? spot : 4;
    PlaceCarAtWorldSpawn(car, rules.kickoff_pos[team][s], rules.kickoff_yaw[team][s], spawn_z, boost);
    car.driving_disabled = driving_disabled;
    car.vel_impulse_cache = Vec3(0.0f, 0.0f, 0.0f);
}

// The respawn spot a demoed car draws when nothing chose one (RespawnStartIndex, arena_ball.hpp:3092-3096).
RG_HD int RespawnStartIndex(uint64_t seed, int env_idx, uint64_t tick, uint32_t car_id) {
    uint32_t state = MixEnvSeed(seed ^ (tick * 0x9e3779b97f4a7c15ull), env_idx, static_cast<int>(car_id * 0x45d9f3bu));
    state = NextLcg(state);
    return LcgPick(state, 4);
}

// The tick's inertia basis of a car that just spawned: its yaw alone, through the game's rotator and matrix forms
// (SpawnInertiaBasis, math.hpp:1513-1517).
RG_HD void SpawnInertiaBasis(const Quat& spawn_rot, Vec3& forward, Vec3& right, Vec3& up) {
    Vec3 f, r, u;
    UeQuatAxes(Ue3QuatFromYawRotator(Ue3QuatToRotator(spawn_rot).yaw), f, r, u);
    GameQuatToColumns(GameMatrixToQuat(f, r, u), forward, right, up);
}

// ------------------------------------------------------------------------------------------------------------
// The car pre-tick (StepEnvPreTickCar).

// Working fields the last report replaced (CarState::RestoreWorkingFields, rocketgpu.h:398-417), or the first-tick
// derivations of a car whose state was just set (env_step.hpp:5914-5925).
RG_HD void RestoreWorkingFlag(uint8_t& field, uint8_t internal, uint8_t bit, uint8_t reported, uint8_t& kept) {
    if ((field != 0) == ((reported & bit) != 0)) {
        field = internal;
        kept |= bit;
    }
}

RG_HD void RestoreStepCarFields(CarState& car) {
    if (car.tick_count_since_update != 0) {
        constexpr uint8_t kVel = 1, kAngVel = 2, kJumped = 4, kDoubleJumped = 8, kFlipped = 16, kSupersonic = 32,
                          kValid = 128;
        if (!(car.reported_flags & kValid)) return;
        uint8_t kept = kValid;
        if (car.vel.x == car.reported_vel.x && car.vel.y == car.reported_vel.y && car.vel.z == car.reported_vel.z) {
            car.vel = car.internal_vel;
            kept |= kVel;
        }
        if (car.ang_vel.x == car.reported_ang_vel.x && car.ang_vel.y == car.reported_ang_vel.y &&
            car.ang_vel.z == car.reported_ang_vel.z) {
            car.ang_vel = car.internal_ang_vel;
            kept |= kAngVel;
        }
        const uint8_t rf = car.reported_flags;
        RestoreWorkingFlag(car.has_jumped, car.internal_has_jumped, kJumped, rf, kept);
        RestoreWorkingFlag(car.has_double_jumped, car.internal_has_double_jumped, kDoubleJumped, rf, kept);
        RestoreWorkingFlag(car.has_flipped, car.internal_has_flipped, kFlipped, rf, kept);
        RestoreWorkingFlag(car.is_supersonic, car.internal_is_supersonic, kSupersonic, rf, kept);
        return;
    }
    if (car.supersonic_time > 0.0f) car.is_supersonic = 1;
    car.internal_has_double_jumped = car.has_double_jumped;
    car.game_has_jumped = car.has_jumped;
    car.double_jump_ticks = 255;
}

// Whether this file steps the car's pre-tick: not a three-wheel body (its jump torques are not built).
RG_HD bool CarPreTickSupported(const CarState& car, const CarConfig& cfg, const CarParams& m) {
    (void)car;
    (void)m;
    return cfg.three_wheels == 0;
}

// One car's pre-tick (StepEnvPreTickCar).  `ws` may be null (no wheel runtime state).
RG_HD void CarPreTick(CarState& car, const CarControls& controls, const CarConfig& cfg, const CarParams& m,
                      WheelRuntimeState* ws, const CarRespawnContext& rc, Vec3* start_vel_out = nullptr) {
    const float dt = m.tick_dt;
    // 1. start.
    car.ball_contact_this_tick = 0;
    RestoreStepCarFields(car);
    if (car.tick_count_since_update == 0) InternalTimersFromSetState(car, m, dt);
    if (car.just_respawned) car.just_respawned = 0;
    bool respawned = false;
    // respawn_slot_override_plus1: 5 + spot is a kickoff re-creation at kickoff spot `spot`, 128 + 8 * spot + dir a
    // respawn 192 uu off respawn spot `spot` (rocketgpu.h:527-533), any other nonzero value 1 + the respawn spot.
    constexpr int kKickoffSpawnOverridePlus1 = 5, kRespawnOffsetOverridePlus1 = 128;
    if (car.is_demoed) {
        car.demo_respawn_timer = Max(0.0f, car.demo_respawn_timer - dt);
        if (car.demo_respawn_timer <= 0.0f) {
            int slot = -1, offset_dir = -1;
            const int ov = car.respawn_slot_override_plus1;
            if (ov >= kRespawnOffsetOverridePlus1 && ov < kRespawnOffsetOverridePlus1 + 32) {
                offset_dir = (ov - kRespawnOffsetOverridePlus1) & 7;
                slot = ((ov - kRespawnOffsetOverridePlus1) >> 3) & 3;
                car.respawn_slot_override_plus1 = 0;
            } else if (ov != 0) {
                slot = ov - 1;
                car.respawn_slot_override_plus1 = 0;
            }
            if (slot < 0) slot = RespawnStartIndex(rc.seed, rc.env_idx, rc.env_tick, rc.car_id);
            if (offset_dir >= 0) {
                RespawnCar(car, *rc.rules, rc.team, slot, cfg.spawn_z, m.boost_spawn_amount, offset_dir);
            } else if (slot >= kKickoffSpawnOverridePlus1 - 1) {
                KickoffSpawnCar(car, *rc.rules, rc.team, slot - (kKickoffSpawnOverridePlus1 - 1), cfg.spawn_z,
                                m.boost_spawn_amount);
            } else {
                RespawnCar(car, *rc.rules, rc.team, slot, cfg.spawn_z, m.boost_spawn_amount, -1);
            }
            respawned = true;
        } else if (!(Max(0.0f, car.demo_respawn_timer - dt) > 0.0f) &&
                   !(car.respawn_slot_override_plus1 >= kKickoffSpawnOverridePlus1 &&
                     car.respawn_slot_override_plus1 < kKickoffSpawnOverridePlus1 + 5)) {
            // The car respawns next tick (not as a kickoff re-creation): it shows the spawn boost already.
            car.boost = m.boost_spawn_amount;
            car.boost_amount = car.boost / 100.0f;
        }
    }
    const Quat spawn_rot = car.rot;
    car.tick_start_pos_bt = car.pos_bt;
    // 2. caps and quantization.
    TickStartLimitAndQuantize(car, m);
    // 3. finish.
    if (respawned) SpawnInertiaBasis(spawn_rot, car.tick_start_basis_forward, car.tick_start_basis_right,
                                     car.tick_start_basis_up);
    const Quat start_rot = car.rot;
    const Vec3 start_vel_uu = car.vel;
    if (start_vel_out) *start_vel_out = start_vel_uu;
    FlipZDampPrepass(car, m, dt);
    car.cal_tick_start_ang_vel = car.ang_vel;
    const Vec3 vel_before = car.vel_bt;
    PreTickCarPrePhysics(car, controls, cfg, m, m.frame_frac, dt, ws, start_rot, start_vel_uu);
    if (!car.is_demoed) {
        const bool impulsed = car.vel_bt.x != vel_before.x || car.vel_bt.y != vel_before.y || car.vel_bt.z != vel_before.z;
        if ((BodyDampingClampCar(car, m) & 1) && impulsed) SetCarVelBt(car, car.vel_bt);
    }
}

// ------------------------------------------------------------------------------------------------------------
// The car post-tick (StepEnvPostCar after the world step, env_step.hpp:6902-6926, with the world-side finish done):
// cooldowns and last controls (PostTickCar, raycast_wheels_and_contact_prep.hpp:4029-4038), the tick's bookkeeping
// (FinishCarTick, :3961-3991) and the fields the game reports (ReportRLBotCarFields / AdvanceRLBotCarFields,
// env_step.hpp:5835-5913).

RG_HD bool HasFlipOrJump(const CarState& car) {
    return car.is_on_ground || (!car.has_flipped && !car.has_double_jumped && car.air_time_since_jump < 1.25f);
}

RG_HD void FinishCarTick(CarState& car) {
    car.restored_inv_inertia_valid = 0;
    if (HasFlipOrJump(car) && !car.flip_or_jump_last_tick) car.last_flip_or_jump_start = car.tick_count_since_update;
    car.flip_or_jump_last_tick = HasFlipOrJump(car);
    car.ball_contact_last_tick = car.ball_contact_this_tick;
    if (car.is_demoed) return;
    if (car.pending_bump_count > 0) {
        Vec3 v = car.vel_bt;
        for (int i = 0; i < car.pending_bump_count; ++i) v = v + car.pending_bump_bt[i];
        SetCarVelBt(car, v);
        car.pending_bump_count = 0;
        car.vel_impulse_cache = Vec3(0.0f, 0.0f, 0.0f);
        car.vel_impulse_cache_bt = Vec3(0.0f, 0.0f, 0.0f);
    } else if (LengthSq(car.vel_impulse_cache_bt) > 0.0f) {
        SetCarVelBt(car, car.vel_bt + car.vel_impulse_cache_bt);
        car.vel_impulse_cache = Vec3(0.0f, 0.0f, 0.0f);
        car.vel_impulse_cache_bt = Vec3(0.0f, 0.0f, 0.0f);
    } else if (LengthSq(car.vel_impulse_cache) > 0.0f) {
        // SetCarLinearVelocityUu: through bt and back, marking the velocity touched.
        const Vec3 v = car.vel + car.vel_impulse_cache;
        car.vel = v;
        car.vel_bt = v * (1.0f / 50.0f);
        car.vel_touched_solve_stage = 1;
        car.vel_impulse_cache = Vec3(0.0f, 0.0f, 0.0f);
    }
    car.tick_count_since_update += 1;
}

// The double-jump tick counter and the game's has_jumped, shared by both report forms.
RG_HD void AdvanceJumpReportState(CarState& car, const CarParams& m, float dt, bool dj_started) {
    car.double_jump_ticks = !car.has_double_jumped
                                ? 255
                                : (dj_started ? 0 : (car.double_jump_ticks < 255 ? car.double_jump_ticks + 1 : 255));
    const bool on_ground_state = car.is_on_ground && !car.is_jumping && !car.is_flipping;
    if (car.is_jumping) {
        car.game_has_jumped = 1;
    } else if (on_ground_state &&
               (!car.has_jumped || car.internal_jump_time - dt > m.jump_min_time + m.jump_reset_time_pad)) {
        car.game_has_jumped = 0;
    }
}

// RLBot's air state of a car: 0 on the ground, 1 jumping, 2 just double-jumped (within 13 ticks), 3 flipping, 4 in
// the air otherwise.
RG_HD uint8_t RlbotAirState(const CarState& car) {
    if (car.is_jumping) return 1;
    if (car.is_flipping) return 3;
    if (car.is_on_ground) return 0;
    return car.double_jump_ticks < 13 ? 2 : 4;
}

// The game's jump flags at the end of a tick: from its own bookkeeping when the wheel runtime state is there (a
// frame boundary reached on three wheels clears them), else from the working flags.
struct GameJumpFlags {
    bool jumped, air_jumped;  // air_jumped: double jump or flip
};

RG_HD GameJumpFlags GameJumpFlagsOf(const CarState& car, const WheelRuntimeState* w) {
    GameJumpFlags g;
    if (w == nullptr) {
        g.jumped = car.has_jumped != 0;
        g.air_jumped = car.has_double_jumped || car.has_flipped;
        return g;
    }
    g.jumped = w->cal_game_jumped != 0;
    g.air_jumped = w->cal_game_double_jumped != 0;
    int wheels = 0;
    for (int k = 0; k < 4; ++k) wheels += car.wheels_with_contact[k] != 0 ? 1 : 0;
    if (w->cal_frame_open == 0 && w->cal_cached_ground == 0 && wheels >= 3) g.jumped = g.air_jumped = false;
    return g;
}

// The RLBot v5 dodge fields: whether the car has dodged (kept until it lands), the time left to dodge after a jump
// (-1 when it cannot), the time since the dodge began, and the dodge direction from the flip's relative torque.
RG_HD void DeriveDodgeFields(CarState& car, const WheelRuntimeState* w, const CarParams& m, float dt) {
    const bool flipped = car.internal_has_flipped != 0;
    const bool dodged = car.air_state != 0 && (car.rlbotv5_dodged || flipped);
    car.rlbotv5_dodged = dodged ? 1 : 0;
    const float since_jump = w != nullptr ? w->cal_jump_inactive_time : car.air_time_since_jump;
    const bool can_dodge = !car.is_on_ground && !car.is_jumping && car.game_has_jumped &&
                           !car.internal_has_double_jumped && !dodged &&
                           since_jump < m.doublejump_max_delay - 0.5f * dt;
    car.dodge_timeout = can_dodge ? m.doublejump_max_delay - since_jump : -1.0f;
    const bool settled = car.is_on_ground && !car.is_jumping && !car.is_flipping;
    if (flipped && !settled) {
        car.dodge_elapsed = car.internal_flip_time;
    } else {
        car.dodge_elapsed = dodged ? car.dodge_elapsed + dt : 0.0f;
    }
    car.dodge_dir_x = car.flip_rel_torque.y;
    car.dodge_dir_y = -car.flip_rel_torque.x;
}

// The state the game reports at the end of a step call (the reference's ReportRLBotCarFields, env_step.hpp:
// 5857-5910): the working velocities and jump flags are kept in the internal_* fields for the next tick, the
// public jump flags become the game's, the RLBot fields are derived, and reported_* records what was reported so
// the next tick can tell a caller's edit from the report.
RG_HD void ReportCarFields(CarState& car, const WheelRuntimeState* jw, const CarParams& m, float dt) {
    if (car.is_demoed && car.tick_count_since_update == 0) {
        car.reported_flags = 0;
        return;
    }
    const bool double_jump_began = car.has_double_jumped && !car.internal_has_double_jumped;
    car.internal_vel = car.vel;
    car.internal_ang_vel = car.ang_vel;
    car.internal_has_jumped = car.has_jumped;
    car.internal_has_double_jumped = car.has_double_jumped;
    car.internal_has_flipped = car.has_flipped;
    car.internal_is_supersonic = car.is_supersonic;
    AdvanceJumpReportState(car, m, dt, double_jump_began);

    const GameJumpFlags g = GameJumpFlagsOf(car, jw);
    car.has_jumped = g.jumped ? 1 : 0;
    car.has_double_jumped = g.air_jumped && car.internal_has_double_jumped ? 1 : 0;
    car.has_flipped = g.air_jumped && !car.has_double_jumped ? 1 : 0;
    car.air_state = RlbotAirState(car);
    DeriveDodgeFields(car, jw, m, dt);

    constexpr uint8_t kValid = 128;
    uint8_t flags = kValid;
    if (car.has_jumped) flags |= 4;
    if (car.has_double_jumped) flags |= 8;
    if (car.has_flipped) flags |= 16;
    if (car.is_supersonic) flags |= 32;
    car.reported_vel = car.vel;
    car.reported_ang_vel = car.ang_vel;
    car.reported_flags = flags;
}

// The in-between ticks of a multi-tick step report nothing; the counters still advance.
RG_HD void AdvanceCarFields(CarState& car, const WheelRuntimeState* jw, const CarParams& m, float dt) {
    if (!(LengthSq(car.vel_bt) < 1.0e38f && LengthSq(car.ang_vel) < 1.0e38f)) {
        ReportCarFields(car, jw, m, dt);
        return;
    }
    car.reported_flags = 0;
    if (car.is_demoed && car.tick_count_since_update == 0) return;
    const bool dj_started = car.has_double_jumped && !car.internal_has_double_jumped;
    car.internal_has_double_jumped = car.has_double_jumped;
    AdvanceJumpReportState(car, m, dt, dj_started);
}

RG_HD void CarPostTick(CarState& car, const CarControls& controls, const CarParams& m, const WheelRuntimeState* jw,
                       bool report_fields) {
    const float dt = m.tick_dt;
    if (!car.is_demoed) {
        if (car.car_contact.cooldown_timer > 0.0f) {
            car.car_contact.cooldown_timer = Max(0.0f, car.car_contact.cooldown_timer - dt);
        }
        car.last_controls = controls;
    }
    FinishCarTick(car);
    if (report_fields) {
        ReportCarFields(car, jw, m, dt);
    } else {
        AdvanceCarFields(car, jw, m, dt);
    }
}

// ------------------------------------------------------------------------------------------------------------
// Force phases of the car's controls step: the jump hold (UpdateJumpState, raycast_wheels_and_contact_prep.hpp:
// 3399-3431, :3629-3652) and boost (ApplyBoost, :3242-3300).  With the wheel runtime state present, as in every batch
// step, a phase leaves the velocity alone and queues its force on the car's frame for the solver
// (host_validation_and_contact_debug.hpp:184-212, :726-745); the caller passes that queue.  Without it (a null queue)
// the force changes the linear velocity at once (ApplyCentralForce, :136-138).

// The frame's force queue: the sums, the number of forces and torques queued with the first four of each in order,
// and the jump force, which a jump replaces rather than adds.
struct CarForceQueue {
    Vec3 force;
    Vec3 ang_accel;
    Vec3 torque_bt;
    Vec3 jump_force;
    int32_t has_jump_force;
    int32_t n_force;
    int32_t n_torque;
    Vec3 force_items[4];
    Vec3 torque_items[4];
};
static_assert(sizeof(CarForceQueue) == 12 * sizeof(Vec3) + 12, "force queue records are packed");

RG_HD void QueueCarForce(CarForceQueue& q, const Vec3& force_bt) {
    q.force = q.force + force_bt;
    if (q.n_force < 4) q.force_items[q.n_force] = force_bt;
    ++q.n_force;
}

RG_HD void QueueCarTorque(CarForceQueue& q, const Vec3& torque_bt) {
    q.torque_bt = q.torque_bt + torque_bt;
    if (q.n_torque < 4) q.torque_items[q.n_torque] = torque_bt;
    ++q.n_torque;
}

// v += F dt / m in game units, with the Bullet-unit copy kept in step (arena_geometry.hpp:1816-1821).
RG_HD void ApplyCarCentralForce(CarState& car, float mass, const Vec3& force_uu, float dt) {
    const Vec3 v = car.vel + force_uu * (dt / mass);
    car.vel = v;
    car.vel_bt = v * (1.0f / 50.0f);
    car.vel_touched_solve_stage = 1;
}

// The car's up axis scaled by scale_uu, in Bullet units (host_validation_and_contact_debug.hpp:505-509).
RG_HD Vec3 GameJumpUpBt(const CarState& car, float scale_uu) {
    Vec3 f, r, u;
    UeQuatAxes(car.rot, f, r, u);
    return Vec3((scale_uu * u.x) * 0.02f, (scale_uu * u.y) * 0.02f, (scale_uu * u.z) * 0.02f);
}

// Jump hold: a grounded car that is not jumping clears its jump once the reset pad has passed; a jump ends when the
// minimum time is over and the button is released (or its hold broken) or the maximum time is reached; a jump that
// goes on pushes along the up axis and counts its time.
RG_HD void CarUpdateJumpState(CarState& car, const CarControls& controls, const CarParams& m, CarForceQueue* q) {
    const float dt = m.tick_dt;
    if (car.is_on_ground && !car.is_jumping &&
        !(car.has_jumped && car.internal_jump_time < m.jump_min_time + m.jump_reset_time_pad)) {
        car.has_jumped = 0;
        car.internal_jump_time = 0.0f;
    }
    const bool held = car.internal_jump_time < m.jump_min_time ||
                      (controls.jump && !car.jump_hold_broken && car.internal_jump_time < m.jump_max_time);
    if (car.is_jumping && !held) car.is_jumping = 0;
    if (car.is_jumping) {
        car.has_jumped = 1;
        const Vec3 f = GameJumpUpBt(car, m.jump_accel * m.car_mass);
        if (q) {
            q->jump_force = f;
            q->has_jump_force = 1;
        } else {
            ApplyCarCentralForce(car, m.car_mass, f * 50.0f, dt);
        }
    }
    if (car.is_jumping || car.has_jumped) car.internal_jump_time += dt;
}

// The boost fraction kept beside the 0..100 amount: trusted while the two agree (raycast_wheels_and_contact_prep.hpp:
// 3242-3250).
RG_HD float CarBoostFraction(const CarState& car) {
    return car.boost_amount * 100.0f == car.boost ? car.boost_amount : car.boost / 100.0f;
}
RG_HD void SetCarBoostFraction(CarState& car, float fraction) {
    car.boost_amount = fraction;
    car.boost = fraction * 100.0f;
}

// Boost: a boosting car spends boost, and keeps boosting while it has boost and either holds the button or has not
// boosted for the minimum time; a car that pushes accelerates along its forward axis.  Otherwise the boost timers
// reset and an enabled recharge refills after its delay.  The amount ends clamped to 0..boost_max.  A demoed car
// (which the tick itself skips) only updates its latch and timers.
RG_HD void CarApplyBoost(CarState& car, const CarControls& controls, const CarParams& m, CarForceQueue* q) {
    const float dt = m.tick_dt;
    car.rlbotv5_boost_hold = 0;
    if (car.is_demoed) {
        if (car.boost > 0.0f) {
            if (car.is_boosting) {
                car.is_boosting = (controls.boost || car.boosting_time < m.boost_min_time) ? 1 : 0;
            } else if (controls.boost) {
                car.is_boosting = 1;
            }
        } else {
            car.is_boosting = 0;
        }
        car.boosting_time = 0.0f;
        car.time_since_boosted += dt;
        if (m.recharge_boost_enabled && car.time_since_boosted >= m.recharge_boost_delay) {
            car.boost = Min(m.boost_max, car.boost + m.recharge_boost_per_second * dt);
        }
        car.boost = Min(car.boost, m.boost_max);
        return;
    }
    bool push;
    if (car.is_boosting) {
        SetCarBoostFraction(car, Max(0.0f, CarBoostFraction(car) - (m.boost_used_per_second / 100.0f) * dt));
        push = !(car.boost == 0.0f || (!controls.boost && car.boosting_time >= m.boost_min_time));
    } else {
        push = controls.boost && car.boost > 0.0f;
    }
    car.is_boosting = push ? 1 : 0;
    if (push) {
        const Vec3 dir = UeQuatRotate(car.rot, Vec3(1.0f, 0.0f, 0.0f));
        const Vec3 f = ((m.boost_accel_ground * m.car_mass) * dir) * (1.0f / 50.0f);
        if (q) {
            QueueCarForce(*q, f);
        } else {
            ApplyCarCentralForce(car, m.car_mass, f * 50.0f, dt);
        }
        car.boosting_time += dt;
        car.time_since_boosted = 0.0f;
    } else {
        car.boosting_time = 0.0f;
        car.time_since_boosted += dt;
        if (m.recharge_boost_enabled && car.time_since_boosted >= m.recharge_boost_delay) {
            SetCarBoostFraction(car, CarBoostFraction(car) + (m.recharge_boost_per_second / 100.0f) * dt);
        }
    }
    if (car.boost < 0.0f || car.boost > m.boost_max) {
        SetCarBoostFraction(car, Min(Max(car.boost, 0.0f), m.boost_max) / 100.0f);
    }
}

// Rotator units per radian-free unit of the game's torque tables: 2 pi / 65536 (one rotator unit), and the same
// scaled by 1000 (the game's torque-to-acceleration factor).
constexpr float kRotatorUnit = 2.0f * Pi() / 65536.0f;
constexpr float kRotatorTorque = (2.0f * Pi() / 65536.0f) * 1000.0f;

// Whether pitch input is locked out: while a flip runs, for flip_pitchlock_extra_time after a flip's torque time,
// and while a restarted flip clock (cal_pre_pitchlock_restart_time) is inside that extra time.
RG_HD bool PitchLocked(const CarState& car, const CarParams& m) {
    if (car.is_flipping) return true;
    if (car.has_flipped && car.internal_flip_time < m.flip_torque_time + m.flip_pitchlock_extra_time) return true;
    return car.cal_pre_pitchlock_restart_time >= 0.0f && car.cal_pre_pitchlock_restart_time < m.flip_pitchlock_extra_time;
}

// A torque written as rates about the game's right, up and forward axes, summed per world component in the game's
// matrix order (right term, a zero term, up term, forward term): the zero term turns a negative zero into +0 as the
// game's product does.
RG_HD Vec3 GameAxesTorque(const Vec3& r, const Vec3& u, const Vec3& f, float about_r, float about_u, float about_f) {
    const auto one = [&](float rc, float uc, float fc) { return ((about_r * rc + 0.0f) + about_u * uc) + about_f * fc; };
    return Vec3(one(r.x, u.x, f.x), one(r.y, u.y, f.y), one(r.z, u.z, f.z));
}

// The damping counterpart: the three rates subtracted from zero in the same axis order.
RG_HD Vec3 GameAxesDamping(const Vec3& r, const Vec3& u, const Vec3& f, float about_r, float about_u, float about_f) {
    const auto one = [&](float rc, float uc, float fc) { return ((0.0f - about_r * rc) - about_u * uc) - about_f * fc; };
    return Vec3(one(r.x, u.x, f.x), one(r.y, u.y, f.y), one(r.z, u.z, f.z));
}

// A game-space torque through the inverse of the frame's Bullet inertia rows, queued when any component reaches
// the game's 1e-4 threshold.
RG_HD bool TorqueAboveGate(const Vec3& t) {
    constexpr float kGate = 1.0e-4f;
    return Abs(t.x) >= kGate || Abs(t.y) >= kGate || Abs(t.z) >= kGate;
}

RG_HD Vec3 ThroughInverseFrame(const CarFrame& frame, const Vec3& t) {
    Vec3 w0, w1, w2;
    InverseRows(frame.inv_row0, frame.inv_row1, frame.inv_row2, w0, w1, w2);
    return MulRows(w0, w1, w2, t);
}

// An angular acceleration in game units as the queued Bullet torque (through the game-unit inertia).
RG_HD Vec3 GameAccelAsTorque(const CarState& car, const CarConfig& cfg, const CarParams& m, const Vec3& accel) {
    Vec3 i0, i1, i2, w0, w1, w2;
    InvInertiaWorldRowsUu(car, cfg, m.car_mass, i0, i1, i2);
    InverseRows(i0, i1, i2, w0, w1, w2);
    return MulRows(w0, w1, w2, accel);
}

// Air control, part one: the control torque about the frame's pitch (-right), yaw (up) and roll (-forward) axes
// minus the damping of the tick-start angular velocity, as a game-unit angular acceleration.
RG_HD Vec3 FrameAirAccel(const CarControls& c, const CarFrame& frame, const Vec3& w, float pitch_gain,
                         const CarParams& m) {
    const Vec3 pitch_axis = -frame.right, yaw_axis = frame.up, roll_axis = -frame.fwd;
    Vec3 input(0.0f, 0.0f, 0.0f);
    if (c.pitch != 0.0f || c.yaw != 0.0f || c.roll != 0.0f) {
        const Vec3 p = c.pitch * pitch_axis * pitch_gain * m.car_air_control_pitch;
        const Vec3 y = c.yaw * yaw_axis * m.car_air_control_yaw;
        const Vec3 r = c.roll * roll_axis * m.car_air_control_roll;
        input = (p + y) + r;
    }
    const float held_pitch = 1.0f - Min(Abs(c.pitch * pitch_gain), 1.0f);
    const float held_yaw = 1.0f - Min(Abs(c.yaw), 1.0f);
    const float damp_pitch = Dot(pitch_axis, w) * m.car_air_damp_pitch * held_pitch;
    const float damp_yaw = Dot(yaw_axis, w) * m.car_air_damp_yaw * held_yaw;
    const float damp_roll = Dot(roll_axis, w) * m.car_air_damp_roll;
    const Vec3 damping = (yaw_axis * damp_yaw + pitch_axis * damp_pitch) + roll_axis * damp_roll;
    return (input - damping) * kRotatorTorque;
}

// Air control, part two (queued steps only): the same input and damping about the game's quaternion axes in
// rotator units, each queued as its own torque through the inverse frame when it passes the gate.
RG_HD void QueueGameAirTorques(const CarState& car, const CarControls& c, const CarFrame& frame, const Vec3& w,
                               float pitch_gain, const CarParams& m, CarForceQueue& q) {
    Vec3 f, r, u;
    UeQuatAxes(car.rot, f, r, u);
    const float pitch = c.pitch * pitch_gain;
    const float rate_r = (m.car_air_control_pitch * -pitch) * kRotatorUnit;
    const float rate_u = (m.car_air_control_yaw * c.yaw) * kRotatorUnit;
    const float rate_f = (-m.car_air_control_roll * c.roll) * kRotatorUnit;
    const Vec3 input = GameAxesTorque(r, u, f, rate_r, rate_u, rate_f) * 1000.0f;
    const float damp_r = (1.0f - Abs(pitch)) * (Dot(w, r) * (m.car_air_damp_pitch * kRotatorUnit));
    const float damp_u = (1.0f - Abs(c.yaw)) * (Dot(w, u) * (m.car_air_damp_yaw * kRotatorUnit));
    const float damp_f = (1.0f * Dot(w, f)) * (m.car_air_damp_roll * kRotatorUnit);
    const Vec3 damping = GameAxesDamping(r, u, f, damp_r, damp_u, damp_f) * 1000.0f;
    const bool queue_input = TorqueAboveGate(input), queue_damping = TorqueAboveGate(damping);
    if (queue_input) QueueCarTorque(q, ThroughInverseFrame(frame, input));
    if (queue_damping) QueueCarTorque(q, ThroughInverseFrame(frame, damping));
}

// Air control (the reference's ApplyAirControl, raycast_wheels_and_contact_prep.hpp:3048-3241): with
// `update_air_control`, the rotation control (part one: applied to the angular velocity when not queued, queued as
// an acceleration when the basis is not authoritative; part two when queued); with `air_throttle`, the throttle
// (full while boosting) as a force along the game's forward axis.
RG_HD void CarApplyAirControl(CarState& car, const CarControls& c, const CarConfig& cfg, const CarParams& m,
                              const CarFrame& frame, bool update_air_control, bool air_throttle, CarForceQueue* q) {
    if (update_air_control) {
        const float pitch_gain = PitchLocked(car, m) ? 0.0f : 1.0f;
        const Vec3 w = car.cal_tick_start_ang_vel;
        if (q == nullptr) {
            car.ang_vel += FrameAirAccel(c, frame, w, pitch_gain, m) * m.tick_dt;
        } else {
            if (!car.basis_authoritative) q->ang_accel = q->ang_accel + GameAccelAsTorque(car, cfg, m, FrameAirAccel(c, frame, w, pitch_gain, m));
            QueueGameAirTorques(car, c, frame, w, pitch_gain, m, *q);
        }
    }
    const float throttle = car.is_boosting ? 1.0f : c.throttle;
    if (!air_throttle || throttle == 0.0f) return;
    Vec3 f, r, u;
    UeQuatAxes(car.rot, f, r, u);
    const Vec3 force_bt = ((f * (m.car_air_throttle_accel * m.car_mass)) * throttle) * (1.0f / 50.0f);
    if (q != nullptr) {
        QueueCarForce(*q, force_bt);
    } else {
        ApplyCarCentralForce(car, m.car_mass, force_bt * 50.0f, m.tick_dt);
    }
}

// Hamilton product in the reference's term order (math.hpp:763-769).
RG_HD Quat QuatProduct(const Quat& a, const Quat& b) {
    return Quat(a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z, a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y,
                a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x, a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w);
}

// v rotated as the vector part of q (0, v) q^-1, with q^-1 the conjugate over |q|^2 (identity below 1e-12)
// (math.hpp:775-786, :866-870).
RG_HD Vec3 QuatSandwichRotate(const Quat& q, const Vec3& v) {
    const float len_sq = q.w * q.w + q.x * q.x + q.y * q.y + q.z * q.z;
    Quat inv(1.0f, 0.0f, 0.0f, 0.0f);
    if (!(len_sq <= 1e-12f)) {
        const float s = 1.0f / len_sq;
        inv = Quat(q.w * s, -q.x * s, -q.y * s, -q.z * s);
    }
    const Quat r = QuatProduct(QuatProduct(q, Quat(0.0f, v.x, v.y, v.z)), inv);
    return Vec3(r.x, r.y, r.z);
}

// A car-local vector in the world: through the basis columns when they are authoritative, else the quaternion
// (math.hpp:984-986, :1402-1405).
RG_HD Vec3 RotateCarVector(const CarState& car, const Vec3& v) {
    if (!car.basis_authoritative) return QuatSandwichRotate(car.rot, v);
    return (car.basis_forward * v.x + car.basis_right * v.y) + car.basis_up * v.z;
}

RG_HD float SignOf(float v) { return static_cast<float>((v > 0.0f) - (v < 0.0f)); }

// The pitch that cancels a dodge: the raw pitch when the caller has one (at most 1), else the control's, once the
// flip has run 0.041 s, quantized toward the game's axis steps (raycast_wheels_and_contact_prep.hpp:3032-3046).
RG_HD float DodgeCancelPitch(const CarState& car, const CarControls& c, float raw_pitch) {
    const float pitch = raw_pitch <= 1.0f ? raw_pitch : c.pitch;
    const float p = car.internal_flip_time >= 0.041f ? pitch : 0.0f;
    return -QuantizeAxisInput(-p);
}

// The flip's local torque: the relative torque scaled by the flip torque constants and rounded to whole units, its
// pitch part reduced while a pitch input against it cancels the dodge.
RG_HD Vec3 FlipLocalTorque(const CarState& car, const CarControls& c, const CarParams& m, float raw_pitch) {
    const Vec3 rel = car.flip_rel_torque;
    float pitch = rintf(rel.y * m.flip_torque_y);
    const float cancel = DodgeCancelPitch(car, c, raw_pitch);
    const bool cancels = rel.y != 0.0f && cancel != 0.0f && SignOf(rel.y) == SignOf(cancel);
    if (cancels) pitch *= 1.0f - Min(Abs(cancel), 1.0f);
    return Vec3(rintf(rel.x * m.flip_torque_x), pitch, 0.0f);
}

// Flip torque (the reference's ApplyFlipTorque, raycast_wheels_and_contact_prep.hpp:3432-3484).  The flip ends once
// its torque time is over.  While it runs, its local torque turns the car: straight into the angular velocity when
// not queued; queued, as an acceleration (basis not authoritative) and as the game-space torque about the
// quaternion axes through the inverse frame when it passes the gate.  The vehicle update queues that torque after
// air control: `deferred` receives it then; without `deferred` a nonzero torque is queued at once.  raw_pitch: the
// control's pitch before quantization, or 2 for none.
RG_HD void CarApplyFlipTorque(CarState& car, const CarControls& c, const CarConfig& cfg, const CarParams& m,
                              const CarFrame& frame, float raw_pitch, CarForceQueue* q, Vec3* deferred = nullptr) {
    if (car.is_flipping) car.is_flipping = car.has_flipped && car.internal_flip_time < m.flip_torque_time ? 1 : 0;
    if (!car.is_flipping) return;
    const Vec3 local = FlipLocalTorque(car, c, m, raw_pitch);
    const Vec3 world = RotateCarVector(car, local);
    if (q == nullptr) {
        car.ang_vel += world * m.tick_dt;
        return;
    }
    if (!car.basis_authoritative) q->ang_accel = q->ang_accel + GameAccelAsTorque(car, cfg, m, world);
    Vec3 f, r, u;
    UeQuatAxes(car.rot, f, r, u);
    // Game order: the pitch (right) term, then roll (forward), then the zero yaw (up) term, then a +0 that clears a
    // negative zero.
    const auto one = [&](float rc, float fc, float uc) { return ((local.y * rc + local.x * fc) + local.z * uc) + 0.0f * 0.0f; };
    const Vec3 game(one(r.x, f.x, u.x), one(r.y, f.y, u.y), one(r.z, f.z, u.z));
    if (!TorqueAboveGate(game)) return;
    const Vec3 t = ThroughInverseFrame(frame, game);
    if (deferred != nullptr) {
        *deferred = t;
    } else if (t.x != 0.0f || t.y != 0.0f || t.z != 0.0f) {
        QueueCarTorque(*q, t);
    }
}

}  // namespace core
}  // namespace rg
