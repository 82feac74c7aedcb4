// One arena tick for an arena without cars: the ball against the arena meshes.
//
// Stages (the reference's StepEnvGeneric with env_car_count 0, include/rocketgpu/detail/step/env_step.hpp:7048-7180):
//   1. env pre-tick (goal flag cleared)                                   env_step.hpp:6652-6699
//   2. ball pre-tick: speed caps, UE3 state quantization, damping         env_step.hpp:6875-6887
//   3. world step (Bullet stepSimulation with one substep)                env_step.hpp:4622-5623
//      a. activation (sleeping) at tick start
//      b. predictive (CCD) ray of the ball                                env_step.hpp:3893-4168
//      c. broadphase update and manifold list, island sort                arena_geometry.hpp:4162-5033
//      d. ball-vs-mesh contacts: persistent manifolds per mesh, aggregated velocity row, per-point split rows
//                                                                         arena_ball.hpp:1868-2437
//      e. predictive contact rows                                          env_step.hpp:4170-4259
//      f. sequential-impulse velocity solve (10 iterations) and split-impulse solve (10 iterations)
//                                                                         arena_geometry.hpp:5527-5866, 5035-5270
//      g. gravity, integration, deactivation
//   4. env tail: ball finish, goals, frame timers, tick counter          env_step.hpp:6968-7045
//
// Everything is evaluated in the reference's float operation order; see rg_math.hpp for the arithmetic contract.
#pragma once

#include <type_traits>

#include "rg_broadphase.hpp"

namespace rg {
namespace core {

// Per-batch constants of the step (the reference's MutatorConfig fields the ball path reads, plus arena rules).
struct StepParams {
    float tick_dt;
    float gravity_z;
    float ball_radius;
    float ball_mass;
    float ball_drag;
    float ball_damping_scale;  // pow(1 - drag, dt) in the UCRT pow, cached per batch like the reference
    float ball_friction;
    float ball_restitution;
    float ball_max_speed;
    float ball_max_ang_speed;
    float frame_frac;          // render-frame fraction (boost_pad_frame_frac)
    uint8_t enable_scoring;
    uint8_t pad[3];
    Vec3 goal_point[2], goal_normal[2], goal_min[2], goal_max[2];
    Vec3 ball_spawn;        // kickoff ball position (uu), from the arena's rules
    int32_t pad_count;      // boost pads per env
    uint64_t pad_big_mask;  // bit k: pad k is a big pad
};

// Where a tick's physics events go (PhysicsEventStream::Emit, rocketgpu.h:75-88): this env's window and its
// capacity records, or nowhere when the stream is off.  `writer` is false on lanes that mirror the lane writing an
// arena's records (the lanes of a group run the same tick on the same data).
struct EventSink {
    PhysicsEventWindow* window = nullptr;
    PhysicsEvent* records = nullptr;
    uint32_t capacity = 0;
    bool goal_only = false;
    bool writer = true;
    RG_HD bool on() const { return window != nullptr; }
    // Events other than goals are wanted (PhysicsEventStream::HasNonGoalEvents).
    RG_HD bool non_goal() const { return window != nullptr && !goal_only; }
    // The env pre-tick stamps the window with the env's tick (env_step.hpp:6695-6697).
    RG_HD void SetTick(uint32_t tick) const {
        if (window != nullptr && writer) window->tick = tick;
    }
    RG_HD void Emit(PhysicsEvent ev) const {
        if (goal_only && ev.type != PhysicsEventType::Goal) return;
        if (window == nullptr || !writer) return;
        ev.tick = window->tick;
        if (window->count >= capacity) {
            window->overflow = 1;
            return;
        }
        records[window->count++] = ev;
    }
    // A ball-vs-world contact: every mesh triangle the ball touches, with the triangle's own contact normal (before
    // the internal-edge correction), in the order the contact query finds them (arena_ball.hpp:1927-1936).
    RG_HD void BallWorldContact(const BallState& ball, const Vec3& normal) const {
        PhysicsEvent ev;
        ev.type = PhysicsEventType::BallWorldContact;
        ev.position = ball.pos;
        ev.normal = normal;
        ev.ball_position = ball.pos;
        ev.ball_velocity = ball.vel;
        Emit(ev);
    }
    // A goal on this tick (env_step.hpp:7025-7032).
    RG_HD void Goal(const BallState& ball, int32_t team) const {
        PhysicsEvent ev;
        ev.type = PhysicsEventType::Goal;
        ev.team = team;
        ev.ball_position = ball.pos;
        ev.ball_velocity = ball.vel;
        Emit(ev);
    }
};

struct BallManifoldPoint {
    Vec3 normal;
    float pen_uu;
    Vec3 local_a;  // contact point on the ball, in the ball's basis (bt)
    Vec3 local_b;  // contact point on the mesh (bt)
};

struct BallMeshManifold {
    uint8_t count;
    uint8_t pad[3];
    BallManifoldPoint pts[4];
};

// A car's persistent contact manifold against one mesh (Bullet's btPersistentManifold between the hitbox and the
// mesh shape): up to four points, each the car-body-frame point, the mesh point and the normal when it was found,
// and its distance at the last refresh (bt).
struct CarWorldPoint {
    Vec3 local_a;
    Vec3 point_b;
    Vec3 normal;
    float distance;
};

struct CarWorldManifold {
    uint8_t count;
    uint8_t mesh;
    uint8_t pad[2];
    CarWorldPoint pts[4];
};

constexpr int kCarWorldManifoldsPerCar = 2;
constexpr int kMaxArenaCarSlots = kMaxBodies - 1;

// The persistent contact manifold of two cars (rocketgpu.h CarCarContact): up to four points, each the point on
// car a and the point on car b in their body frames, and the normal from a toward b when it was found (bt).  The
// world points and distances are recomputed from these at every refresh, so only these are kept between ticks.
struct CarPairPoint {
    Vec3 local_a;
    Vec3 local_b;
    Vec3 normal;
};

struct CarPairManifold {
    uint8_t count;
    uint8_t pad[3];
    CarPairPoint pts[4];
};

// One manifold per pair of car slots lo < hi, at hi * (hi - 1) / 2 + lo.
constexpr int kCarPairSlots = kMaxArenaCarSlots * (kMaxArenaCarSlots - 1) / 2;
RG_HD constexpr int CarPairSlot(int lo, int hi) { return hi * (hi - 1) / 2 + lo; }

// Internal per-arena state that persists across ticks (the parts of the reference's WorldSolverScratch the ball
// path carries).
struct ArenaInternal {
    BroadphaseState bp;
    uint16_t ball_mesh_retained;  // bit m: mesh m's persistent ball manifold may hold points
    // Windows the last tick's contact walk loaded in a group kernel (saturating; 0 after a one-lane kernel).  The
    // launch uses it to give the arenas with the longest walks more lanes; it does not affect the physics.
    uint16_t walk_rounds;
    // Car arenas: the physics the tick met that this build does not step (kUnsupported* bits), sticky until reset.
    uint32_t unsupported;
    BallMeshManifold ball_mesh[kMaxArenaMeshes];
    // Car arenas: each car slot's hitbox-vs-mesh manifolds (two slots, each bound to one mesh while it holds points).
    CarWorldManifold car_world[kMaxArenaCarSlots][kCarWorldManifoldsPerCar];
    // Car arenas: each pair of car slots' hitbox-vs-hitbox manifold.
    CarPairManifold car_pair[kCarPairSlots];
};

// ------------------------------------------------------------------------------------------------------------
// Ball state helpers.

RG_HD void SetBallPosBt(BallState& b, const Vec3& p) {
    b.pos_bt = p;
    b.pos = p * 50.0f;
}
RG_HD void SetBallVelBt(BallState& b, const Vec3& v) {
    b.vel_bt = v;
    b.vel = v * 50.0f;
}
RG_HD Vec3 UuToBtDiv(const Vec3& v) { return Vec3(v.x / 50.0f, v.y / 50.0f, v.z / 50.0f); }

RG_HD void BallBasis(const BallState& b, Vec3& f, Vec3& r, Vec3& u) {
    if (b.basis_authoritative) {
        f = b.basis_forward;
        r = b.basis_right;
        u = b.basis_up;
        return;
    }
    QuatColumns(b.rot, f, r, u);
}

RG_HD void BallInertiaBasis(const BallState& b, Vec3& f, Vec3& r, Vec3& u) {
    if (b.inertia_basis_valid) {
        f = b.inertia_basis_forward;
        r = b.inertia_basis_right;
        u = b.inertia_basis_up;
        return;
    }
    BallBasis(b, f, r, u);
}

RG_HD Vec3 Mul3x3(const Vec3& r0, const Vec3& r1, const Vec3& r2, const Vec3& v) {
    return Vec3(Dot(r0, v), Dot(r1, v), Dot(r2, v));
}

// World inverse inertia of the ball (a solid sphere of radius r) applied to a torque, in bt: B diag(1/I) B^T t.
RG_HD Vec3 BallInvInertiaBt(const BallState& b, const StepParams& p, const Vec3& t) {
    if (b.restored_inv_inertia_valid) {
        return Mul3x3(b.restored_inv_inertia_bt[0], b.restored_inv_inertia_bt[1], b.restored_inv_inertia_bt[2], t);
    }
    Vec3 f(1.0f, 0.0f, 0.0f), r(0.0f, 1.0f, 0.0f), u(0.0f, 0.0f, 1.0f);
    BallInertiaBasis(b, f, r, u);
    const float radius_bt = p.ball_radius * (1.0f / 50.0f);
    const float inertia = 0.4f * p.ball_mass * radius_bt * radius_bt;
    const float inv = inertia != 0.0f ? (1.0f / inertia) : 0.0f;
    const float b00 = f.x, b01 = r.x, b02 = u.x;
    const float b10 = f.y, b11 = r.y, b12 = u.y;
    const float b20 = f.z, b21 = r.z, b22 = u.z;
    const float s00 = b00 * inv, s01 = b01 * inv, s02 = b02 * inv;
    const float s10 = b10 * inv, s11 = b11 * inv, s12 = b12 * inv;
    const float s20 = b20 * inv, s21 = b21 * inv, s22 = b22 * inv;
    const Vec3 row0(s00 * b00 + s01 * b01 + s02 * b02, s00 * b10 + s01 * b11 + s02 * b12,
                    s00 * b20 + s01 * b21 + s02 * b22);
    const Vec3 row1(s10 * b00 + s11 * b01 + s12 * b02, s10 * b10 + s11 * b11 + s12 * b12,
                    s10 * b20 + s11 * b21 + s12 * b22);
    const Vec3 row2(s20 * b00 + s21 * b01 + s22 * b02, s20 * b10 + s21 * b11 + s22 * b12,
                    s20 * b20 + s21 * b21 + s22 * b22);
    return Mul3x3(row0, row1, row2, t);
}

// Same in uu with the scalar sphere inertia (the reference's uu-side bookkeeping, used for row validity).
RG_HD Vec3 BallInvInertiaUu(const StepParams& p, const Vec3& t) {
    const float inertia = 0.4f * p.ball_mass * p.ball_radius * p.ball_radius;
    return inertia > 0.0f ? (t / inertia) : Vec3(0.0f, 0.0f, 0.0f);
}

// ------------------------------------------------------------------------------------------------------------
// Pre-tick: speed caps and the game's replicated-state quantization (raycast_wheels_and_contact_prep.hpp:3757-3898).

RG_HD void BallPreTick(BallState& ball, const StepParams& p, Vec3& predict_vel_bt) {
    // Speed caps.
    Vec3 vel_uu = ball.vel_bt * 50.0f;
    const float len_sq = LengthSq(vel_uu);
    if (len_sq > p.ball_max_speed * p.ball_max_speed) {
        vel_uu = (vel_uu * Rsqrt(len_sq)) * p.ball_max_speed;
        SetBallVelBt(ball, vel_uu * (1.0f / 50.0f));
    }
    ball.ang_vel = ClampLength(ball.ang_vel, p.ball_max_ang_speed);
    // Quantization of position (0.01 uu, round), velocity (0.01 uu/s, truncate+offset), angular velocity, rotation.
    const Vec3 pu = ball.pos_bt * 50.0f;
    const Vec3 qp(QuantRoundPos(pu.x, 100.0f), QuantRoundPos(pu.y, 100.0f), QuantRoundPos(pu.z, 100.0f));
    SetBallPosBt(ball, qp * (1.0f / 50.0f));
    const Vec3 qv(QuantTruncOffset(vel_uu.x, 100.0f), QuantTruncOffset(vel_uu.y, 100.0f),
                  QuantTruncOffset(vel_uu.z, 100.0f));
    SetBallVelBt(ball, qv * (1.0f / 50.0f));
    ball.pos = qp;
    ball.vel = qv;
    ball.ang_vel = Vec3(QuantTruncOffset(ball.ang_vel.x, 10000.0f), QuantTruncOffset(ball.ang_vel.y, 10000.0f),
                        QuantTruncOffset(ball.ang_vel.z, 10000.0f));
    if (!ball.basis_authoritative) {
        QuatColumns(ball.rot, ball.basis_forward, ball.basis_right, ball.basis_up);
        ball.basis_authoritative = 1;
    }
    ball.inertia_basis_forward = ball.basis_forward;
    ball.inertia_basis_right = ball.basis_right;
    ball.inertia_basis_up = ball.basis_up;
    ball.inertia_basis_valid = 1;
    ball.rot = GameQuantizeColumns(ball.basis_forward, ball.basis_right, ball.basis_up);
    // Body velocity clamp (Bullet-side), then linear damping with the cached scale.
    {
        const Vec3 v = ball.vel_bt;
        const float max_lin = p.ball_max_speed * 0.02f;
        const float vsq = (v.x * v.x + v.y * v.y) + v.z * v.z;
        if (0.0f < max_lin && max_lin * max_lin < vsq) {
            const float s = 1.0f / sqrtf(vsq);
            SetBallVelBt(ball, Vec3((v.x * s) * max_lin, (v.y * s) * max_lin, (v.z * s) * max_lin));
        }
        const Vec3 w = ball.ang_vel;
        const float max_ang = p.ball_max_ang_speed;
        const float lsq = (w.x * w.x + w.y * w.y) + w.z * w.z;
        if (0.0f < max_ang && max_ang * max_ang < lsq) {
            const float s = 1.0f / sqrtf(lsq);
            ball.ang_vel = Vec3((w.x * s) * max_ang, (w.y * s) * max_ang, (w.z * s) * max_ang);
        }
    }
    if (p.ball_drag > 0.0f) SetBallVelBt(ball, ball.vel_bt * p.ball_damping_scale);
    predict_vel_bt = ball.vel_bt;
}

// ------------------------------------------------------------------------------------------------------------
// Solver rows (ball only; stored in bt units, which is what the reference's solver runs in for these rows).

constexpr int kMaxBallSplitRows = 20;
constexpr int kMaxBallVelRows = 4;

struct SplitRow {
    Vec3 normal;
    Vec3 rel_cross_bt;
    Vec3 ang_bt;
    float inv_mass;
    float jac_bt;
    float rhs_bt;
    float applied;
    int16_t static_id;
};

struct VelRow {
    Vec3 normal;
    Vec3 rel_cross_bt;
    Vec3 ang_bt;
    float inv_mass;
    float jac;
    float rhs_bt;
    float applied;
    uint8_t predictive;
};

struct FricRow {
    Vec3 tangent;
    Vec3 rel_cross_bt;
    Vec3 ang_bt;
    float inv_mass;
    float jac;
    float rhs_bt;
    float applied;
    float friction;
    int32_t normal_row;
};

struct BallRows {
    SplitRow split[kMaxBallSplitRows];
    VelRow vel[kMaxBallVelRows];
    FricRow fric[kMaxBallVelRows];
    int n_split, n_vel, n_fric;
};

// Split (position-correction) row against a static body (InitializeBallSplitContactRow with the ball's bt terms).
RG_HD void AppendStaticSplitRow(BallRows& rows, float pen_uu, const Vec3& n, const Vec3& rel_uu, const StepParams& p,
                                float dt, const Vec3& rel_bt, float pen_bt_in, const Vec3& ang_bt, int16_t static_id) {
    if (rows.n_split >= kMaxBallSplitRows) return;
    if (pen_uu <= 0.0f || dt <= 0.0f) return;
    const float inv_mass = 1.0f / p.ball_mass;
    const Vec3 ang_uu = BallInvInertiaUu(p, Cross(rel_uu, n));
    float denom = 0.0f;
    if (inv_mass > 0.0f) denom += inv_mass + Dot(n, Cross(ang_uu, rel_uu));
    if (denom <= 1e-8f) return;
    const float jac = 1.0f / denom;
    SplitRow& r = rows.split[rows.n_split++];
    r.normal = n;
    r.rel_cross_bt = Cross(rel_bt, n);
    r.ang_bt = ang_bt;
    r.inv_mass = inv_mass;
    float denom_bt = 0.0f;
    if (inv_mass > 0.0f) denom_bt += inv_mass + Dot(n, Cross(ang_bt, rel_bt));
    r.jac_bt = denom_bt > 1e-8f ? (1.0f / denom_bt) : jac;
    const float pen_bt = pen_bt_in > 0.0f ? pen_bt_in : pen_uu * (1.0f / 50.0f);
    r.rhs_bt = ((pen_bt * 0.8f) / dt) * r.jac_bt;
    r.applied = 0.0f;
    r.static_id = static_id;
}

// Normal + friction rows of the ball against a static body (InitializeBallVelocityContactRow /
// InitializeBallFrictionContactRow with no other body; restitution without Bullet's velocity threshold).
RG_HD void AppendStaticVelocityRows(BallRows& rows, const Vec3& n, float restitution, float friction,
                                    const Vec3& rel_uu, const Vec3& ang_n_uu, const Vec3& ang_t_uu, const Vec3& ball_vel_uu,
                                    const Vec3& ball_ang, const Vec3& ext_lin_uu, const Vec3& rel_bt,
                                    const Vec3& bt_lin_override, const Vec3& ang_n_bt, const Vec3& ang_t_bt,
                                    const Vec3& tangent_basis_bt, const StepParams& p, bool predictive) {
    if (rows.n_vel >= kMaxBallVelRows) return;
    const float inv_mass = 1.0f / p.ball_mass;
    const Vec3 zero(0.0f, 0.0f, 0.0f);
    const Vec3 rel_other_uu = zero;
    const Vec3 bt_rel_other = UuToBtDiv(rel_other_uu);
    float denom = 0.0f;
    if (inv_mass > 0.0f) denom += inv_mass + Dot(n, Cross(ang_n_uu, rel_uu));
    if (denom <= 1e-8f) return;
    float jac = 1.0f / denom;
    {
        float bt_denom = 0.0f;
        if (inv_mass > 0.0f) bt_denom += inv_mass + Dot(n, Cross(ang_n_bt, rel_bt));
        if (bt_denom > 1e-8f) jac = 1.0f / bt_denom;
    }
    const Vec3 bt_rel_ball_cross_n = Cross(rel_bt, n);
    const Vec3 bt_rel_other_cross_n = Cross(bt_rel_other, n);
    const Vec3 bt_lin_b = bt_lin_override;
    const Vec3 bt_ang_b = ball_ang;
    const Vec3 bt_ext_b = UuToBtDiv(ext_lin_uu);
    const Vec3 bt_ext_torque_b = zero;
    const Vec3 bt_lin_a = UuToBtDiv(zero);
    const Vec3 bt_ang_a = zero;
    const Vec3 bt_ext_a = UuToBtDiv(zero);
    const Vec3 bt_ext_torque_a = zero;
    const Vec3 rhs_b = bt_lin_b + bt_ext_b;
    const Vec3 rhs_a = bt_lin_a + bt_ext_a;
    const float rhs_b_n = Dot(rhs_b, n) + Dot(bt_rel_ball_cross_n, bt_ang_b + bt_ext_torque_b);
    const float rhs_a_n = Dot(rhs_a, n) + Dot(bt_rel_other_cross_n, bt_ang_a + bt_ext_torque_a);
    const float rhs_rel_n = rhs_b_n - rhs_a_n;
    const Vec3 rest_b = bt_lin_b + Cross(bt_ang_b, rel_bt);
    const Vec3 rest_a = bt_lin_a + Cross(bt_ang_a, bt_rel_other);
    const float rest_rel_n = Dot(rest_b - rest_a, n);
    const float rest = Max(restitution * -rest_rel_n, 0.0f);
    const int vi = rows.n_vel++;
    VelRow& v = rows.vel[vi];
    v.normal = n;
    v.rel_cross_bt = bt_rel_ball_cross_n;
    v.ang_bt = ang_n_bt;
    v.inv_mass = inv_mass;
    v.jac = jac;
    v.rhs_bt = (rest - rhs_rel_n) * jac;
    v.applied = 0.0f;
    v.predictive = predictive ? 1 : 0;
    (void)ball_vel_uu;

    // Friction row along the tangential relative velocity.
    if (rows.n_fric >= kMaxBallVelRows || friction <= 0.0f) return;
    Vec3 tangent;
    {
        const float vn = Dot(n, tangent_basis_bt);
        const Vec3 lat = tangent_basis_bt - n * vn;
        const float lat2 = Dot(lat, lat);
        if (lat2 > 1.1920928955078125e-7f) {
            tangent = lat * (1.0f / sqrtf(lat2));
        } else {
            const Vec3 axis = Abs(n.z) < 0.70710678f ? Vec3(0.0f, 0.0f, 1.0f) : Vec3(1.0f, 0.0f, 0.0f);
            tangent = Normalize(Cross(axis, n), Vec3(1.0f, 0.0f, 0.0f));
        }
    }
    const Vec3 axis_n = -tangent;
    float d0 = 0.0f;
    float d1 = 0.0f;
    if (inv_mass > 0.0f) d1 = inv_mass + Dot(axis_n, Cross(-ang_t_uu, rel_uu));
    const float fd = d0 + d1;
    if (fd <= 1e-8f) return;
    float fjac = 1.0f / fd;
    {
        float bd0 = 0.0f;
        float bd1 = 0.0f;
        if (inv_mass > 0.0f) bd1 = inv_mass + Dot(axis_n, Cross(-ang_t_bt, rel_bt));
        const float bd = bd0 + bd1;
        if (bd > 1e-8f) fjac = 1.0f / bd;
    }
    FricRow& f = rows.fric[rows.n_fric++];
    f.tangent = tangent;
    f.rel_cross_bt = Cross(rel_bt, tangent);
    f.ang_bt = ang_t_bt;
    f.inv_mass = inv_mass;
    f.jac = fjac;
    const Vec3 cn1 = tangent;
    const Vec3 rel1_cross = Cross(rel_bt, cn1);
    const float vel1 = Dot(cn1, bt_lin_b + bt_ext_b) + Dot(rel1_cross, bt_ang_b);
    const float rel_vel = vel1 + 0.0f;
    f.rhs_bt = (-rel_vel) * fjac;
    f.applied = 0.0f;
    f.friction = friction;
    f.normal_row = vi;
}

// ------------------------------------------------------------------------------------------------------------
// Persistent ball-vs-mesh manifold (Bullet's btPersistentManifold: 4 points, deepest kept, largest area on
// replacement, merge within the breaking threshold; reference arena_ball.hpp:1146-1290).

struct BallPointFull {
    BallManifoldPoint p;
    Vec3 rel_pos1_bt;
    float rel_len_bt;
    float distance_bt;
};

RG_HD int SortCachedPoints(const BallManifoldPoint* c, const BallManifoldPoint& np) {
    int max_idx = -1;
    float max_pen = np.pen_uu;
    for (int i = 0; i < 4; ++i) {
        if (c[i].pen_uu > max_pen) {
            max_idx = i;
            max_pen = c[i].pen_uu;
        }
    }
    float r0 = 0.0f, r1 = 0.0f, r2 = 0.0f, r3 = 0.0f;
    if (max_idx != 0) {
        const Vec3 x = Cross(np.local_a - c[1].local_a, c[3].local_a - c[2].local_a);
        r0 = Dot(x, x);
    }
    if (max_idx != 1) {
        const Vec3 x = Cross(np.local_a - c[0].local_a, c[3].local_a - c[2].local_a);
        r1 = Dot(x, x);
    }
    if (max_idx != 2) {
        const Vec3 x = Cross(np.local_a - c[0].local_a, c[3].local_a - c[1].local_a);
        r2 = Dot(x, x);
    }
    if (max_idx != 3) {
        const Vec3 x = Cross(np.local_a - c[0].local_a, c[2].local_a - c[1].local_a);
        r3 = Dot(x, x);
    }
    int biggest = 0;
    float best = r0;
    if (r1 > best) {
        best = r1;
        biggest = 1;
    }
    if (r2 > best) {
        best = r2;
        biggest = 2;
    }
    if (r3 > best) {
        biggest = 3;
    }
    return biggest;
}

RG_HD void InsertManifoldPoint(BallManifoldPoint* c, int& count, const BallManifoldPoint& np, float merge_th) {
    if (merge_th >= 0.0f) {
        float shortest = merge_th * merge_th;
        int nearest = -1;
        for (int i = 0; i < count; ++i) {
            const Vec3 d = c[i].local_a - np.local_a;
            const float dd = Dot(d, d);
            if (dd < shortest) {
                shortest = dd;
                nearest = i;
            }
        }
        if (nearest >= 0) {
            c[nearest] = np;
            return;
        }
    }
    int at = count;
    if (at == 4) {
        at = SortCachedPoints(c, np);
    } else {
        ++count;
    }
    if (at < 0) at = 0;
    c[at] = np;
}

// Recompute each point's distance from the current ball pose and drop points beyond the breaking threshold
// (last point first; a dropped point is replaced by the current last one).  Fills the per-point derived terms into
// `out`.
RG_HD void RefreshManifold(BallManifoldPoint* c, int& count, const BallState& ball, const Vec3& center_bt,
                           float threshold, BallPointFull* out) {
    Vec3 bx, by, bz;
    BallBasis(ball, bx, by, bz);
    for (int i = count - 1; i >= 0; --i) {
        const BallManifoldPoint& mp = c[i];
        const Vec3 rot_a(bx.x * mp.local_a.x + by.x * mp.local_a.y + bz.x * mp.local_a.z,
                         bx.y * mp.local_a.x + by.y * mp.local_a.y + bz.y * mp.local_a.z,
                         bx.z * mp.local_a.x + by.z * mp.local_a.y + bz.z * mp.local_a.z);
        const Vec3 world_a = rot_a + center_bt;
        const Vec3 world_b = mp.local_b;
        const float dist = Dot(world_a - world_b, mp.normal);
        c[i].pen_uu = -dist * kBtToUu;
        out[i].distance_bt = dist;
        out[i].rel_pos1_bt = world_a - center_bt;
        out[i].rel_len_bt = Length(world_a - center_bt);
        bool remove = dist > threshold;
        if (!remove) {
            const Vec3 proj = world_a - mp.normal * dist;
            const Vec3 diff = world_b - proj;
            remove = Dot(diff, diff) > threshold * threshold;
        }
        if (remove) {
            c[i] = c[count - 1];
            out[i] = out[count - 1];
            --count;
        }
    }
    for (int i = 0; i < count; ++i) out[i].p = c[i];
}

// ------------------------------------------------------------------------------------------------------------
// The tick.

// With the ball alone the manifold list holds at most one manifold per static mesh plus the predictive one.
constexpr int kMaxBallManifoldEntries = kMaxArenaMeshes + 2;

// A retained ball-mesh contact point with what its rows need.
struct BallContactPoint {
    Vec3 normal;
    float pen_uu;
    float rel_len_bt;
    Vec3 rel_pos1_bt;
    float pen_bt;
    int16_t static_id;
};

constexpr int kMaxBallContactPoints = 16;

// Per-tick working data of one arena (no state survives the tick).  Kept in one place so a launch can put it in LDS.
struct BallTickScratch {
    BodyProxy ball_proxy;
    Vec3 predict_vel_bt;
    uint8_t predictive_hit;
    Vec3 predictive_normal;
    float predictive_distance_bt;
    ManifoldEntry entries[kMaxBallManifoldEntries];
    int entry_count;
    BallRows rows;
    BallContactPoint cps[kMaxBallContactPoints];  // the contact points of the tick's ball-mesh manifolds
    BallPointFull full[4];                        // one manifold's refreshed points
};

// Predictive contact ray of the ball (Bullet's createPredictiveContacts for the CCD body, as a ray from the center;
// env_step.hpp:3893-4127, ball branch).
template <typename Walk = SerialWalk>
RG_HD bool BallPredictiveRay(const BallState& ball, float dt, const ArenaMeshView& mesh, Vec3& n_out, float& d_out) {
    constexpr float kCcdMotionThresholdBt = 0.02f;
    constexpr float kRayBackoffBt = 0.04f;
    const Vec3 cur = ball.pos_bt;
    const Vec3 to = ball.pos_bt + ball.vel_bt * dt;
    const Vec3 m(to.x - cur.x, to.y - cur.y, to.z - cur.z);
    const float m2 = m.x * m.x + m.y * m.y + m.z * m.z;
    if (!(m2 > kCcdMotionThresholdBt * kCcdMotionThresholdBt)) return false;
    const float inv_len = 1.0f / sqrtf(m2);
    const Vec3 from(cur.x - inv_len * m.x * kRayBackoffBt, cur.y - inv_len * m.y * kRayBackoffBt,
                    cur.z - inv_len * m.z * kRayBackoffBt);
    const Vec3 delta(to.x - from.x, to.y - from.y, to.z - from.z);
    const RayHit hit = CastArenaRay<Walk>(mesh, from, delta);
    if (!hit.hit || !(hit.fraction < 1.0f)) return false;
    Vec3 n = hit.normal;
    if (hit.tri >= 0 && hit.tri < mesh.tri_count) {
        const MeshTriangle& tri = mesh.tris[hit.tri];
        n = Cross(tri.b - tri.a, tri.c - tri.a);
        const bool behind = Dot(n, from) - Dot(tri.a, n) <= 0.0f;
        const float il = 1.0f / sqrtf((n.y * n.y + n.x * n.x) + n.z * n.z);
        n = Vec3(n.x * il, n.y * il, n.z * il);
        if (behind) n = -n;
    }
    const float f = hit.fraction;
    const Vec3 d(delta.x * f, delta.y * f, delta.z * f);
    const float d_len = sqrtf(d.x * d.x + d.y * d.y + d.z * d.z);
    const float s = d_len > kRayBackoffBt ? d_len - kRayBackoffBt : 0.0f;
    const float inv_d = 1.0f / d_len;
    const Vec3 ds(inv_d * d.x * s, inv_d * d.y * s, inv_d * d.z * s);
    n_out = n;
    d_out = ((delta.y - ds.y) * n.y + (delta.x - ds.x) * n.x) + (delta.z - ds.z) * n.z;
    return true;
}

// Broadphase update for the ball alone and the sorted manifold list (BuildBulletManifoldList with no cars).  A
// sleeping ball gets no new ball-static manifolds.
RG_HD void BallBroadphase(ArenaInternal& st, const ArenaMeshView& mesh, BallTickScratch& t, bool static_pairs_skipped) {
    BroadphaseState& bp = st.bp;
    SweepState& h = bp.sweep;
    const int body_count = 1;
    if (h.handle_count == 0) {
        SweepInitStatics(bp, mesh);
        bp.pair_count = 0;
        bp.manifold_count = 0;
        bp.predictive_count = 0;
    }
    Sweep s(bp, body_count);
    // Insert the ball when it has no handle.
    if (t.ball_proxy.valid && h.body_handle[0] == 0 && h.handle_count < kSweepMaxHandles - 1) {
        uint16_t qmin[3], qmax[3];
        const float mn[3] = {t.ball_proxy.child_min.x, t.ball_proxy.child_min.y, t.ball_proxy.child_min.z};
        const float mx[3] = {t.ball_proxy.child_max.x, t.ball_proxy.child_max.y, t.ball_proxy.child_max.z};
        for (int a = 0; a < 3; ++a) {
            qmin[a] = SweepQuantize(mn[a], 0);
            qmax[a] = SweepQuantize(mx[a], 1);
        }
        h.body_handle[0] = h.free_count > 0 ? h.free_list[h.free_count - 1] : h.next_fresh;
        s.AddHandle(qmin, qmax, static_cast<int16_t>(-1), true);
        h.obj[h.obj_count++] = 0;
    }
    // Predictive manifolds are rebuilt every tick at the end of the list.
    for (int i = 0; i < bp.predictive_count; ++i) {
        SwapRemove(bp.manifolds, bp.manifold_count, FindKey(bp.manifolds, bp.manifold_count, bp.predictive[i]));
    }
    bp.predictive_count = 0;
    for (int obj = 0; obj < h.obj_count && t.predictive_hit; ++obj) {
        const int body = h.obj[obj];
        if (body != 0 || bp.manifold_count >= kMaxManifoldEntries) continue;
        const uint32_t key = PackPairKey(0, -1, -1, kPredictiveStaticId);
        bp.manifolds[bp.manifold_count++] = key;
        bp.predictive[bp.predictive_count++] = key;
    }
    // Move the swept boxes.
    for (int obj = 0; obj < h.obj_count; ++obj) {
        const int body = h.obj[obj];
        if (body >= body_count) continue;
        const int handle = h.body_handle[body];
        if (!t.ball_proxy.valid || handle == 0) continue;
        const float mn[3] = {t.ball_proxy.aabb_min.x, t.ball_proxy.aabb_min.y, t.ball_proxy.aabb_min.z};
        const float mx[3] = {t.ball_proxy.aabb_max.x, t.ball_proxy.aabb_max.y, t.ball_proxy.aabb_max.z};
        for (int axis = 0; axis < 3; ++axis) {
            const uint16_t qmin = SweepQuantize(mn[axis], 0);
            const uint16_t qmax = SweepQuantize(mx[axis], 1);
            const int emin = h.min_edge[handle][axis];
            const int emax = h.max_edge[handle][axis];
            const int dmin = static_cast<int>(qmin) - static_cast<int>(h.edge_pos[axis][emin]);
            const int dmax = static_cast<int>(qmax) - static_cast<int>(h.edge_pos[axis][emax]);
            if (dmin == 0 && dmax == 0) continue;
            if (dmin != 0) h.edge_pos[axis][emin] = qmin;
            if (dmax != 0) h.edge_pos[axis][emax] = qmax;
            if (dmin < 0) s.SortDown(axis, emin, false, true);
            if (dmax > 0) s.SortUp(axis, emax, true, true);
            if (dmin > 0) s.SortUp(axis, emin, false, true);
            if (dmax < 0) s.SortDown(axis, emax, true, true);
        }
    }
    // Pairs -> manifolds (ball-static pairs always want a manifold while the ball has a proxy).
    for (int i = 0; i < bp.pair_count; ++i) {
        const uint32_t key = bp.pairs[i];
        const int at = FindKey(bp.manifolds, bp.manifold_count, key);
        const bool want = KeyKind(key) == 0 ? t.ball_proxy.valid != 0 : false;
        if (want && at < 0 && static_pairs_skipped) continue;
        if (want && at < 0) {
            bp.manifolds[bp.manifold_count++] = key;
        } else if (!want) {
            SwapRemove(bp.manifolds, bp.manifold_count, at);
        }
    }
    // Islands: with the ball alone every manifold belongs to the ball's island (its object index).
    int8_t tag[kMaxBodies];
    for (int b = 0; b < kMaxBodies; ++b) tag[b] = -1;
    for (int obj = 0; obj < h.obj_count; ++obj) tag[h.obj[obj]] = static_cast<int8_t>(obj);
    const int n = bp.manifold_count < kMaxBallManifoldEntries ? bp.manifold_count : kMaxBallManifoldEntries;
    int8_t island[kMaxBallManifoldEntries];
    int8_t order[kMaxBallManifoldEntries];
    for (int i = 0; i < n; ++i) {
        const uint32_t key = bp.manifolds[i];
        island[i] = tag[KeyKind(key) <= 1 ? 0 : KeyBodyA(key) + 1];
        order[i] = static_cast<int8_t>(i);
    }
    if (n > 1) QuickSortByIsland(order, island, 0, n - 1);
    for (int i = 0; i < n; ++i) {
        t.entries[i].key = bp.manifolds[order[i]];
        t.entries[i].island = island[order[i]];
    }
    t.entry_count = n;
}

// Ball-vs-mesh contacts and their rows (AppendBallWorldSpecialContactRows for meshes only).
// A leaf's contact for the cooperative walk: what its commit needs, plus the triangle's own normal when events are
// written.
struct LeafContactPlain {
    BallManifoldPoint np;
    int32_t hit;
    RG_HD bool keep() const { return hit != 0; }  // commit changes nothing otherwise
};
struct LeafContactWithNormal {
    BallManifoldPoint np;
    Vec3 raw_normal;
    int32_t hit;
    RG_HD bool keep() const { return hit != 0; }
};

// kEvents: the caller writes physics events (compiled out otherwise).
template <typename Walk = SerialWalk, bool kEvents = false>
RG_HD void BallMeshContacts(BallState& ball, const StepParams& p, ArenaInternal& st, const ArenaMeshView& mesh,
                            BallTickScratch& t, const Vec3& ext_lin_uu, const Vec3& solver_vel_bt, float dt,
                            const EventSink& events) {
    constexpr int kMaxCps = kMaxBallContactPoints;
    BallContactPoint* cps = t.cps;
    int n_cp = 0;
    int agg_count = 0;
    Vec3 agg_normal(0.0f, 0.0f, 0.0f);
    float agg_rel_len = 0.0f;
    const float r = p.ball_radius;
    const float radius_bt = r * (1.0f / 50.0f);
    const float th = (radius_bt + 0.08f) * 0.02f;
    const float r_th = radius_bt + th;
    const Vec3 center = ball.pos_bt;
    const Vec3 qmin_bt = center - Vec3(r_th, r_th, r_th);
    const Vec3 qmax_bt = center + Vec3(r_th, r_th, r_th);
    constexpr float kSkip = 0.01f;
    uint32_t next_retained = st.ball_mesh_retained;
    int walk_rounds = 0;
    for (int mi = 0; mi < mesh.mesh_count; ++mi) {
        const MeshInfo& info = mesh.meshes[mi];
        BallMeshManifold& man = st.ball_mesh[mi];
        const int retained = ((st.ball_mesh_retained >> mi) & 1u) != 0u ? man.count : 0;
        const bool walk = !(info.aabb_max.x < qmin_bt.x - kSkip || info.aabb_min.x > qmax_bt.x + kSkip ||
                            info.aabb_max.y < qmin_bt.y - kSkip || info.aabb_min.y > qmax_bt.y + kSkip ||
                            info.aabb_max.z < qmin_bt.z - kSkip || info.aabb_min.z > qmax_bt.z + kSkip);
        if (retained == 0 && !walk) {
            next_retained &= ~(1u << mi);
            continue;
        }
        int count = retained;
        if (walk) {
            uint16_t qmin[3], qmax[3];
            QuantizePoint(info.q, center - Vec3(r_th, r_th, r_th), 0, qmin);
            QuantizePoint(info.q, center + Vec3(r_th, r_th, r_th), 1, qmax);
            Vec3 bx, by, bz;
            BallBasis(ball, bx, by, bz);
            // A leaf's contact point (independent of the other leaves), then its insertion into the manifold (in
            // node order).
            const auto contact = [&](int32_t local, BallManifoldPoint& np, Vec3& normal) {
                const MeshTriangle& tri = mesh.tris[info.tri_start + local];
                Vec3 point;
                float depth, pen_uu;
                if (!SphereTriangleContact(center, radius_bt, th, tri.a, tri.b, tri.c, normal, point, depth, pen_uu)) {
                    return false;
                }
                const Vec3 on_a = point + normal * depth;
                const Vec3 rel = on_a - center;
                np.local_a = Vec3(Dot(rel, bx), Dot(rel, by), Dot(rel, bz));
                Vec3 on_b = point;
                Vec3 cn = normal;
                AdjustInternalEdgeContact(tri, on_a, depth, on_b, cn);
                np.normal = cn;
                np.pen_uu = pen_uu;
                np.local_b = on_b;
                return true;
            };
            using LeafContact = typename std::conditional<kEvents, LeafContactWithNormal, LeafContactPlain>::type;
            constexpr float kLeafReachMarginBt = 0.01f;
            const SphereNearLeaf near_leaf(info.q, center, r_th + kLeafReachMarginBt);
            walk_rounds += Walk::Run(
                mesh.nodes, info.node_start, info.node_start + info.node_count, qmin, qmax, near_leaf,
                [&](int32_t local) {
                    BallManifoldPoint np;
                    Vec3 raw_normal;
                    if (contact(local, np, raw_normal)) {
                        if (kEvents && events.non_goal()) events.BallWorldContact(ball, raw_normal);
                        InsertManifoldPoint(man.pts, count, np, th);
                    }
                },
                [&](int32_t local) {
                    LeafContact c;
                    Vec3 raw_normal;
                    c.hit = contact(local, c.np, raw_normal) ? 1 : 0;
                    if constexpr (kEvents) c.raw_normal = raw_normal;
                    return c;
                },
                [&](const LeafContact& c) {
                    if (c.hit) {
                        if constexpr (kEvents) {
                            if (events.non_goal()) events.BallWorldContact(ball, c.raw_normal);
                        }
                        InsertManifoldPoint(man.pts, count, c.np, th);
                    }
                });
        }
        BallPointFull* full = t.full;
        RefreshManifold(man.pts, count, ball, center, th, full);
        man.count = static_cast<uint8_t>(count);
        next_retained = count > 0 ? (next_retained | (1u << mi)) : (next_retained & ~(1u << mi));
        for (int i = 0; i < count; ++i) {
            if (n_cp < kMaxCps) {
                BallContactPoint& c = cps[n_cp++];
                c.normal = full[i].p.normal;
                c.pen_uu = full[i].p.pen_uu;
                c.rel_len_bt = full[i].rel_len_bt;
                c.rel_pos1_bt = full[i].rel_pos1_bt;
                c.pen_bt = -full[i].distance_bt;
                c.static_id = static_cast<int16_t>(info.mesh_index);
            } else {
                ++agg_count;
                agg_normal += full[i].p.normal;
                agg_rel_len += full[i].rel_len_bt;
            }
        }
    }
    st.ball_mesh_retained = static_cast<uint16_t>(next_retained);
    st.walk_rounds = static_cast<uint16_t>(walk_rounds < 65535 ? walk_rounds : 65535);

    // Aggregate the contact points in manifold-list order (kind-0 entries), then the rest.
    uint32_t consumed = 0u;
    for (int e = 0; e < t.entry_count; ++e) {
        const uint32_t key = t.entries[e].key;
        if (KeyKind(key) != 0) continue;
        for (int i = 0; i < n_cp; ++i) {
            if (((consumed >> i) & 1u) == 0u && cps[i].static_id == KeyStatic(key)) {
                consumed |= 1u << i;
                ++agg_count;
                agg_normal += cps[i].normal;
                agg_rel_len += cps[i].rel_len_bt;
            }
        }
    }
    for (int i = 0; i < n_cp; ++i) {
        if (((consumed >> i) & 1u) == 0u) {
            ++agg_count;
            agg_normal += cps[i].normal;
            agg_rel_len += cps[i].rel_len_bt;
        }
    }
    if (agg_count == 0) return;

    for (int i = 0; i < n_cp; ++i) {
        const Vec3 rel_uu = -cps[i].normal * p.ball_radius;
        const Vec3 ang_bt = BallInvInertiaBt(ball, p, Cross(cps[i].rel_pos1_bt, cps[i].normal));
        AppendStaticSplitRow(t.rows, cps[i].pen_uu, cps[i].normal, rel_uu, p, dt, cps[i].rel_pos1_bt, cps[i].pen_bt,
                             ang_bt, cps[i].static_id);
    }
    const float inv_count = 1.0f / static_cast<float>(agg_count);
    Vec3 normal = agg_normal;
    float dist_bt = agg_rel_len;
    float fric = p.ball_friction;
    float rest = ball.constrained_2d ? 0.0f : p.ball_restitution;
    if (agg_count > 1) {
        float fs = fric, rs = rest;
        for (int i = 1; i < agg_count; ++i) {
            fs += fric;
            rs += rest;
        }
        fric = inv_count * fs;
        rest = inv_count * rs;
        const float l2 = Dot(normal, normal);
        normal = l2 >= kSimdEpsilon * kSimdEpsilon ? normal * (1.0f / sqrtf(l2)) : Vec3(1.0f, 0.0f, 0.0f);
        dist_bt = inv_count * dist_bt;
    }
    const Vec3 rel_bt = normal * (-dist_bt);
    const Vec3 rel_uu = rel_bt * kBtToUu;
    const Vec3 ang_n_uu = BallInvInertiaUu(p, Cross(rel_uu, normal));
    const Vec3 tangent_basis = (solver_vel_bt + UuToBtDiv(ext_lin_uu)) + Cross(ball.ang_vel, rel_bt);
    Vec3 tangent;
    {
        const float vn = Dot(normal, tangent_basis);
        const Vec3 lat = tangent_basis - normal * vn;
        const float lat2 = Dot(lat, lat);
        if (lat2 > 1.1920928955078125e-7f) {
            tangent = lat * (1.0f / sqrtf(lat2));
        } else {
            const Vec3 axis = Abs(normal.z) < 0.70710678f ? Vec3(0.0f, 0.0f, 1.0f) : Vec3(1.0f, 0.0f, 0.0f);
            tangent = Normalize(Cross(axis, normal), Vec3(1.0f, 0.0f, 0.0f));
        }
    }
    const Vec3 ang_t_uu = BallInvInertiaUu(p, Cross(rel_uu, tangent));
    const Vec3 ang_n_bt = BallInvInertiaBt(ball, p, Cross(rel_bt, normal));
    const Vec3 ang_t_bt = BallInvInertiaBt(ball, p, Cross(rel_bt, tangent));
    AppendStaticVelocityRows(t.rows, normal, rest, fric, rel_uu, ang_n_uu, ang_t_uu, ball.vel, ball.ang_vel, ext_lin_uu,
                             rel_bt, solver_vel_bt, ang_n_bt, ang_t_bt, tangent_basis, p, false);
}

RG_HD void BallPredictiveRows(const BallState& ball, const StepParams& p, const ArenaInternal& st,
                              BallTickScratch& t, const Vec3& ext_lin_uu, const Vec3& solver_vel_bt, float dt) {
    const Vec3 zero(0.0f, 0.0f, 0.0f);
    for (int k = 0; k < st.bp.predictive_count; ++k) {
        if (!t.predictive_hit) continue;
        const Vec3 n = t.predictive_normal;
        const float pen_bt = -t.predictive_distance_bt;
        AppendStaticSplitRow(t.rows, pen_bt * 50.0f, n, zero, p, dt, zero, pen_bt, zero,
                             static_cast<int16_t>(kPredictiveStaticId));
        const Vec3 basis = solver_vel_bt + UuToBtDiv(ext_lin_uu);
        AppendStaticVelocityRows(t.rows, n, 0.0f, p.ball_friction, zero, zero, zero, ball.vel, ball.ang_vel, ext_lin_uu,
                                 zero, solver_vel_bt, zero, zero, basis, p, true);
    }
}

RG_HD void SolveBallVelocity(BallState& ball, BallTickScratch& t, const Vec3& solver_vel_bt) {
    BallRows& R = t.rows;
    if (R.n_vel <= 0 && R.n_fric <= 0) return;
    int nord[kMaxBallVelRows], ford[kMaxBallVelRows];
    int nn = 0, nf = 0;
    if (R.n_vel + R.n_fric <= 2 && R.n_vel <= 1 && R.n_fric <= 1) {
        if (R.n_vel == 1) nord[nn++] = 0;
        if (R.n_fric == 1) ford[nf++] = 0;
    } else {
        uint32_t used = 0u;
        for (int e = 0; e < t.entry_count; ++e) {
            const uint32_t key = t.entries[e].key;
            if (KeyKind(key) == 0 && KeyStatic(key) == kPredictiveStaticId) {
                for (int i = 0; i < R.n_vel; ++i) {
                    if (((used >> i) & 1u) || !R.vel[i].predictive) continue;
                    used |= 1u << i;
                    nord[nn++] = i;
                }
            }
        }
        for (int i = 0; i < R.n_vel; ++i) {
            if (!((used >> i) & 1u)) nord[nn++] = i;
        }
        uint32_t fused = 0u;
        for (int k = 0; k < nn; ++k) {
            for (int i = 0; i < R.n_fric; ++i) {
                if (!((fused >> i) & 1u) && R.fric[i].normal_row == nord[k]) {
                    fused |= 1u << i;
                    ford[nf++] = i;
                }
            }
        }
        for (int i = 0; i < R.n_fric; ++i) {
            if (!((fused >> i) & 1u)) ford[nf++] = i;
        }
    }
    Vec3 dl(0.0f, 0.0f, 0.0f), da(0.0f, 0.0f, 0.0f);
    for (int iter = 0; iter < 10; ++iter) {
        bool changed = false;
        for (int k = 0; k < nn; ++k) {
            VelRow& r = R.vel[nord[k]];
            float di = r.rhs_bt;
            const float dv = ((r.normal.x * dl.x + r.normal.y * dl.y) + r.normal.z * dl.z) +
                             ((r.rel_cross_bt.x * da.x + r.rel_cross_bt.y * da.y) + r.rel_cross_bt.z * da.z);
            di = di - dv * r.jac;
            const float before = r.applied;
            const float sum = before + di;
            if (sum < 0.0f) {
                di = -before;
                r.applied = 0.0f;
            } else {
                r.applied = sum;
            }
            if (di == 0.0f) {
                changed = (FloatBits(before) != FloatBits(r.applied)) || changed;
                continue;
            }
            dl += (r.normal * r.inv_mass) * di;
            da += r.ang_bt * di;
            changed = true;
        }
        for (int k = 0; k < nf; ++k) {
            FricRow& r = R.fric[ford[k]];
            if (r.normal_row < 0 || r.normal_row >= R.n_vel) continue;
            const float total = R.vel[r.normal_row].applied;
            if (total <= 0.0f) continue;
            const float max_i = r.friction * total;
            float di = r.rhs_bt;
            const float dv = ((r.tangent.x * dl.x + r.tangent.y * dl.y) + r.tangent.z * dl.z) +
                             ((r.rel_cross_bt.x * da.x + r.rel_cross_bt.y * da.y) + r.rel_cross_bt.z * da.z);
            di = di - dv * r.jac;
            const float before = r.applied;
            const float sum = before + di;
            if (sum < -max_i) {
                di = -max_i - before;
                r.applied = -max_i;
            } else if (sum > max_i) {
                di = max_i - before;
                r.applied = max_i;
            } else {
                r.applied = sum;
            }
            if (di == 0.0f) {
                changed = (FloatBits(before) != FloatBits(r.applied)) || changed;
                continue;
            }
            dl += (r.tangent * r.inv_mass) * di;
            da += r.ang_bt * di;
            changed = true;
        }
        if (!changed) break;
    }
    SetBallVelBt(ball, solver_vel_bt + dl);
    ball.ang_vel += da;
}

RG_HD void SolveBallSplit(BallState& ball, BallTickScratch& t, float dt) {
    BallRows& R = t.rows;
    if (R.n_split <= 0) return;
    int ord[kMaxBallSplitRows];
    int n = 0;
    if (t.entry_count > 0 && R.n_split > 1) {
        uint32_t used = 0u;
        for (int e = 0; e < t.entry_count; ++e) {
            const uint32_t key = t.entries[e].key;
            if (KeyKind(key) != 0) continue;
            for (int i = 0; i < R.n_split; ++i) {
                if (((used >> i) & 1u) || R.split[i].static_id < 0 || R.split[i].static_id != KeyStatic(key)) continue;
                used |= 1u << i;
                ord[n++] = i;
            }
        }
        for (int i = 0; i < R.n_split; ++i) {
            if (!((used >> i) & 1u)) ord[n++] = i;
        }
    } else {
        for (int i = 0; i < R.n_split; ++i) ord[n++] = i;
    }
    Vec3 push(0.0f, 0.0f, 0.0f), turn(0.0f, 0.0f, 0.0f);
    for (int iter = 0; iter < 10; ++iter) {
        bool changed = false;
        for (int k = 0; k < n; ++k) {
            SplitRow& r = R.split[ord[k]];
            const float before = r.applied;
            float di = r.rhs_bt;
            const float term = (((r.normal.x * push.x + r.normal.y * push.y) + r.normal.z * push.z) +
                                ((r.rel_cross_bt.x * turn.x + r.rel_cross_bt.y * turn.y) + r.rel_cross_bt.z * turn.z)) *
                               r.jac_bt;
            di -= term;
            const float sum = r.applied + di;
            if (sum < 0.0f) {
                di = -r.applied;
                r.applied = 0.0f;
            } else {
                r.applied = sum;
            }
            if (di == 0.0f) {
                changed = (FloatBits(before) != FloatBits(r.applied)) || changed;
                continue;
            }
            push += (r.normal * r.inv_mass) * di;
            turn += r.ang_bt * di;
            changed = true;
        }
        if (!changed) break;
    }
    if (!(push.x == 0.0f && push.y == 0.0f && push.z == 0.0f)) SetBallPosBt(ball, ball.pos_bt + push * dt);
}

// One tick of an arena holding only the ball.
// The round countdown on render-frame time (CountdownUpdateTimers, env_step.hpp:6516-6538): phase 1 waits one
// second, phase 2 fires three times a second apart, then the countdown ends and the ball's 2D constraint lifts.
RG_HD void CountdownUpdateTimers(EnvState& env, BallState& ball, float frame_dt) {
    if (env.countdown_phase == 0) return;
    env.countdown_count += frame_dt;
    if (env.countdown_count > 1.0f) {
        if (env.countdown_phase == 1) {
            env.countdown_phase = 2;
            env.countdown_count = 0.0f;
        } else {
            env.countdown_count -= 1.0f;
            if (++env.countdown_fires == 3) {
                env.countdown_phase = 0;
                env.countdown_fires = 0;
                ball.constrained_2d = 0;
            }
        }
    }
}

// ContactWalk answers the contact query's mesh walks, RayWalk the predictive ray's.  With kEvents the tick writes
// its physics events to `events` (an off sink writes nothing); without it event code is compiled out.
template <typename ContactWalk = SerialWalk, typename RayWalk = ContactWalk, bool kEvents = false>
RG_HD void StepBallArena(BallState& ball, EnvState& env, ArenaInternal& st, const ArenaMeshView& mesh,
                         const StepParams& p, BallTickScratch& t, const EventSink& events = EventSink{}) {
    const float dt = p.tick_dt;
    // 1. env pre-tick.
    if (kEvents) events.SetTick(env.tick);
    env.goal_scored_this_tick = 0;
    // 2. ball pre-tick.
    BallPreTick(ball, p, t.predict_vel_bt);
    // 3a. activation.
    const bool vehicle_woke = (ball.sleeping & 4u) != 0;
    ball.sleeping &= 3u;
    if (ball.sleeping == 2u) ball.sleeping = 0u;  // an uploaded state starts awake on a custom arena
    if (ball.sleeping == 0u && ball.deactivation_time > 2.0f) {
        bool car_in_island = false;
        for (int i = 0; i < st.bp.pair_count; ++i) car_in_island = car_in_island || KeyKind(st.bp.pairs[i]) == 1;
        if (!car_in_island) {
            ball.sleeping = 1u;
            SetBallVelBt(ball, Vec3(0.0f, 0.0f, 0.0f));
            ball.ang_vel = Vec3(0.0f, 0.0f, 0.0f);
        }
    }
    const bool sleeping = ball.sleeping == 1u;
    if (vehicle_woke) ball.deactivation_time = 0.0f;
    const Vec3 ext_lin_uu = sleeping ? Vec3(0.0f, 0.0f, 0.0f) : Vec3(0.0f, 0.0f, p.gravity_z * dt);
    const Vec3 solver_vel_bt = ball.vel_bt;
    t.rows.n_split = t.rows.n_vel = t.rows.n_fric = 0;
    // 3b. predictive ray.
    t.predictive_hit = 0;
    if (!(sleeping && !vehicle_woke)) {
        t.predictive_hit = BallPredictiveRay<RayWalk>(ball, dt, mesh, t.predictive_normal, t.predictive_distance_bt) ? 1 : 0;
    }
    // 3c. broadphase.
    {
        constexpr float kContactThresholdBt = 0.02f;
        constexpr float kSphereAabbExtraBt = 0.08f;
        const Vec3 pos = ball.pos_bt;
        const Vec3 pred = pos + t.predict_vel_bt * dt;
        const float extent = p.ball_radius * (1.0f / 50.0f) + kSphereAabbExtraBt;
        const float pe = (p.ball_radius * (1.0f / 50.0f) + 0.5f * kSphereAabbExtraBt) + 0.5f * kSphereAabbExtraBt;
        BodyProxy& px = t.ball_proxy;
        px.aabb_min = Vec3(Min(pos.x, pred.x) - pe - kContactThresholdBt, Min(pos.y, pred.y) - pe - kContactThresholdBt,
                           Min(pos.z, pred.z) - pe - kContactThresholdBt);
        px.aabb_max = Vec3(Max(pos.x, pred.x) + pe + kContactThresholdBt, Max(pos.y, pred.y) + pe + kContactThresholdBt,
                           Max(pos.z, pred.z) + pe + kContactThresholdBt);
        px.child_min = Vec3(pos.x - extent, pos.y - extent, pos.z - extent);
        px.child_max = Vec3(pos.x + extent, pos.y + extent, pos.z + extent);
        px.valid = 1;
    }
    BallBroadphase(st, mesh, t, sleeping && !vehicle_woke);
    // 3d. contacts.
    const bool awake_rows = !sleeping || vehicle_woke;
    bool any_retained = false;
    for (int mi = 0; mi < mesh.mesh_count; ++mi) {
        any_retained = any_retained || (((st.ball_mesh_retained >> mi) & 1u) != 0u && st.ball_mesh[mi].count != 0);
    }
    const bool dispatched = !sleeping || vehicle_woke || any_retained;
    st.walk_rounds = 0;
    if (awake_rows && dispatched) {
        BallMeshContacts<ContactWalk, kEvents>(ball, p, st, mesh, t, ext_lin_uu, solver_vel_bt, dt, events);
    }
    // 3e. predictive rows.
    if (t.predictive_hit) BallPredictiveRows(ball, p, st, t, ext_lin_uu, solver_vel_bt, dt);
    // 3f. solve.
    SolveBallVelocity(ball, t, solver_vel_bt);
    SolveBallSplit(ball, t, dt);
    // 3g. gravity, integration, deactivation.
    if (!sleeping) SetBallVelBt(ball, ball.vel_bt + UuToBtDiv(ext_lin_uu));
    if (awake_rows) {
        const Vec3 v = ball.vel_bt;
        if (!(v.x == 0.0f && v.y == 0.0f && v.z == 0.0f)) SetBallPosBt(ball, ball.pos_bt + v * dt);
        if (!ball.basis_authoritative) {
            QuatColumns(ball.rot, ball.basis_forward, ball.basis_right, ball.basis_up);
            ball.basis_authoritative = 1;
        }
        ball.restored_inv_inertia_valid = 0;
        ball.sleeping = 0;
        const bool below = LengthSq(ball.vel_bt) < 0.8f * 0.8f && LengthSq(ball.ang_vel) < 1.0f * 1.0f;
        ball.deactivation_time = below && !sleeping ? ball.deactivation_time + dt : 0.0f;
    }
    // 4. env tail.
    if (LengthSq(ball.vel_impulse_cache_bt) > 0.0f || LengthSq(ball.vel_impulse_cache) > 0.0f) {
        SetBallVelBt(ball, ball.vel_bt + ball.vel_impulse_cache_bt);
        ball.vel_impulse_cache = Vec3(0.0f, 0.0f, 0.0f);
        ball.vel_impulse_cache_bt = Vec3(0.0f, 0.0f, 0.0f);
    }
    ball.inertia_basis_valid = 0;
    ball.tick_count_since_update += 1;
    if (ball.constrained_2d) --ball.constrained_2d;
    if (p.enable_scoring) {
        for (int g = 0; g < 2; ++g) {
            const Vec3& lo = p.goal_min[g];
            const Vec3& hi = p.goal_max[g];
            if (Dot(ball.pos - p.goal_point[g], p.goal_normal[g]) >= p.ball_radius && ball.pos.x >= lo.x &&
                ball.pos.y >= lo.y && ball.pos.z >= lo.z && ball.pos.x <= hi.x && ball.pos.y <= hi.y &&
                ball.pos.z <= hi.z) {
                const int scorer = g == 0 ? 1 : 0;
                env.last_scored_team = scorer;
                env.goal_scored_this_tick = 1;
                if (!env.goal_latched) (scorer == 0 ? env.blue_score : env.orange_score) += 1;
                break;
            }
        }
        env.goal_latched = env.goal_scored_this_tick;
    }
    if (kEvents && env.goal_scored_this_tick && events.on()) events.Goal(ball, env.last_scored_team);
    env.frame_ticks += 1;
    if (p.frame_frac >= 0.0f) {
        // The render frame ends on this tick: the countdown timers advance by its DeltaSeconds.
        const float frame_dt = (static_cast<float>(env.frame_ticks) + p.frame_frac - env.frame_frac_prev) * p.tick_dt;
        env.frame_ticks = 0;
        env.frame_frac_prev = p.frame_frac;
        CountdownUpdateTimers(env, ball, frame_dt);
    }
    env.tick += 1;
}

}  // namespace core
}  // namespace rg
