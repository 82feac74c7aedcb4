// The car half of an arena tick: the vehicle update between the car pre-tick and the world step (wheel rays,
// suspension, tyre friction, drive, sticky force, air control, flip torque, jump hold, auto-roll, boost), and the car
// rigid body inside the world step (external forces before the contacts, the broadphase proxy, integration).
//
// Reference (the reference CPU build's StepEnvGeneric path, include/rocketgpu/detail/step/):
//   vehicle update   PreTickCar env_step.hpp:3115-3510 -> PreTickVehicleDynamics raycast_wheels_and_contact_prep.hpp:
//                    2618-2975 (wheel rays :2074-2238 and :641-862, drive set-up :1697-1830, sticky force), then
//                    ApplyFlipTorque :3432, ApplyAirControl :3048, UpdateJumpState, ApplyAutoRollReference :1021,
//                    ApplyWheelSuspensionAndFriction :1475, ApplyBoost :3252
//   world step       WorldStepPrestepCar env_step.hpp:3737, WorldStepFinishCar :3789, ComputeBulletBodyProxy
//                    arena_geometry.hpp:3873-4033, IntegrateCarTransformOnly raycast_wheels...:3807, orientation
//                    math.hpp:1202-1227
//
// Layout of this file is its own: a car's tick works on a CarKin (the car record plus the per-tick frame and force
// queue), and a wheel is a WheelRay holding only what the later phases read.  Float expressions keep the
// reference's operation order, which bit-exactness fixes (no contraction, correctly rounded / and sqrt).
//
// What the tick does not step yet is reported, not guessed: every path that would need a body contact this build
// does not compute (a car's hitbox on a triangle, a wheel ray on another car, a car's predictive contact, car-ball or
// car-car contacts) sets a bit of the arena's `unsupported` word and leaves the state as the rest of the tick makes
// it.  Callers read the word; an arena with a bit set is no longer the reference's.
#pragma once

#include "rg_ball_step.hpp"
#include "rg_car_phases.hpp"
#include "rg_gjk.hpp"

namespace rg {
namespace core {

// Bits of an arena's unsupported-physics word.
enum : uint32_t {
    kUnsupportedCarWorld = 1u << 0,        // a car's hitbox may touch a mesh triangle
    kUnsupportedBallCar = 1u << 1,         // the ball's and a car's boxes overlap
    kUnsupportedCarCar = 1u << 2,          // two cars' boxes overlap
    kUnsupportedWheelOnCar = 1u << 3,      // unused: wheel rays are cast against the other cars
    kUnsupportedCarPredictive = 1u << 4,   // a car's swept center crosses a triangle within the tick
    kUnsupportedRespawnSpot = 1u << 5,     // a demolished car picks its respawn spot this tick
    kUnsupportedSetWheelNormals = 1u << 6, // an uploaded car carries set-state wheel normals
    kUnsupportedThreeWheels = 1u << 7,     // a three-wheel body
};

// ------------------------------------------------------------------------------------------------------------
// Per-tick car frame: the orientation columns read after the pre-tick, the Bullet inverse inertia (local diagonal
// and world rows from the tick-start basis), and the force queue the vehicle update fills for the solver.

struct CarKin {
    Vec3 fwd, right, up;
    Vec3 inv_local_bt;
    Vec3 inv_row0, inv_row1, inv_row2;
    CarForceQueue q;
};

RG_HD void MakeCarKin(const CarState& car, const CarConfig& cfg, float mass, CarKin& k) {
    CarBasis(car, k.fwd, k.right, k.up);
    k.inv_local_bt = InvInertiaLocalBt(cfg, mass);
    if (car.restored_inv_inertia_valid) {
        k.inv_row0 = car.restored_inv_inertia_bt[0];
        k.inv_row1 = car.restored_inv_inertia_bt[1];
        k.inv_row2 = car.restored_inv_inertia_bt[2];
    } else {
        InvInertiaWorldRows(car.tick_start_basis_forward, car.tick_start_basis_right, car.tick_start_basis_up,
                            k.inv_local_bt, k.inv_row0, k.inv_row1, k.inv_row2);
    }
    k.q = CarForceQueue{};
}

RG_HD CarFrame KinFrame(const CarKin& k) {
    CarFrame f;
    f.fwd = k.fwd;
    f.right = k.right;
    f.up = k.up;
    f.inv_row0 = k.inv_row0;
    f.inv_row1 = k.inv_row1;
    f.inv_row2 = k.inv_row2;
    return f;
}

// The world inverse inertia (Bullet rows) applied to a torque, row by row.
RG_HD Vec3 KinInvInertia(const CarKin& k, const Vec3& t) { return MulRows(k.inv_row0, k.inv_row1, k.inv_row2, t); }

// Puts `force` and `torque` ahead of everything queued so far (QueueFirstForceAndTorqueBt,
// host_validation_and_contact_debug.hpp:771-796): the sums restart from them and add the recorded items in order.
// With more than four items of either kind recorded, the two are simply queued.
RG_HD void QueueFirst(CarForceQueue& q, const Vec3& force, const Vec3& torque) {
    if (q.n_force > 4 || q.n_torque > 4) {
        QueueCarForce(q, force);
        QueueCarTorque(q, torque);
        return;
    }
    Vec3 f = force;
    for (int i = 0; i < q.n_force; ++i) f += q.force_items[i];
    Vec3 t = torque;
    for (int i = 0; i < q.n_torque; ++i) t += q.torque_items[i];
    q.force = f;
    q.torque_bt = t;
    q.n_force = 5;
    q.n_torque = 5;
}

// ------------------------------------------------------------------------------------------------------------
// Curves of the vehicle (constants.hpp:116-187): piecewise linear over sorted knots, clamped at both ends.

template <int N>
RG_HD float Piecewise(const float (&x)[N], const float (&y)[N], float v) {
    if (x[0] >= v) return y[0];
    if (v >= x[N - 1]) return y[N - 1];
    for (int i = 1; i < N; ++i) {
        if (x[i] > v) return (y[i] - y[i - 1]) * ((v - x[i - 1]) / (x[i] - x[i - 1])) + y[i - 1];
    }
    return y[N - 1];
}

RG_HD float SteerDegAtSpeed(const float (&curve)[6], float speed) {
    const float knots[6] = {0.0f, 500.0f, 1000.0f, 1500.0f, 1750.0f, 3000.0f};
    return Piecewise(knots, curve, Abs(speed));
}
RG_HD float PowerslideSteerDegAtSpeed(const float (&curve)[2], float speed) {
    const float knots[2] = {0.0f, 2500.0f};
    return Piecewise(knots, curve, Abs(speed));
}
RG_HD float DriveTorqueAtSpeed(float forward_speed) {
    const float knots[5] = {-1410.0f, -1400.0f, 0.0f, 1400.0f, 1410.0f};
    const float scale[5] = {0.0f, 0.1f, 1.0f, 0.1f, 0.0f};
    return Piecewise(knots, scale, forward_speed);
}
RG_HD float NonStickyFrictionAt(float normal_z) {
    const float knots[3] = {0.0f, 0.707f, 1.0f};
    const float scale[3] = {0.1f, 0.5f, 1.0f};
    return Piecewise(knots, scale, normal_z);
}
RG_HD float LateralFrictionAt(float slip) {
    const float knots[2] = {0.0f, 1.0f};
    const float scale[2] = {1.0f, 0.2f};
    return Piecewise(knots, scale, slip);
}
RG_HD float HandbrakeLongFrictionAt(float slip) {
    const float knots[2] = {0.0f, 1.0f};
    const float scale[2] = {0.5f, 0.9f};
    return Piecewise(knots, scale, slip);
}

// ------------------------------------------------------------------------------------------------------------
// Wheels.

constexpr float kWheelMaxTravel = 14.5f;
constexpr float kWheelTravelTrim = 2.5f;  // WheelSuspensionSubtraction
constexpr float kSuspensionStiffness = 500.0f;
constexpr float kSuspensionCompression = 25.0f;
constexpr float kSuspensionRelaxation = 40.0f;

enum : uint8_t { kGroundNone = 0, kGroundMesh = 1, kGroundBall = 2, kGroundCar = 3 };

// One wheel for one tick: its ray, what the ray found, and the friction terms the drive set-up leaves for the
// friction impulse.  Wheel order: 0 front left-of-offset, 1 front mirrored, 2 back, 3 back mirrored.
struct WheelRay {
    Vec3 from;       // ray source (bt)
    Vec3 down;       // ray direction: -up
    Vec3 axle;       // -right
    Vec3 hit_point;  // bt
    Vec3 normal;
    float trace_len;     // source to hit along up (bt)
    float susp_len;      // trace_len - radius (bt)
    float rel_vel_bt;    // suspension velocity along the normal, over the normal's cosine
    float rel_vel_uu;
    float inv_cos;       // 1 / (normal . up), 10 when flatter than 0.1
    float pushback;
    float radius;        // uu
    float rest;          // suspension rest length (uu)
    float force_scale;   // share of the car's mass on this wheel
    Vec3 fric_axle, fric_fwd;
    float fric_roll, fric_lat, fric_long, fric_scale;
    int32_t static_id;
    uint8_t ground;      // kGround*
    uint8_t front, contact, on_world;
    int8_t ground_car;   // the car slot under the wheel (kGroundCar)
};

// Bone height of the suspension, rounded to 0.01 uu (raycast_wheels_and_contact_prep.hpp:458-464).
RG_HD float WheelBoneZ(const CarConfig& cfg, bool front) {
    const float rest = (front ? cfg.front_suspension_rest : cfg.back_suspension_rest) - 12.0f;
    return floorf((cfg.hitbox_offset.z - rest) * 100.0f + 0.5f) / 100.0f;
}

// The mass each front / back wheel carries (CalSuspensionForceScales, :627-639).
RG_HD void WheelMassShares(const CarConfig& cfg, float car_mass, float& front, float& back) {
    const float mass = 1.0f / (1.0f / car_mass);
    const float fx = cfg.front_wheels_offset.x * (1.0f / 50.0f);
    const float bx = -(cfg.back_wheels_offset.x * (1.0f / 50.0f));
    if (fx < 0.0f || bx < 0.0f) {
        front = back = (mass * 0.5f) * 0.5f;
    } else {
        const float back_share = (fx / (bx + fx)) * mass;
        front = (mass - back_share) * 0.5f;
        back = back_share * 0.5f;
    }
}

// Ray length of a wheel: rest + radius + travel - trim, summed in bt as the reference does (:1-8).
RG_HD float WheelRayLenBt(const WheelRay& w) {
    const float rest_bt = w.rest * (1.0f / 50.0f);
    const float travel_bt = (kWheelMaxTravel - kWheelTravelTrim) * (1.0f / 50.0f);
    const float radius_bt = w.radius * (1.0f / 50.0f);
    return (rest_bt + radius_bt) + travel_bt;
}

// Set a wheel up from the car's pose (InitWheelStepState, :420-456).
RG_HD void InitWheel(WheelRay& w, const CarState& car, const CarConfig& cfg, const CarKin& k, int idx, float bone_z) {
    const bool front = idx < 2;
    Vec3 off = front ? cfg.front_wheels_offset : cfg.back_wheels_offset;
    if (idx & 1) off.y = -off.y;
    off.z = cfg.hitbox_offset.z;
    Vec3 f(1.0f, 0.0f, 0.0f), r(0.0f, 1.0f, 0.0f), u(0.0f, 0.0f, 1.0f);
    if (car.basis_authoritative) {
        f = k.fwd;
        r = k.right;
        u = k.up;
    } else {
        QuatColumns(car.rot, f, r, u);
    }
    const Vec3 off_bt = off * (1.0f / 50.0f);
    w.from = car.pos_bt + ((f * off_bt.x + r * off_bt.y) + u * off_bt.z);
    w.down = -u;
    w.axle = -r;
    w.radius = front ? cfg.front_wheel_radius : cfg.back_wheel_radius;
    w.rest = off.z - bone_z;
    w.front = front ? 1 : 0;
    w.normal = -w.down;
    w.hit_point = Vec3(0.0f, 0.0f, 0.0f);
    w.trace_len = 0.0f;
    w.susp_len = 0.0f;
    w.rel_vel_bt = 0.0f;
    w.rel_vel_uu = 0.0f;
    w.inv_cos = 1.0f;
    w.pushback = 0.0f;
    w.fric_axle = Vec3(0.0f, -1.0f, 0.0f);
    w.fric_fwd = Vec3(1.0f, 0.0f, 0.0f);
    w.fric_roll = w.fric_lat = w.fric_long = w.fric_scale = 0.0f;
    w.static_id = -1;
    w.ground_car = -1;
    w.ground = kGroundNone;
    w.contact = 0;
    w.on_world = 0;
}

RG_HD void WheelMiss(WheelRay& w) {
    w.rel_vel_uu = 0.0f;
    w.inv_cos = 1.0f;
    w.pushback = 0.0f;
    w.contact = 0;
    w.on_world = 0;
    w.ground = kGroundNone;
    w.normal = -w.down;
}

// Suspension velocity terms at the hit (ComputeWheelRayVelocityTerms, :846-862).
RG_HD void WheelHitVelocity(const CarState& car, WheelRay& w) {
    const Vec3 up = -w.down;
    const float cos_n = Dot(w.normal, up);
    const Vec3 rel = w.hit_point - car.pos_bt;
    const Vec3 v = car.vel_bt + Cross(car.ang_vel, rel);
    const float along = Dot(w.normal, v);
    if (cos_n > 0.1f) {
        w.inv_cos = 1.0f / cos_n;
        w.rel_vel_bt = along * w.inv_cos;
        w.rel_vel_uu = w.rel_vel_bt * 50.0f;
    } else {
        w.inv_cos = 10.0f;
        w.rel_vel_uu = 0.0f;
        w.rel_vel_bt = 0.0f;
    }
}

// Record a hit at `p` with normal `n`.
RG_HD void WheelHit(const CarState& car, WheelRay& w, const Vec3& p, const Vec3& n, uint8_t ground, bool world,
                    int32_t static_id) {
    const Vec3 up = -w.down;
    w.hit_point = p;
    w.normal = n;
    w.contact = 1;
    w.on_world = world ? 1 : 0;
    w.ground = ground;
    w.static_id = world ? static_id : -1;
    w.trace_len = Dot(w.from - p, up);
    w.susp_len = w.trace_len - w.radius * (1.0f / 50.0f);
    WheelHitVelocity(car, w);
}

// Ray against a sphere by conservative advancement, Bullet's btGjkConvexCast of a point (CastGjkRayVsSphere,
// :190-262): step the ray parameter by distance over approach speed until within 0.001 bt.
RG_HD bool RaySphereAdvance(const Vec3& from, const Vec3& to, const Vec3& center, float radius, float& frac,
                            Vec3& n) {
    const auto closest = [&](const Vec3& p, const Vec3& c, float& dist, Vec3& nn) {
        const Vec3 mid = (p + c) * 0.5f;
        const Vec3 w = (p - mid) - (c - mid);
        const float len2 = Dot(w, w);
        if (!(len2 > 1.1920929e-7f * 1.1920929e-7f)) return false;
        const float rlen = 1.0f / sqrtf(len2);
        nn = w * rlen;
        dist = (1.0f / rlen) - radius;
        return true;
    };
    const Vec3 r = to - from;
    float lambda = 0.0f, last = 0.0f, dist = 0.0f;
    Vec3 nn(0.0f, 0.0f, 0.0f);
    if (!closest(from, center, dist, nn)) return false;
    int iter = 0;
    while (dist > 0.001f) {
        if (++iter > 64) return false;
        lambda = lambda - dist / Dot(r, nn);
        if (lambda > 1.0f || lambda < 0.0f || lambda <= last) return false;
        last = lambda;
        const Vec3 p = from * (1.0f - lambda) + to * lambda;
        const Vec3 c = center * lambda + center * (1.0f - lambda);
        float next = 0.0f;
        if (!closest(p, c, next, nn)) return false;
        if (next < 0.0f) {
            frac = last;
            n = nn;
            return true;
        }
        dist = next;
    }
    if (Dot(nn, r) >= 0.0f) return false;
    frac = lambda;
    n = nn;
    return true;
}

// The wheel ray against the ball (CastWheelRayAgainstBall, :264-327).  Returns the hit fraction or -1.
RG_HD float WheelRayBall(const WheelRay& w, const BallState& ball, float ball_radius, float len_bt, Vec3& n_out,
                         Vec3& p_out) {
    const float rb = ball_radius * (1.0f / 50.0f);
    const Vec3 d = ball.pos_bt - w.from;
    const float reject = rb + len_bt + 1.0e-4f;
    if (Dot(d, d) > reject * reject) return -1.0f;
    const Vec3 to = w.from + w.down * len_bt;
    float frac = 0.0f;
    Vec3 n(0.0f, 0.0f, 0.0f);
    const bool hit = RaySphereAdvance(w.from, to, ball.pos_bt, rb, frac, n);
    const float n2 = Dot(n, n);
    if (!hit || n2 <= 0.0001f) return -1.0f;
    n = n * (1.0f / sqrtf(n2));
    if (frac >= 1.0f) return -1.0f;
    n_out = n;
    p_out = (1.0f - frac) * w.from + frac * to;
    return frac;
}

// Pushback of a wheel sunk below its threshold: the static-contact impulse a Bullet solver step would give at the
// hit, a quarter per wheel (SetWorldWheelPushback, :2050-2071, with ResolveStaticCollisionImpulseMagnitudeBtInputs,
// host_validation_and_contact_debug.hpp:435-457).
RG_HD void WheelPushback(const CarState& car, float mass, const CarKin& k, WheelRay& w) {
    constexpr float kSolverDt = 1.0f / 60.0f;
    const float thresh = (w.rest * (1.0f / 50.0f) + w.radius * (1.0f / 50.0f)) - kWheelTravelTrim * (1.0f / 50.0f);
    if (!(w.trace_len < thresh)) return;
    const float depth = w.trace_len - thresh;
    const Vec3 rel = w.hit_point - car.pos_bt;
    const Vec3 v = car.vel_bt + Cross(car.ang_vel, rel);
    const float rel_vel = Dot(w.normal, v);
    // Denominator: 1/m + n . ((I^-1 (r x n)) x r), I^-1 applied by columns of the rows.
    const Vec3 c0 = Cross(w.hit_point - car.pos_bt, w.normal);
    const Vec3 ang(k.inv_row0.x * c0.x + k.inv_row1.x * c0.y + k.inv_row2.x * c0.z,
                   k.inv_row0.y * c0.x + k.inv_row1.y * c0.y + k.inv_row2.y * c0.z,
                   k.inv_row0.z * c0.x + k.inv_row1.z * c0.y + k.inv_row2.z * c0.z);
    const float denom = 1.0f / mass + Dot(w.normal, Cross(ang, w.hit_point - car.pos_bt));
    float impulse = 0.0f;
    if (!(denom <= 0.0f)) {
        const float inv = 1.0f / denom;
        const float pen = (0.2f * -depth / kSolverDt) * inv;
        const float vel = -rel_vel * inv;
        const float sum = pen + vel;
        impulse = sum > 0.0f ? sum : 0.0f;
    }
    w.pushback = impulse / 4.0f;
}

// The hitbox size Bullet is given: the published size when the config carries it, grown by the game's 0.025 bt
// collision skin (BulletHitboxSize, math.hpp:1554-1560).
RG_HD Vec3 BulletHitboxSizeUu(const CarConfig& cfg) {
    const Vec3 sized = cfg.hitbox_published + Vec3(2.5f, 2.5f, 2.5f);
    const bool has = sized.x == cfg.hitbox_size.x && sized.y == cfg.hitbox_size.y && sized.z == cfg.hitbox_size.z;
    const Vec3 published = has ? cfg.hitbox_published : cfg.hitbox_size - Vec3(2.5f, 2.5f, 2.5f);
    const Vec3 half_bt = (published * 0.02f) * 0.5f + Vec3(0.025f, 0.025f, 0.025f);
    return (half_bt * 50.0f) * 2.0f;
}

// Bullet half extents of the hitbox box shape, the collision margin taken off and put back
// (BulletBoxHalfExtentWithMarginBt, host_validation_and_contact_debug.hpp:43-54).
RG_HD Vec3 HitboxHalfBt(const CarConfig& cfg) {
    const Vec3 half = (BulletHitboxSizeUu(cfg) * (1.0f / 50.0f)) / 2.0f;
    const float safe = 0.1f * Min(half.x, Min(half.y, half.z));
    const float margin = safe < 0.04f ? safe : 0.04f;
    return (half - Vec3(margin, margin, margin)) + Vec3(margin, margin, margin);
}

// Closest points of a point and a box with a collision margin by Bullet's GJK pair detector (Voronoi simplex, the two
// shapes shifted by their midpoint, the normal turned to point from the box's center toward the point); the result
// is kept when nearer than `dist` (WheelGjkPointBoxClosest, raycast_wheels_and_contact_prep.hpp:4758-4855).
RG_HD bool PointBoxClosest(const Vec3& point, const Vec3& origin, const Vec3& f, const Vec3& r, const Vec3& u,
                           const Vec3& core, float box_margin, bool& has, float& dist, Vec3& normal) {
    const Vec3 offset = (point + origin) * 0.5f;
    const Vec3 la = point - offset;
    const Vec3 lb = origin - offset;
    const float margin = 0.0f + box_margin;
    Vec3 axis(0.0f, 1.0f, 0.0f);
    float sq = 1.0e18f;
    gjk::Simplex sx;
    gjk::SimplexReset(sx);
    bool check = false;
    int iter = 0;
    const Vec3 pw = Vec3(0.0f, 0.0f, 0.0f) + la;
    while (true) {
        const Vec3 sb(Dot(f, axis), Dot(r, axis), Dot(u, axis));
        const Vec3 ql(sb.x >= 0.0f ? core.x : -core.x, sb.y >= 0.0f ? core.y : -core.y,
                      sb.z >= 0.0f ? core.z : -core.z);
        const Vec3 qw((f.x * ql.x + r.x * ql.y) + u.x * ql.z + lb.x, (f.y * ql.x + r.y * ql.y) + u.y * ql.z + lb.y,
                      (f.z * ql.x + r.z * ql.y) + u.z * ql.z + lb.z);
        const Vec3 w = pw - qw;
        const float delta = Dot(axis, w);
        if (delta > 0.0f && delta * delta > sq * 1.0e18f) {
            check = true;
            break;
        }
        if (gjk::SimplexHas(sx, w)) {
            check = true;
            break;
        }
        if (sq - delta <= sq * 1.0e-6f) {
            check = true;
            break;
        }
        gjk::SimplexAdd(sx, w, pw, qw);
        if (!gjk::SimplexUpdate(sx)) {
            check = true;
            break;
        }
        const Vec3 next = sx.near_v;
        if (LengthSq(next) < 1.0e-6f) {
            axis = next;
            check = true;
            break;
        }
        const float prev = sq;
        sq = LengthSq(next);
        if (prev - sq <= 1.1920929e-7f * prev) {
            check = true;
            break;
        }
        axis = next;
        if (iter++ > 1000) break;
        if (sx.n == 4) break;
    }
    if (!check) return has;
    gjk::SimplexUpdate(sx);
    const float len_sq = LengthSq(axis);
    if (!(len_sq > 1.1920929e-7f * 1.1920929e-7f)) return has;
    const float rlen = 1.0f / sqrtf(len_sq);
    Vec3 n = axis * rlen;
    const float d = (1.0f / rlen) - margin;
    if (!(d < 0.0f || d * d < 1.0e18f)) return has;
    const Vec3 hm = core + Vec3(box_margin, box_margin, box_margin);
    const Vec3 ext((hm.x * Abs(f.x) + hm.y * Abs(r.x)) + hm.z * Abs(u.x),
                   (hm.x * Abs(f.y) + hm.y * Abs(r.y)) + hm.z * Abs(u.y),
                   (hm.x * Abs(f.z) + hm.y * Abs(r.z)) + hm.z * Abs(u.z));
    const Vec3 zero(0.0f, 0.0f, 0.0f);
    const Vec3 ca = ((la + zero) + (la - zero)) * 0.5f;
    const Vec3 cb = ((lb - ext) + (lb + ext)) * 0.5f;
    if (Dot(ca - cb, n) < 0.0f) n = n * -1.0f;
    if (d < dist) {
        has = true;
        normal = n;
        dist = d;
    }
    return has;
}

// A wheel's ray against another car's hitbox (CastWheelRayAgainstCarHitboxGjkCast,
// raycast_wheels_and_contact_prep.hpp:4857-4979): a sphere and slab rejection, then Bullet's GJK convex cast of the
// ray's point against the box (advance by distance over approach speed until within 0.001 bt).  On a hit: the
// distance along the ray in uu, the point (bt) and the box's normal there.
RG_HD bool WheelRayCarHit(const Vec3& src, const Vec3& dir, float len_uu, float len_bt, const CarState& other,
                          const CarConfig& ocfg, float& t_uu, Vec3& point, Vec3& normal) {
    const Vec3 to = src + dir * len_bt;
    const Vec3 half = BulletHitboxSizeUu(ocfg) * (0.5f * (1.0f / 50.0f));
    const float margin = Min(0.1f * Min(half.x, Min(half.y, half.z)), 0.04f);
    const Vec3 inner = half - Vec3(margin, margin, margin);
    Vec3 f, r, u;
    CarBasis(other, f, r, u);
    const Vec3 off = ocfg.hitbox_offset * (1.0f / 50.0f);
    const Vec3 center = other.pos_bt + ((f * off.x + r * off.y) + u * off.z);
    {
        const float rr = Length(half) + len_bt + 1.0e-3f;
        if (LengthSq(center - src) > rr * rr) return false;
    }
    {
        const Vec3 rf = src - other.pos_bt, rt = to - other.pos_bt;
        const Vec3 lf(Dot(rf, f), Dot(rf, r), Dot(rf, u));
        const Vec3 lt(Dot(rt, f), Dot(rt, r), Dot(rt, u));
        const Vec3 ext = inner + Vec3(margin, margin, margin);
        const Vec3 bmin = off - ext, bmax = off + ext;
        if ((lf.x < bmin.x && lt.x < bmin.x) || (lf.x > bmax.x && lt.x > bmax.x) || (lf.y < bmin.y && lt.y < bmin.y) ||
            (lf.y > bmax.y && lt.y > bmax.y) || (lf.z < bmin.z && lt.z < bmin.z) || (lf.z > bmax.z && lt.z > bmax.z)) {
            return false;
        }
        Vec3 ld = lt - lf;
        const float ll = Length(ld);
        if (!(ll > 0.0f)) return false;
        ld = ld * (1.0f / ll);
        const float lambda_max = Dot(ld, lt - lf);
        const float inv[3] = {ld.x == 0.0f ? 1.0e30f : 1.0f / ld.x, ld.y == 0.0f ? 1.0e30f : 1.0f / ld.y,
                              ld.z == 0.0f ? 1.0e30f : 1.0f / ld.z};
        const float from[3] = {lf.x, lf.y, lf.z};
        const float lo[3] = {bmin.x, bmin.y, bmin.z};
        const float hi[3] = {bmax.x, bmax.y, bmax.z};
        float tmin = ((inv[0] < 0.0f ? hi[0] : lo[0]) - from[0]) * inv[0];
        float tmax = ((inv[0] < 0.0f ? lo[0] : hi[0]) - from[0]) * inv[0];
        for (int k = 1; k < 3; ++k) {
            const float t0 = ((inv[k] < 0.0f ? hi[k] : lo[k]) - from[k]) * inv[k];
            const float t1 = ((inv[k] < 0.0f ? lo[k] : hi[k]) - from[k]) * inv[k];
            if (tmin > t1 || t0 > tmax) return false;
            if (t0 > tmin) tmin = t0;
            if (t1 < tmax) tmax = t1;
        }
        if (!(tmin < lambda_max && tmax > 0.0f)) return false;
    }
    const Vec3 ray(to.x - src.x - (center.x - center.x), to.y - src.y - (center.y - center.y),
                   to.z - src.z - (center.z - center.z));
    bool has = false;
    float dist = 1.0e18f;
    Vec3 n(0.0f, 0.0f, 0.0f);
    if (!PointBoxClosest(src, center, f, r, u, inner, margin, has, dist, n)) return false;
    float lambda = 0.0f;
    int iter = 0;
    while (true) {
        if (dist <= 0.001f) {
            if (-0.0f <= Dot(ray, n)) return false;
            break;
        }
        if (++iter > 64) return false;
        const float next = lambda - dist / Dot(ray, n);
        if (next > 1.0f || next < 0.0f || next <= lambda) return false;
        const float s = 1.0f - next;
        const Vec3 pa(next * to.x + s * src.x, next * to.y + s * src.y, next * to.z + s * src.z);
        const Vec3 bo(next * center.x + s * center.x, next * center.y + s * center.y, next * center.z + s * center.z);
        if (!PointBoxClosest(pa, bo, f, r, u, inner, margin, has, dist, n)) return false;
        lambda = next;
        if (!(0.0f <= dist)) break;
    }
    const float n2 = LengthSq(n);
    if (!(n2 > 1.0e-4f) || !(lambda < 1.0f)) return false;
    const float inv_len = 1.0f / sqrtf(n2);
    const float s = 1.0f - lambda;
    t_uu = lambda * len_uu;
    normal = Vec3(n.x * inv_len, n.y * inv_len, n.z * inv_len);
    point = Vec3(s * src.x + lambda * to.x, s * src.y + lambda * to.y, s * src.z + lambda * to.z);
    return true;
}

// The wheel's segment against the arena meshes: its length and the nearest fraction a hit may have.
RG_HD void WheelMeshSegment(const WheelRay& w, float min_hit_sq, Vec3& delta, float& min_frac_sq) {
    delta = w.down * WheelRayLenBt(w);
    min_frac_sq = min_hit_sq > 0.0f ? min_hit_sq / LengthSq((w.from + delta) - w.from) : 0.0f;
}

// The wheel's ray against the arena, the other cars and the ball; nearest wins (RayCastWheelAgainstArena, :641-844,
// for a custom arena: meshes only, no analytic planes).  `mh`: the ray's nearest mesh hit (CastArenaRay over the
// segment WheelMeshSegment gives).  `others_near`: the other bodies whose last swept boxes the car's rays reach
// (bit 0 ball, bit 1 + j car j); a car there cannot be cast here yet, which the arena reports.
RG_HD void CastWheel(WheelRay& w, const RayHit& mh, const CarState& car, const CarState* cars, const CarConfig* cfgs,
                     int count, int self, const ArenaMeshView& mesh, const BallState& ball, float ball_radius,
                     uint32_t others_near, float min_hit_sq, uint32_t& unsupported) {
    const float len_bt = WheelRayLenBt(w);
    const float len_uu = ((w.rest + kWheelMaxTravel) + w.radius) - kWheelTravelTrim;
    const Vec3 delta = w.down * len_bt;
    int kind = mh.hit ? kGroundMesh : kGroundNone;
    float best_t = mh.hit ? mh.fraction * len_uu : 0.0f;
    Vec3 p(0.0f, 0.0f, 0.0f), n = mh.normal;
    if (mh.hit) p = (1.0f - mh.fraction) * w.from + mh.fraction * (w.from + delta);
    // Other cars within reach of the ray: the nearest hitbox hit (CastWheelRayAgainstCars, :329-418), taken when
    // nearer than the mesh's.  A car that gives no contact response (demolished this tick) voids the wheel's contact.
    int car_hit = -1;
    bool no_response = false;
    {
        float car_t = 0.0f;
        Vec3 car_p(0.0f, 0.0f, 0.0f), car_n(0.0f, 0.0f, 0.0f);
        for (int j = 0; j < count; ++j) {
            if (j == self || ((others_near >> (j + 1)) & 1u) == 0u || cars[j].is_demoed) continue;
            const CarConfig& oc = cfgs[j];
            const float reach = 0.5f * (Abs(oc.hitbox_size.x) + Abs(oc.hitbox_size.y) + Abs(oc.hitbox_size.z)) +
                                Abs(oc.hitbox_offset.x) + Abs(oc.hitbox_offset.y) + Abs(oc.hitbox_offset.z);
            const float rr = reach * (1.0f / 50.0f) + len_bt + 1.0e-2f;
            if (LengthSq(cars[j].pos_bt - w.from) > rr * rr) continue;
            float t = 0.0f;
            Vec3 hp, hn;
            if (!WheelRayCarHit(w.from, w.down, len_uu, len_bt, cars[j], oc, t, hp, hn)) continue;
            if (min_hit_sq > 0.0f && LengthSq(hp - w.from) < min_hit_sq) continue;
            if (car_hit < 0 || t < car_t) {
                car_hit = j;
                car_t = t;
                car_p = hp;
                car_n = hn;
                no_response = cars[j].demo_no_contact_response != 0;
            }
        }
        if (car_hit >= 0 && (kind == kGroundNone || car_t < best_t)) {
            kind = kGroundCar;
            best_t = car_t;
            p = car_p;
            n = car_n;
        }
    }
    if (others_near & 1u) {
        Vec3 bn, bp;
        const float bf = WheelRayBall(w, ball, ball_radius, len_bt, bn, bp);
        if (bf >= 0.0f) {
            const float t = bf * len_uu;
            const bool too_close = min_hit_sq > 0.0f && LengthSq(bp - w.from) < min_hit_sq;
            if (!(t < 0.0f || t > len_uu) && !too_close && (kind == kGroundNone || t < best_t)) {
                kind = kGroundBall;
                best_t = t;
                p = bp;
                n = bn;
            }
        }
    }
    if (kind == kGroundNone || (kind == kGroundCar && no_response)) {
        WheelMiss(w);
        return;
    }
    const int32_t mesh_id = kind == kGroundMesh ? static_cast<int32_t>((mesh.tris[mh.tri].flags >> 8) & 0xffu) : -1;
    WheelHit(car, w, p, n, static_cast<uint8_t>(kind), kind == kGroundMesh, mesh_id);
    w.ground_car = static_cast<int8_t>(kind == kGroundCar ? car_hit : -1);
}

// A wheel resting on another car pushes it (WheelRayPhase, raycast_wheels_and_contact_prep.hpp:2197-2219): when
// sunk below its threshold the wheel's pushback is a quarter of the two-body contact impulse that a Bullet solver
// step would give (ResolveCarCarWheelPushbackMagnitudeBtInputs, host_validation_and_contact_debug.hpp:461-484),
// applied in full to the other car at the hit, and every such wheel also gives the other car a unit impulse along
// -normal there.
RG_HD void WheelPushCar(const CarState& car, const CarKin& k, float mass, WheelRay& w, CarState& other,
                        const CarConfig& ocfg) {
    constexpr float kSolverDt = 1.0f / 60.0f;
    const Vec3 n = w.normal, cp = w.hit_point;
    const float thresh = (w.rest * (1.0f / 50.0f) + w.radius * (1.0f / 50.0f)) - 0.05f;
    Vec3 o0, o1, o2;
    if (other.restored_inv_inertia_valid) {
        o0 = other.restored_inv_inertia_bt[0];
        o1 = other.restored_inv_inertia_bt[1];
        o2 = other.restored_inv_inertia_bt[2];
    } else {
        InvInertiaWorldRows(other.tick_start_basis_forward, other.tick_start_basis_right, other.tick_start_basis_up,
                            InvInertiaLocalBt(ocfg, mass), o0, o1, o2);
    }
    w.pushback = 0.0f;
    if (w.trace_len < thresh) {
        const float depth = w.trace_len - thresh;
        const Vec3 v1 = car.vel_bt + Cross(car.ang_vel, cp - car.pos_bt);
        const Vec3 v2 = other.vel_bt + Cross(other.ang_vel, cp - other.pos_bt);
        const float rel_vel = Dot(n, v1 - v2);
        const auto denom = [&](const Vec3& pos, const Vec3& r0, const Vec3& r1, const Vec3& r2) {
            const Vec3 rel = cp - pos;
            const Vec3 c0 = Cross(rel, n);
            const Vec3 ang(r0.x * c0.x + r1.x * c0.y + r2.x * c0.z, r0.y * c0.x + r1.y * c0.y + r2.y * c0.z,
                           r0.z * c0.x + r1.z * c0.y + r2.z * c0.z);
            return 1.0f / mass + Dot(n, Cross(ang, rel));
        };
        const float d0 = denom(car.pos_bt, k.inv_row0, k.inv_row1, k.inv_row2);
        const float d1 = denom(other.pos_bt, o0, o1, o2);
        float impulse = 0.0f;
        if (!(d0 + d1 <= 0.0f)) {
            const float jac = 1.0f / (d0 + d1);
            const float sum = (0.2f * -depth / kSolverDt) * jac + -rel_vel * jac;
            impulse = sum > 0.0f ? sum : 0.0f;
        }
        w.pushback = impulse / 4.0f;
    }
    const auto apply = [&](const Vec3& impulse) {
        SetCarVelBt(other, other.vel_bt + impulse * (1.0f / mass));
        other.vel_touched_solve_stage = 1;
        other.ang_vel += MulRows(o0, o1, o2, Cross(cp - other.pos_bt, impulse));
    };
    if (w.pushback > 0.0f) apply(-n * (w.pushback * 4.0f));
    apply(-n);
}

// The right column of a wheel's frame turned by the steer angle about the car's up axis
// (ComputeWheelSteeredRightCol, :1501-1536).
RG_HD Vec3 SteeredRight(const WheelRay& w, float steer) {
    const Vec3 up = -w.down;
    const Vec3 right = w.axle;
    const float half = steer * 0.5f;
    const float s = (half == 0.0f && LengthSq(up) > 0.0f) ? half : (GameSinf(half) / Length(up));
    const float c = half == 0.0f ? 1.0f : GameCosf(half);
    const Quat q(c, up.x * s, up.y * s, up.z * s);
    Vec3 c0, c1, c2;
    GameQuatToColumns(q, c0, c1, c2);
    const Vec3 nr = -right;
    return Vec3(c0.x * nr.x + c1.x * nr.y + c2.x * nr.z, c0.y * nr.x + c1.y * nr.y + c2.y * nr.z,
                c0.z * nr.x + c1.z * nr.y + c2.z * nr.z);
}

struct DriveInputs {
    float steer;
    float engine;
    float brake;
    float throttle;
    float friction_scale;
    Vec3 front_right, back_right;
};

// Per-wheel drive set-up (WheelDriveSetupPhase, :1697-1830): the runtime record's steer/engine/brake, the slip
// friction curves, and the rolling friction for the friction impulse.
RG_HD void WheelDriveSetup(const CarState& car, const CarConfig& cfg, const DriveInputs& d, WheelRay& w, int i,
                           WheelRuntimeState& ws, float dt) {
    float lat = ws.lat_friction[i];
    float lon = ws.long_friction[i];
    ws.engine_force[i] = d.engine;
    ws.brake[i] = d.brake;
    ws.steer_angle[i] = w.front ? d.steer : 0.0f;
    Vec3 right_col(0.0f, 0.0f, 0.0f);
    if (w.contact) {
        right_col = w.front ? d.front_right : d.back_right;
        const Vec3 wv = (Cross(car.ang_vel, w.from - car.pos_bt) + car.vel_bt) * 50.0f;
        const Vec3 long_dir = Cross(right_col, w.normal);
        const float side_speed = Abs(Dot(wv, right_col));
        const float long_speed = Abs(Dot(wv, long_dir));
        float slip = 0.0f;
        if (side_speed > 5.0f) slip = side_speed / (long_speed + side_speed);
        lat = cfg.three_wheels ? Piecewise({0.0f, 1.0f}, {0.30f, 0.25f}, slip) : LateralFrictionAt(slip);
        lon = 1.0f;
        if (car.handbrake_val > 0.0f) {
            lat *= (0.1f - 1.0f) * car.handbrake_val + 1.0f;
            lon *= (HandbrakeLongFrictionAt(slip) - 1.0f) * car.handbrake_val + 1.0f;
        } else {
            lon = 1.0f;
        }
        const bool sticky = d.throttle != 0.0f && w.on_world;
        if (!sticky) {
            const float k = NonStickyFrictionAt(w.normal.z);
            lat *= k;
            lon *= k;
        }
    }
    ws.lat_friction[i] = lat;
    ws.long_friction[i] = lon;
    if (!w.contact) return;
    const Vec3 c = Cross(w.normal, right_col);
    const Vec3 fwd = c * (1.0f / sqrtf(LengthSq(c)));
    float roll = 0.0f;
    if (d.engine == 0.0f) {
        if (d.brake != 0.0f) {
            const Vec3 rel = w.hit_point - car.pos_bt;
            const Vec3 cv = car.vel_bt + Cross(car.ang_vel, rel);
            float along = Dot(cv, fwd);
            if (dt > (1.0f / 80.0f)) {
                const float th = -(1.0f / (dt * 150.0f)) + 0.8f;
                if (Abs(along) < th) along = 0.0f;
            }
            roll = Clamp(-along * 113.73963f, -d.brake, d.brake) * (dt * d.friction_scale);
        }
    } else {
        roll = -d.engine * dt;
    }
    w.fric_axle = right_col;
    w.fric_fwd = fwd;
    w.fric_roll = roll;
    w.fric_lat = lat;
    w.fric_long = lon;
    w.fric_scale = d.friction_scale;
}

// An impulse at a point relative to the car's center: velocity change and angular velocity change through the frame's
// Bullet rows (WheelImpulseAtPointBt, :1083-1097).
struct WheelImpulse {
    Vec3 dv, dw;
    uint8_t on;
};

RG_HD WheelImpulse ImpulseAt(const CarKin& k, float mass, const Vec3& j, const Vec3& rel) {
    WheelImpulse o;
    o.dv = j * (1.0f / mass);
    o.dw = KinInvInertia(k, Cross(rel, j));
    o.on = 1;
    return o;
}

RG_HD void AddWheelImpulses(CarState& car, const WheelImpulse* o) {
    Vec3 v = car.vel_bt, w = car.ang_vel;
    bool any = false;
    for (int k = 0; k < 4; ++k) {
        const int i = k ^ 1;
        if (!o[i].on) continue;
        v = v + o[i].dv;
        w += o[i].dw;
        any = true;
    }
    if (any) {
        SetCarVelBt(car, v);
        car.vel_touched_solve_stage = 1;
        car.ang_vel = w;
    }
}

// Suspension spring-damper impulses, then the tyre friction impulses against the updated velocity
// (ApplyWheelSuspensionAndFriction, :1475-1499; WheelSuspensionPhase :1116-1143; WheelFrictionPhase :1220-1254 with
// ResolveSingleBilateralAgainstFixedBt :1197-1218).
RG_HD void SuspensionAndFriction(CarState& car, const CarConfig& cfg, const CarKin& k, float mass, float dt,
                                 WheelRay* w) {
    if (!(w[0].contact || w[1].contact || w[2].contact || w[3].contact)) return;
    WheelImpulse out[4];
    for (int i = 0; i < 4; ++i) {
        out[i].on = 0;
        if (!w[i].contact) continue;
        const float spring = (w[i].rest * (1.0f / 50.0f) - w[i].susp_len) * kSuspensionStiffness * w[i].inv_cos;
        const float damp = w[i].rel_vel_uu < 0.0f ? kSuspensionCompression : kSuspensionRelaxation;
        float f = (spring - damp * w[i].rel_vel_bt) * w[i].force_scale;
        if (f < 0.0f) f = 0.0f;
        if (f * 50.0f > 0.0f) {
            const Vec3 rel = w[i].hit_point - car.pos_bt;
            const float j = (f * dt) + w[i].pushback;
            out[i] = ImpulseAt(k, mass, w[i].normal * j, rel);
        }
    }
    AddWheelImpulses(car, out);
    const Vec3 com = car.pos_bt, lin = car.vel_bt, ang = car.ang_vel;
    const float inv_mass = mass != 0.0f ? 1.0f / mass : 0.0f;
    for (int i = 0; i < 4; ++i) {
        out[i].on = 0;
        if (!w[i].contact) continue;
        // Side impulse of a bilateral constraint against a fixed body along the axle (ERP 0.2).
        const Vec3 rp = w[i].hit_point - com;
        const Vec3 pv = lin + Cross(ang, rp);
        const Vec3 rxn = Cross(rp, w[i].fric_axle);
        const Vec3 aj(Dot(k.fwd, rxn), Dot(k.right, rxn), Dot(k.up, rxn));
        const Vec3 minv = Mul3(k.inv_local_bt, aj);
        const float diag = inv_mass + Dot(minv, aj);
        const float along = Dot(w[i].fric_axle, pv);
        float side_j = 0.0f;
        if (!(diag <= 0.0f)) side_j = -0.2f * along * (1.0f / diag);
        const float roll = w[i].fric_roll * w[i].fric_long;
        const float side = (side_j * (dt * w[i].fric_scale)) * w[i].fric_lat;
        const Vec3 j(roll * w[i].fric_fwd.x + side * w[i].fric_axle.x, roll * w[i].fric_fwd.y + side * w[i].fric_axle.y,
                     roll * w[i].fric_fwd.z + side * w[i].fric_axle.z);
        if (LengthSq(j) == 0.0f) continue;
        const Vec3 rel = w[i].hit_point - car.pos_bt;
        Vec3 rel_plane = rel - k.up * Dot(k.up, rel);
        if (cfg.three_wheels) rel_plane = rel_plane - k.up * (4.0f * 0.02f);
        out[i] = ImpulseAt(k, mass, j, rel_plane);
    }
    AddWheelImpulses(car, out);
}

// ------------------------------------------------------------------------------------------------------------
// Forward speed of the car's replicated snapshot: the velocity capped and quantized as the game sends it, along the
// forward axis of the quantized rotation (FreshCarRbSnapshot / SnapshotForwardSpeed, :912-1015).
RG_HD float SnapshotForwardSpeedNow(const CarState& car, const CarKin& k, float max_speed) {
    Vec3 v = car.vel_bt * 50.0f;
    const float len2 = (v.y * v.y + v.x * v.x) + v.z * v.z;
    if (max_speed > 0.0f && max_speed * max_speed < len2) {
        const float inv = 1.0f / sqrtf(len2);
        v = Vec3(max_speed * (inv * v.x), max_speed * (inv * v.y), max_speed * (inv * v.z));
    }
    const Vec3 sv(QuantTruncOffset(v.x, 100.0f), QuantTruncOffset(v.y, 100.0f), QuantTruncOffset(v.z, 100.0f));
    const Quat q = GameQuantizeQuat(GameMatrixToQuat(k.fwd, k.right, k.up));
    const float w2 = q.w + q.w;
    const float s = q.w * q.w - ((q.y * q.y + q.x * q.x) + q.z * q.z);
    float dq = (1.0f * q.x + 0.0f * q.y) + 0.0f * q.z;
    dq = dq + dq;
    const Vec3 f(((0.0f * q.y - 0.0f * q.z) * w2 + s * 1.0f) + dq * q.x,
                 ((1.0f * q.z - 0.0f * q.x) * w2 + s * 0.0f) + dq * q.y,
                 ((0.0f * q.x - 1.0f * q.y) * w2 + s * 0.0f) + dq * q.z);
    return (f.x * sv.x + f.y * sv.y) + f.z * sv.z;
}

RG_HD Vec3 NormalizeOrZero(const Vec3& sum) {
    const float l2 = LengthSq(sum);
    if (l2 >= 1e-8f) return sum * (1.0f / sqrtf(l2));
    return Vec3(0.0f, 0.0f, 0.0f);
}

// Auto-roll (ApplyAutoRollReference, :1021-1076): a grounded-ish car with throttle is pushed onto the surface and
// torqued toward it.  The force and the game-axis torque go first in the queue; the angular acceleration form is
// queued only for a car without an authoritative basis.
RG_HD void AutoRoll(CarState& car, const CarConfig& cfg, const CarParams& m, CarKin& k, int contacts,
                    const Vec3& normal_sum) {
    const Vec3 g_up = contacts > 0 ? NormalizeOrZero(normal_sum) : car.world_contact.contact_normal;
    const Vec3 g_down = -g_up;
    if (!car.basis_authoritative) {
        const Vec3 cr = Cross(g_up, k.fwd);
        const Vec3 cf = Cross(g_down, cr);
        const float rf = 1.0f - Clamp(Dot(k.right, cr), 0.0f, 1.0f);
        const float ff = 1.0f - Clamp(Dot(k.fwd, cf), 0.0f, 1.0f);
        const Vec3 tr = k.fwd * (Dot(k.right, g_up) >= 0.0f ? -1.0f : 1.0f);
        const Vec3 tf = k.right * (Dot(k.fwd, g_up) >= 0.0f ? 1.0f : -1.0f);
        const Vec3 torque = (tf * ff + tr * rf) * 80.0f;
        Vec3 i0, i1, i2, w0, w1, w2;
        InvInertiaWorldRowsUu(car, cfg, m.car_mass, i0, i1, i2);
        InverseRows(i0, i1, i2, w0, w1, w2);
        k.q.ang_accel = k.q.ang_accel + MulRows(w0, w1, w2, torque);
    }
    const Vec3 force = ((g_down * 100.0f) * (1.0f / 50.0f)) * m.car_mass;
    Vec3 F, R, U;
    UeQuatAxes(car.rot, F, R, U);
    const Vec3& G = g_up;
    const float crx = F.z * G.y - F.y * G.z;
    const float cry = F.x * G.z - F.z * G.x;
    const float crz = F.y * G.x - F.x * G.y;
    const float r_dot = Clamp((crx * R.x + R.y * cry) + R.z * crz, 0.0f, 1.0f);
    const float nx = -G.x, ny = -G.y, nz = -G.z;
    const float f_dot = Clamp(((ny * crz - nz * cry) * F.x + (crx * nz - nx * crz) * F.y) + (nx * cry - crx * ny) * F.z,
                              0.0f, 1.0f);
    const Vec3 f_axis = 0.0f <= (R.y * G.y + R.x * G.x) + R.z * G.z ? -F : F;
    const Vec3 r_axis = (F.y * G.y + F.x * G.x) + F.z * G.z < 0.0f ? -R : R;
    const float rf = (1.0f - r_dot) * 80.0f;
    const float ff = (1.0f - f_dot) * 80.0f;
    const Vec3 t(ff * r_axis.x + rf * f_axis.x, ff * r_axis.y + rf * f_axis.y, ff * r_axis.z + rf * f_axis.z);
    Vec3 w0, w1, w2;
    InverseRows(k.inv_row0, k.inv_row1, k.inv_row2, w0, w1, w2);
    QueueFirst(k.q, force, MulRows(w0, w1, w2, t));
}

// What the vehicle update needs from the rest of the arena.
struct VehicleWorld {
    const ArenaMeshView* mesh;
    BallState* ball;
    float ball_radius;
    CarState* cars;           // the arena's cars (a wheel on another car pushes it)
    const CarConfig* cfgs;
    int count;
    int self;
    uint32_t others_near;  // bodies the car's rays are cast against (bit 0 ball, 1 + j car j)
    int ray_lanes = 1;     // lanes of a wavefront stepping this arena together (1, 2 or 4)
    int ray_sub = 0;       // this lane's index among them
};

#if defined(__HIP_DEVICE_COMPILE__)
// Ray hit `h` of lane `src` of this lane's group of `width` lanes (the mesh-hit fields CastWheel reads).
__device__ inline RayHit ShuffleRayHit(const RayHit& h, int src, int width) {
    RayHit o;
    o.fraction = __shfl(h.fraction, src, width);
    o.normal = Vec3(__shfl(h.normal.x, src, width), __shfl(h.normal.y, src, width), __shfl(h.normal.z, src, width));
    o.tri = __shfl(h.tri, src, width);
    o.hit = static_cast<uint8_t>(__shfl(static_cast<int>(h.hit), src, width));
    o.min_fraction_sq = __shfl(h.min_fraction_sq, src, width);
    return o;
}
#endif

// The four wheel segments' nearest mesh hits, each CastArenaRay's.  When `lanes` (2 or 4) lanes of a wavefront
// step the same arena with the same data, lane `sub` walks rays sub (and sub + 2 for two lanes) and the group's
// lanes exchange the hits: the serial chain of tree loads per arena is a half or a quarter as long.
RG_HD void CastWheelRays(const ArenaMeshView& mesh, const WheelRay* w, const Vec3* delta, const float* min_sq,
                         int lanes, int sub, RayHit* out) {
#if defined(__HIP_DEVICE_COMPILE__)
    if (lanes == 2 || lanes == 4) {
        const int r0 = sub, r1 = sub + 2;
        Vec3 f0 = w[0].from, d0 = delta[0], f1 = w[2].from, d1 = delta[2];
        float m0 = min_sq[0], m1 = min_sq[2];
        for (int q = 1; q < 4; ++q) {
            if (q == r0) {
                f0 = w[q].from;
                d0 = delta[q];
                m0 = min_sq[q];
            }
        }
        if (r1 == 3) {
            f1 = w[3].from;
            d1 = delta[3];
            m1 = min_sq[3];
        }
        const RayHit h0 = CastArenaRay(mesh, f0, d0, m0);
        const RayHit h1 = lanes == 2 ? CastArenaRay(mesh, f1, d1, m1) : h0;
        for (int r = 0; r < 4; ++r) out[r] = ShuffleRayHit(r < lanes ? h0 : h1, r % lanes, lanes);
        return;
    }
#else
    (void)lanes;
    (void)sub;
#endif
    for (int r = 0; r < 4; ++r) out[r] = CastArenaRay(mesh, w[r].from, delta[r], min_sq[r]);
}

// The vehicle update of one car (PreTickCar after its pre-physics part, env_step.hpp:3115-3510, with
// PreTickVehicleDynamics, raycast_wheels_and_contact_prep.hpp:2618-2975).  Controls are the step's raw controls.
RG_HD void CarVehicleUpdate(CarState& car, const CarControls& raw, const CarConfig& cfg, const CarParams& m,
                            WheelRuntimeState& ws, const VehicleWorld& world, uint32_t& unsupported) {
    const float dt = m.tick_dt;
    car.demo_no_contact_response = car.is_demoed ? 1 : 0;
    if (car.is_demoed) return;
    if (cfg.three_wheels) unsupported |= kUnsupportedThreeWheels;
    if (car.set_wheel_normals_mask != 0) unsupported |= kUnsupportedSetWheelNormals;
    float raw_pitch = 2.0f;
    const CarControls c = PreTickControls(car, raw, raw_pitch);
    const uint8_t pre = car.cal_pre_flags;
    CarKin k;
    MakeCarKin(car, cfg, m.car_mass, k);
    ws.pre_solve_rot = car.rot;
    ws.pre_solve_basis_forward = car.tick_start_basis_forward;
    ws.pre_solve_basis_right = car.tick_start_basis_right;
    ws.pre_solve_basis_up = car.tick_start_basis_up;
    ws.pre_solve_basis_authoritative = 1;
    const bool set_v3_tick = (pre & kPreSupersonicFromSet) != 0;
    const uint8_t flipping_at_start = (pre & kPreFlippingAtStart) ? 1 : 0;
    const bool flip_started = (pre & kPreFlipStarted) != 0;
    const bool air_jump_started = (pre & kPreGroundJumpStarted) != 0;
    const uint8_t was_jumping = (pre & kPreWasJumping) ? 1 : 0;
    const uint8_t air_jump_used_before = (pre & kPreHadAirJump) ? 1 : 0;
    const bool flip_vel_changed = (pre & kPreFlipMovedCar) != 0;
    (void)flip_vel_changed;
    const int prev_contacts = (car.wheels_with_contact[0] != 0) + (car.wheels_with_contact[1] != 0) +
                              (car.wheels_with_contact[2] != 0) + (car.wheels_with_contact[3] != 0);
    const bool roll_prev = car.tick_count_since_update != 0;
    Vec3 roll_sum = roll_prev ? ws.cal_prev_wheel_normal_sum : Vec3(0.0f, 0.0f, 0.0f);
    int roll_world = roll_prev ? ws.cal_prev_world_wheel_contacts : 0;

    // --- vehicle dynamics: wheel rays, drive, sticky force.
    car.world_contact.wheel_ball_mask = 0;
    const uint8_t sticky_prev_world = ws.cal_prev_world_wheel_contacts;
    const Vec3 sticky_prev_sum = ws.cal_prev_wheel_normal_sum;
    float front_share, back_share;
    WheelMassShares(cfg, m.car_mass, front_share, back_share);
    const float min_hit = (cfg.hitbox_size.z * 0.5f - 2.25f) * (1.0f / 50.0f);
    const float min_hit_sq = min_hit * min_hit;
    const float front_bone = WheelBoneZ(cfg, true);
    const float back_bone = cfg.back_suspension_rest == cfg.front_suspension_rest ? front_bone : WheelBoneZ(cfg, false);
    // The four rays against the meshes first (a ray's mesh hit depends only on the car's own pose).
    WheelRay w[4];
    Vec3 seg_delta[4];
    float seg_min_sq[4];
    for (int i = 0; i < 4; ++i) {
        InitWheel(w[i], car, cfg, k, i, i < 2 ? front_bone : back_bone);
        WheelMeshSegment(w[i], min_hit_sq, seg_delta[i], seg_min_sq[i]);
    }
    RayHit mesh_hit[4];
    CastWheelRays(*world.mesh, w, seg_delta, seg_min_sq, world.ray_lanes, world.ray_sub, mesh_hit);
    for (int kk = 0; kk < 4; ++kk) {
        const int i = kk ^ 1;
        w[i].force_scale = w[i].front ? front_share : back_share;
        w[i].pushback = ws.extra_pushback[i];
        CastWheel(w[i], mesh_hit[i], car, world.cars, world.cfgs, world.count, world.self, *world.mesh,
                  *world.ball, world.ball_radius, world.others_near, min_hit_sq, unsupported);
        if (w[i].contact) w[i].pushback = 0.0f;
        if (w[i].on_world) WheelPushback(car, m.car_mass, k, w[i]);
        if (w[i].contact && !w[i].on_world && w[i].ground == kGroundCar) {
            WheelPushCar(car, k, m.car_mass, w[i], world.cars[w[i].ground_car], world.cfgs[w[i].ground_car]);
        }
        if (w[i].contact && !w[i].on_world && w[i].ground == kGroundBall) {
            // A wheel on the ball pushes it with a unit impulse along -normal.
            BallState& b = *world.ball;
            b.sleeping |= 4u;
            const Vec3 push = -w[i].normal;
            SetBallVelBt(b, b.vel_bt + push * (1.0f / m.ball_mass));
            StepParams bp{};
            bp.ball_radius = world.ball_radius;
            bp.ball_mass = m.ball_mass;
            b.ang_vel += BallInvInertiaBt(b, bp, Cross(w[i].hit_point - b.pos_bt, push));
        }
        ws.extra_pushback[i] = w[i].pushback;
    }
    // Wheel contacts summed (CombineWheelContacts, :2240-2285).
    int n_contact = 0;
    bool world_contact_any = false;
    Vec3 normal_sum(0.0f, 0.0f, 0.0f), world_sum(0.0f, 0.0f, 0.0f);
    uint8_t world_n = 0;
    for (int kk = 0; kk < 4; ++kk) {
        const int i = kk ^ 1;
        car.wheels_with_contact[i] = w[i].contact;
        car.world_contact.wheel_ball_mask |=
            static_cast<uint8_t>((w[i].contact && w[i].ground == kGroundBall) << i);
        n_contact += w[i].contact ? 1 : 0;
        world_contact_any = world_contact_any || w[i].on_world;
        if (w[i].contact) normal_sum += w[i].normal;
    }
    car.is_on_ground = n_contact >= 3 ? 1 : 0;
    for (int kk = 0; kk < 4; ++kk) {
        const int i = kk ^ 1;
        if (w[i].contact && w[i].on_world) {
            world_sum += w[i].normal;
            ++world_n;
        }
    }
    ws.cal_prev_wheel_normal_sum = world_sum;
    ws.cal_prev_world_wheel_contacts = world_n;
    (void)world_contact_any;

    const float snapshot_speed_abs = Abs(car.cal_pre_forward_speed_uu);
    const float fwd_speed = SnapshotForwardSpeedNow(car, k, m.car_max_speed);
    const float fwd_speed_abs = Abs(fwd_speed);
    car.handbrake_val = c.handbrake ? Clamp(car.handbrake_val + 5.0f * dt, 0.0f, 1.0f)
                                    : Clamp(car.handbrake_val - 2.0f * dt, 0.0f, 1.0f);
    const float throttle = c.throttle;
    float brake = 0.0f;
    float engine = car.driving_disabled ? 0.0f : throttle;
    if (!c.handbrake && !car.driving_disabled) {
        if (Abs(throttle) > 0.01f) {
            const float st = static_cast<float>((throttle > 0.0f) - (throttle < 0.0f));
            const float sv = static_cast<float>((fwd_speed > 0.0f) - (fwd_speed < 0.0f));
            if (fwd_speed_abs > 25.0f && st != sv) {
                brake = 1.0f;
                if (fwd_speed_abs > 0.01f) engine = 0.0f;
            }
        } else {
            engine = 0.0f;
            brake = fwd_speed_abs < 25.0f ? 1.0f : 0.15f;
        }
    }
    float engine_uu = (DriveTorqueAtSpeed(fwd_speed) * (engine * 288000.0f)) * 0.25f;
    if (n_contact < 3) engine_uu *= 0.25f;
    DriveInputs d;
    d.engine = engine_uu * (1.0f / 50.0f);
    d.brake = ((brake * 10500.0f) * 0.25f) * (1.0f / 50.0f);
    {
        float deg = SteerDegAtSpeed(cfg.steer_angle_curve_deg, fwd_speed_abs);
        if (car.handbrake_val > 0.0f) {
            deg += (PowerslideSteerDegAtSpeed(cfg.powerslide_steer_angle_curve_deg, fwd_speed_abs) - deg) *
                   car.handbrake_val;
        }
        d.steer = (deg * c.steer) * 0.01745329238474369f;
    }
    d.throttle = throttle;
    d.friction_scale = m.car_mass / 3.0f;
    d.front_right = Vec3(0.0f, 0.0f, 0.0f);
    d.back_right = Vec3(0.0f, 0.0f, 0.0f);
    const bool front_on = w[0].contact || w[1].contact;
    const bool back_on = w[2].contact || w[3].contact;
    if (front_on) d.front_right = SteeredRight(w[0], d.steer);
    if (back_on) {
        const bool plus_zero = d.steer == 0.0f && (FloatBits(d.steer) >> 31) == 0u;
        d.back_right = (front_on && plus_zero) ? d.front_right : SteeredRight(w[2], 0.0f);
    }
    for (int i = 0; i < 4; ++i) WheelDriveSetup(car, cfg, d, w[i], i, ws, dt);
    // Sticky force from last tick's world wheel normals (this tick's on a freshly set state).
    {
        Vec3 sum = sticky_prev_sum;
        bool gate = sticky_prev_world > 0;
        if (car.tick_count_since_update == 0) {
            sum = world_sum;
            gate = world_n > 0;
        }
        if (gate) {
            const Vec3 up = NormalizeOrZero(sum);
            const bool full = (c.throttle != 0.0f) || (snapshot_speed_abs > 25.0f);
            const float base = cfg.three_wheels ? 0.0f : 0.5f;
            const float full_scale = cfg.three_wheels ? 1.5f : base + 1.0f;
            const float target = full ? full_scale : base;
            const float scale = Abs(up.z) * (base - target) + target;
            const float accel = (m.gravity_z * scale) * (m.gravity_z > 0.0f ? -1.0f : 1.0f);
            QueueCarForce(k.q, ((accel * up) * (1.0f / 50.0f)) * m.car_mass);
        }
    }

    // --- after the wheels: surface history, air control, flip torque, jump, auto-roll, suspension, boost.
    bool world_wheel = false;
    for (int i = 0; i < 4; ++i) world_wheel = world_wheel || (w[i].contact && w[i].on_world);
    if (world_wheel) {
        car.last_world_surface_contact = car.tick_count_since_update;
        if (car.last_ball_contact_start != ~0ULL && car.first_world_after_last_ball == ~0ULL &&
            car.tick_count_since_update > car.last_ball_contact_start) {
            car.first_world_after_last_ball = car.tick_count_since_update;
        }
    }
    if (HasFlipOrJump(car) && !car.flip_or_jump_last_tick) car.last_flip_or_jump_start = car.tick_count_since_update;
    const int air_gate = (car.tick_count_since_update != 0 || set_v3_tick)
                             ? prev_contacts
                             : (prev_contacts > n_contact ? prev_contacts : n_contact);
    const int landing_prev = (car.tick_count_since_update == 0 && !set_v3_tick) ? MaxI(n_contact, prev_contacts)
                                                                               : prev_contacts;
    const uint8_t flipping_before_air = car.is_flipping;
    const CarFrame frame = KinFrame(k);
    if (air_gate < 3) {
        Vec3 dodge_torque(0.0f, 0.0f, 0.0f);
        CarApplyFlipTorque(car, c, cfg, m, frame, raw_pitch, &k.q, &dodge_torque);
        if (!car.driving_disabled) {
            CarApplyAirControl(car, c, cfg, m, frame, landing_prev == 0, landing_prev == 0, &k.q);
        }
        if (dodge_torque.x != 0.0f || dodge_torque.y != 0.0f || dodge_torque.z != 0.0f) {
            QueueCarTorque(k.q, dodge_torque);
        }
    } else {
        car.is_flipping = 0;
        if (landing_prev == 0 && !car.driving_disabled) CarApplyAirControl(car, c, cfg, m, frame, true, true, &k.q);
    }
    const bool flip_ended = air_gate < 3 && flipping_before_air && !car.is_flipping;
    CarUpdateJumpState(car, c, m, &k.q);
    if ((car.is_jumping && !was_jumping) || (!air_jump_used_before && (car.has_double_jumped || car.has_flipped))) {
        ws.cal_game_jumped = 1;
    }
    if (!air_jump_used_before && (car.has_double_jumped || car.has_flipped)) ws.cal_game_double_jumped = 1;
    ws.cal_jump_inactive_time = car.is_jumping ? 0.0f : ws.cal_jump_inactive_time + dt;
    const bool jump_changed = air_jump_started || (was_jumping && !car.is_jumping);
    const bool flip_changed = flip_started || car.is_flipping != flipping_at_start;
    if (jump_changed || !car.just_respawned) {
        const float ticks = rintf(car.jump_time / dt) + 1.0f;
        car.jump_time = jump_changed ? dt : (ticks <= 1024.0f ? car.jump_time + dt : ticks * dt);
    }
    if (flip_changed || !car.just_respawned) car.flip_time = flip_changed ? dt : car.flip_time + dt;
    if (air_jump_started) {
        car.jump_hold_broken = 0;
    } else if (was_jumping && !c.jump) {
        car.jump_hold_broken = 1;
    }
    if (!roll_prev) {
        roll_sum = ws.cal_prev_wheel_normal_sum;
        roll_world = ws.cal_prev_world_wheel_contacts;
    }
    int roll_contacts = roll_world;
    bool roll_gate = roll_contacts < 4 && (roll_contacts > 0 || car.world_contact.has_contact);
    if (!roll_prev && prev_contacts > 0 && prev_contacts < 4) {
        const bool sticky = roll_world > 0;
        roll_gate = sticky || car.world_contact.has_contact;
        roll_contacts = sticky ? prev_contacts : 0;
    }
    if (c.throttle != 0.0f && !car.is_flipping && !flipping_at_start && !flip_ended && roll_gate && !cfg.three_wheels) {
        AutoRoll(car, cfg, m, k, roll_contacts, roll_sum);
    }
    car.set_wheel_normals_mask = 0;
    car.world_contact.has_contact = 0;
    if (!car.just_respawned) SuspensionAndFriction(car, cfg, k, m.car_mass, dt, w);
    CarApplyBoost(car, c, m, &k.q);
    if (k.q.has_jump_force) k.q.force += k.q.jump_force;
    ws.accumulated_force = k.q.force;
    ws.accumulated_ang_accel = k.q.ang_accel;
    ws.accumulated_torque_bt = k.q.torque_bt;
    ws.solver_time_step = dt;
    for (int i = 0; i < 4; ++i) {
        const bool on = w[i].contact && w[i].on_world;
        ws.event_world_ids[i] = on ? w[i].static_id : -1;
        ws.event_world_normals[i] = on ? w[i].normal : Vec3(0.0f, 0.0f, 0.0f);
    }
}

// ------------------------------------------------------------------------------------------------------------
// Car rigid body in the world step.

// What the world step keeps of a car from before its contacts (PreContactCar, arena_geometry.hpp:887-893).
struct CarPreContact {
    Quat rot;
    Vec3 vel, vel_bt, ang_vel;
    uint8_t is_demoed;
};

// External forces of the step as velocity impulses, fixed before the contacts (WorldStepPrestepCar, env_step.hpp:
// 3737-3787): gravity plus the queued force over the mass, the queued torque through the pre-solve inverse inertia.
RG_HD void CarPrestep(CarState& car, const CarConfig& cfg, const WheelRuntimeState& ws, const CarParams& m, float dt,
                      CarPreContact& pc) {
    const bool inactive = car.is_demoed || car.just_respawned;
    Vec3 ext_force(0.0f, 0.0f, 0.0f);
    if (!inactive) {
        const float inv_mass = 1.0f / m.car_mass;
        const Vec3 gravity(0.0f, 0.0f, (m.gravity_z * (1.0f / 50.0f)) * m.car_mass);
        ext_force = ((ws.accumulated_force + gravity) * inv_mass) * dt;
    }
    Vec3 ext_torque(0.0f, 0.0f, 0.0f);
    if (!inactive) {
        const Vec3& t = ws.accumulated_torque_bt;
        if (t.x != 0.0f || t.y != 0.0f || t.z != 0.0f) {
            Vec3 r0, r1, r2;
            if (car.restored_inv_inertia_valid) {
                r0 = car.restored_inv_inertia_bt[0];
                r1 = car.restored_inv_inertia_bt[1];
                r2 = car.restored_inv_inertia_bt[2];
            } else {
                Vec3 f = ws.pre_solve_basis_forward, r = ws.pre_solve_basis_right, u = ws.pre_solve_basis_up;
                if (!ws.pre_solve_basis_authoritative) QuatBasis(ws.pre_solve_rot, f, r, u);
                InvInertiaWorldRows(f, r, u, InvInertiaLocalBt(cfg, m.car_mass), r0, r1, r2);
            }
            ext_torque = Vec3(r0.x * t.x + r1.x * t.y + r2.x * t.z, r0.y * t.x + r1.y * t.y + r2.y * t.z,
                              r0.z * t.x + r1.z * t.y + r2.z * t.z) * dt;
        }
    }
    pc.rot = car.rot;
    pc.vel = car.vel;
    pc.vel_bt = car.vel_bt;
    pc.ang_vel = car.ang_vel;
    pc.is_demoed = car.is_demoed;
    car.ext_force_impulse_bt = ext_force;
    car.ext_force_impulse_bt_valid = 1;
    car.solve_delta_linear_bt = Vec3(0.0f, 0.0f, 0.0f);
    car.solve_delta_angular = Vec3(0.0f, 0.0f, 0.0f);
    car.ext_torque_impulse = ext_torque;
}

// The game's orientation integration on basis columns (GameIntegrateOrientationBasisColumns, math.hpp:1202-1227):
// the rotation quaternion of the basis advanced by the half-angle formula, normalized, and written back as columns.
RG_HD void IntegrateBasis(Vec3& fwd, Vec3& right, Vec3& up, Quat& q_out, const Vec3& w, float dt) {
    constexpr float kMaxAngle = 0.5f * Pi() * 0.5f;
    float angle = sqrtf((w.x * w.x + w.y * w.y) + w.z * w.z);
    if (angle * dt > kMaxAngle) angle = kMaxAngle / dt;
    float f;
    if (angle < 0.001f) {
        f = dt * 0.5f - ((((dt * dt) * dt) * 0.020833333333f) * angle) * angle;
    } else {
        f = GameSinf((angle * 0.5f) * dt) / angle;
    }
    const float ax = w.x * f, ay = w.y * f, az = w.z * f;
    const float c = GameCosf((angle * dt) * 0.5f);
    const Quat o = GameMatrixToQuat(fwd, right, up);
    const Quat p(((c * o.w - ax * o.x) - ay * o.y) - az * o.z, ((ax * o.w + c * o.x) + ay * o.z) - az * o.y,
                 ((c * o.y + ay * o.w) + az * o.x) - ax * o.z, ((c * o.z + az * o.w) + ax * o.y) - ay * o.x);
    q_out = GameQuatNormalize(p, -1.0f);
    GameQuatToColumns(q_out, fwd, right, up);
}

// After the contacts: the external impulses, then position and orientation (WorldStepFinishCar, env_step.hpp:
// 3789-3806; ApplyDeferredVelocityDeltas host_validation_and_contact_debug.hpp:986-1029; IntegrateCarTransformOnly).
RG_HD bool CarFinishSupported(const CarState& car) { return car.basis_authoritative != 0; }

RG_HD void CarFinish(CarState& car, const CarPreContact& pc, float dt) {
    if (pc.is_demoed) return;
    if (car.just_respawned) return;
    if (car.ext_force_impulse_bt_valid && !car.is_demoed) {
        SetCarVelBt(car, car.vel_bt + car.ext_force_impulse_bt);
        car.ang_vel += car.ext_torque_impulse;
    }
    if (!(car.vel_bt.x == 0.0f && car.vel_bt.y == 0.0f && car.vel_bt.z == 0.0f)) {
        SetCarPosBt(car, car.pos_bt + car.vel_bt * dt);
    }
    IntegrateBasis(car.basis_forward, car.basis_right, car.basis_up, car.rot, car.ang_vel, dt);
}

// ------------------------------------------------------------------------------------------------------------
// Broadphase proxy of a car (ComputeBulletBodyProxy, car branch, arena_geometry.hpp:3910-4033): the hitbox's box now
// (insertion box) and the union of the compound's box now and after one predicted step (swept box), each grown by
// the contact threshold.

struct CarProxyPose {
    Vec3 axis[3];   // hitbox axes now
    Vec3 center;    // hitbox center now (bt)
    Vec3 half;      // hitbox half extents (bt)
};

RG_HD void CarBroadphaseProxy(const CarState& car, const CarConfig& cfg, const Vec3& pred_vel_bt,
                              const Vec3& pred_ang_vel, float dt, BodyProxy& px, CarProxyPose& pose) {
    constexpr float kThreshold = 0.02f;
    const Vec3 half = HitboxHalfBt(cfg);
    const Vec3 child = cfg.hitbox_offset * (1.0f / 50.0f);
    Vec3 ax = car.basis_forward, ay = car.basis_right, az = car.basis_up;
    Vec3 px_ = ax, py_ = ay, pz_ = az;
    Quat unused;
    IntegrateBasis(px_, py_, pz_, unused, pred_ang_vel, dt);
    const Vec3 origin = car.pos_bt;
    const Vec3 pred_origin = origin + pred_vel_bt * dt;
    const auto center_of = [](const Vec3& o, const Vec3& a, const Vec3& b, const Vec3& c, const Vec3& l) {
        return Vec3((a.x * l.x + b.x * l.y) + c.x * l.z + o.x, (a.y * l.x + b.y * l.y) + c.y * l.z + o.y,
                    (a.z * l.x + b.z * l.y) + c.z * l.z + o.z);
    };
    const auto extent_of = [](const Vec3& a, const Vec3& b, const Vec3& c, const Vec3& h) {
        return Vec3(Abs(a.x) * h.x + Abs(b.x) * h.y + Abs(c.x) * h.z, Abs(a.y) * h.x + Abs(b.y) * h.y + Abs(c.y) * h.z,
                    Abs(a.z) * h.x + Abs(b.z) * h.y + Abs(c.z) * h.z);
    };
    const Vec3 cn = center_of(origin, ax, ay, az, child);
    const Vec3 en = extent_of(ax, ay, az, half);
    px.child_min = Vec3(cn.x - en.x, cn.y - en.y, cn.z - en.z);
    px.child_max = Vec3(cn.x + en.x, cn.y + en.y, cn.z + en.z);
    const Vec3 lmin = child - half, lmax = child + half;
    const Vec3 lc = (lmax + lmin) * 0.5f, lh = (lmax - lmin) * 0.5f;
    const Vec3 on = center_of(origin, ax, ay, az, lc);
    const Vec3 op = center_of(pred_origin, px_, py_, pz_, lc);
    const Vec3 ean = extent_of(ax, ay, az, lh);
    const Vec3 eap = extent_of(px_, py_, pz_, lh);
    px.aabb_min = Vec3(Min(on.x - ean.x, op.x - eap.x) - kThreshold, Min(on.y - ean.y, op.y - eap.y) - kThreshold,
                       Min(on.z - ean.z, op.z - eap.z) - kThreshold);
    px.aabb_max = Vec3(Max(on.x + ean.x, op.x + eap.x) + kThreshold, Max(on.y + ean.y, op.y + eap.y) + kThreshold,
                       Max(on.z + ean.z, op.z + eap.z) + kThreshold);
    px.valid = 1;
    pose.axis[0] = ax;
    pose.axis[1] = ay;
    pose.axis[2] = az;
    pose.center = cn;
    pose.half = half;
}

}  // namespace core
}  // namespace rg
