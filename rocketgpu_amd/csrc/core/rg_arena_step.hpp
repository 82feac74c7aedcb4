// One arena tick with cars: the reference's StepEnvGeneric for one env and one tick
// (include/rocketgpu/detail/step/env_step.hpp:7048-7180), composed from the ball pieces of rg_ball_step.hpp, the car
// pre/post-tick of rg_car_phases.hpp and the vehicle update and car body of rg_car_tick.hpp:
//
//   1. env pre-tick: car visit order, respawn spots of due demolished cars       env_step.hpp:6652-6700
//   2. car pre-ticks in visit order, ball pre-tick, pad pre-tick                  env_step.hpp:6751-6900, :2004
//   3. vehicle updates in visit order                                             env_step.hpp:6176-6263
//   4. world step: ball activation, car external impulses, predictive rays, broadphase with every body and the
//      island-sorted manifold list, ball-mesh contacts, ball rows solve, integration of ball and cars
//                                                                                 env_step.hpp:4622-5623
//   5. car post-ticks in visit order                                              env_step.hpp:6902-6926
//   6. env tail: boost pads, driving lock, ball finish, goals, frame timers       env_step.hpp:6968-7045
//
// Contacts between bodies and car-mesh contacts are not computed here; a tick that would need one sets the arena's
// unsupported word (rg_car_tick.hpp lists the bits).
#pragma once

#include "rg_car_car.hpp"

namespace rg {
namespace core {

constexpr int kMaxArenaCars = kMaxBodies - 1;

// Per-tick working data of a car arena.
struct CarArenaTick {
    BallTickScratch ball;
    BodyProxy proxy[kMaxBodies];
    CarProxyPose pose[kMaxArenaCars];
    CarPreContact pre[kMaxArenaCars];
    Vec3 predict_vel[kMaxArenaCars];
    Vec3 predict_ang[kMaxArenaCars];
    ManifoldEntry entries[kMaxManifoldEntries];
    int entry_count;
    // Car-mesh contacts: each car's solved velocity change and split pseudo-velocities, and the rows of the car
    // being processed.
    Vec3 world_dlin[kMaxArenaCars], world_dang[kMaxArenaCars];
    Vec3 world_push[kMaxArenaCars], world_turn[kMaxArenaCars];
    uint8_t world_vel[kMaxArenaCars], world_split[kMaxArenaCars];
    CarWorldRows world_rows;
    // Car-car contacts: each car's tick-start velocity (uu) and the pair rows.
    Vec3 start_vel[kMaxArenaCars];
    CarPairRows pair_rows;
};

// Where an arena's records live for one tick.
struct CarArenaRefs {
    BallState* ball;
    EnvState* env;
    ArenaInternal* st;
    BoostPadState* pads;
    CarState* cars;              // count cars
    WheelRuntimeState* wheels;   // count
    const CarConfig* cfgs;       // count
    const uint8_t* teams;        // count
    const CarControls* controls; // count
    int count;
    int env_idx;
    int car_base;  // batch index of this arena's car slot 0 (pad locks and respawns store base + slot + 1)
    int ray_lanes = 1;  // lanes of a wavefront stepping this arena together, each with the same data (1, 2, 4)
    int ray_sub = 0;    // this lane's index among them (the wheel rays it walks)
};

// ------------------------------------------------------------------------------------------------------------
// Visit order (rocketgpu.h:575-626).

// Mode 3 keeps the game's actor list: a car that comes back goes after the last present car, a car demolished this
// tick swaps places with the last present car.
RG_HD void UpdateVisitOrder(EnvState& env, const CarState* cars, int count) {
    if (env.car_visit_order_mode != 3) return;
    int order[8];
    for (int k = 0; k < 8; ++k) order[k] = static_cast<int>((env.car_visit_order_packed >> (4 * k)) & 0xFu);
    uint32_t present = env.car_visit_present_mask;
    const auto is_present = [&](int s) { return s < count && ((present >> s) & 1u) != 0u; };
    const auto last_present = [&]() {
        int q = -1;
        for (int k = 0; k < 8; ++k) {
            if (is_present(order[k])) q = k;
        }
        return q;
    };
    for (int s = 0; s < count && s < 8; ++s) {
        if (((present >> s) & 1u) || cars[s].is_demoed) continue;
        int at = 0;
        while (order[at] != s) ++at;
        const int dst = last_present() + 1;
        for (int k = at; k > dst; --k) order[k] = order[k - 1];
        order[dst < at ? dst : at] = s;
        present |= 1u << s;
    }
    for (int s = 0; s < count && s < 8; ++s) {
        if (!((present >> s) & 1u) || !cars[s].is_demoed) continue;
        int at = 0;
        while (order[at] != s) ++at;
        const int q = last_present();
        order[at] = order[q];
        order[q] = s;
        present &= ~(1u << s);
    }
    uint32_t packed = 0;
    for (int k = 0; k < 8; ++k) packed |= (static_cast<uint32_t>(order[k]) & 0xFu) << (4 * k);
    env.car_visit_order_packed = packed;
    env.car_visit_present_mask = static_cast<uint8_t>(present);
}

// ------------------------------------------------------------------------------------------------------------
// Boost pads (env_step.hpp:1945-2016, :2300-2610, :2706-2860).  A pad is a vertical cylinder the car's render-frame
// position sweeps through; pickup goes to the first eligible car by sweep order, and the pad locks for the rest of the
// tick.  Pads are the arena's own (positions from the layout).

RG_HD uint16_t PadFracQ(float frame_frac) {
    return static_cast<uint16_t>(Min(Max(frame_frac, 0.0f), 1.0f) * 65535.0f + 0.5f);
}

RG_HD void PadPreTick(BoostPadState& pad, float dt) {
    if (!pad.is_active || pad.cooldown > 0.0f) {
        pad.cooldown -= dt;
        pad.is_active = 0;
    }
    if (pad.cur_locked_car_id != 0) pad.cur_locked_car_id = 0;
}

RG_HD bool PadRespawnDue(const BoostPadState& pad, float dt, float frame_frac) {
    const float left = floorf(pad.cooldown / dt + 0.5f) + 1.0f;
    return left < 0.0f || (left == 0.0f && pad.pickup_frac_q < PadFracQ(frame_frac));
}

// Entry parameter of the segment from -> to into the pad's pickup cylinder (radius + 48 uu, +-96 uu tall).
RG_HD bool PadSweepEnter(const Vec3& pad_uu, const BoostPadState& pad, const Vec3& from, const Vec3& to, float& enter) {
    const float rad = (pad.is_big ? 160.0f + 48.0f : 128.0f + 48.0f) * (1.0f / 50.0f);
    const float half_z = (64.0f + 32.0f) * (1.0f / 50.0f);
    const float pad_z = (pad_uu.z - (pad.is_big ? 1.0f : 5.92f)) * (1.0f / 50.0f);
    float s0 = 0.0f, s1 = 1.0f;
    const float z0 = from.z - pad_z;
    const float dz = to.z - from.z;
    if (Abs(dz) < 1e-7f) {
        if (Abs(z0) >= half_z) return false;
    } else {
        const float ta = (-half_z - z0) / dz;
        const float tb = (half_z - z0) / dz;
        s0 = Max(s0, Min(ta, tb));
        s1 = Min(s1, Max(ta, tb));
    }
    const float ax = from.x - pad_uu.x * (1.0f / 50.0f);
    const float ay = from.y - pad_uu.y * (1.0f / 50.0f);
    const float dx = to.x - from.x;
    const float dy = to.y - from.y;
    const float a = dx * dx + dy * dy;
    const float b = ax * dx + ay * dy;
    const float c = ax * ax + ay * ay - rad * rad;
    if (a < 1e-12f) {
        enter = s0;
        return false;
    }
    const float disc = b * b - a * c;
    if (disc <= 0.0f) return false;
    const float root = sqrtf(disc);
    enter = Max(s0, (-b - root) / a);
    return enter < Min(s1, (-b + root) / a);
}

// Separating-axis test of the car's hitbox (shrunk by 1.25 uu, at its render-frame pose) against the pad's box.
RG_HD bool PadBoxOverlap(const Vec3& pad_uu, bool big, const Vec3& car_uu, const Vec3 axes[3], const CarConfig& cfg) {
    const float pr = big ? 160.0f : 128.0f;
    const float h[3] = {pr, pr, 64.0f};
    const float e[3] = {cfg.hitbox_size.x * 0.5f - 1.25f, cfg.hitbox_size.y * 0.5f - 1.25f,
                        cfg.hitbox_size.z * 0.5f - 1.25f};
    const Vec3 center = car_uu + axes[0] * cfg.hitbox_offset.x + axes[1] * cfg.hitbox_offset.y +
                        axes[2] * cfg.hitbox_offset.z;
    const float t[3] = {center.x - pad_uu.x, center.y - pad_uu.y, center.z - (pad_uu.z - (big ? 1.0f : 5.917984f))};
    float r[3][3], ar[3][3];
    for (int j = 0; j < 3; ++j) {
        const float c[3] = {axes[j].x, axes[j].y, axes[j].z};
        for (int i = 0; i < 3; ++i) {
            r[i][j] = c[i];
            ar[i][j] = Abs(c[i]) + 1e-6f;
        }
    }
    for (int i = 0; i < 3; ++i) {
        if (Abs(t[i]) > h[i] + e[0] * ar[i][0] + e[1] * ar[i][1] + e[2] * ar[i][2]) return false;
    }
    for (int j = 0; j < 3; ++j) {
        const float tb = t[0] * r[0][j] + t[1] * r[1][j] + t[2] * r[2][j];
        if (Abs(tb) > h[0] * ar[0][j] + h[1] * ar[1][j] + h[2] * ar[2][j] + e[j]) return false;
    }
    for (int i = 0; i < 3; ++i) {
        const int i1 = (i + 1) % 3, i2 = (i + 2) % 3;
        for (int j = 0; j < 3; ++j) {
            const int j1 = (j + 1) % 3, j2 = (j + 2) % 3;
            const float ra = h[i1] * ar[i2][j] + h[i2] * ar[i1][j];
            const float rb = e[j1] * ar[i][j2] + e[j2] * ar[i][j1];
            if (Abs(t[i2] * r[i1][j] - t[i1] * r[i2][j]) > ra + rb) return false;
        }
    }
    return true;
}

// The car's position as the game's render frame sees it: the replicated (0.01 uu) positions at the tick's start and
// end, interpolated by the frame fraction.
RG_HD Vec3 PadRbPos(const Vec3& pos_bt) {
    const Vec3 p = pos_bt * 50.0f;
    return Vec3(QuantRoundPos(p.x, 100.0f), QuantRoundPos(p.y, 100.0f), QuantRoundPos(p.z, 100.0f));
}
RG_HD Vec3 PadFramePos(const CarState& car, float frame_frac) {
    const Vec3 p0 = PadRbPos(car.tick_start_pos_bt);
    const Vec3 p1 = PadRbPos(car.pos_bt);
    return p0 + (p1 - p0) * frame_frac;
}
RG_HD void PadFrameAxes(const CarState& car, const CarParams& m, float frame_frac, Vec3 axes[3]) {
    CarBasis(car, axes[0], axes[1], axes[2]);
    const Vec3 w = car.is_demoed ? car.ang_vel : ClampLength(car.ang_vel, m.car_max_ang_speed);
    const Vec3 back = w * (-(1.0f - frame_frac) * m.tick_dt);
    for (int a = 0; a < 3; ++a) axes[a] = axes[a] + Cross(back, axes[a]);
}
RG_HD Vec3 PadSweepFrom(const CarState& car, const Vec3& car_bt) {
    if (car.pad_frame_pos_valid) {
        const Vec3 moved = car_bt - car.pad_frame_pos_bt;
        if (Dot(moved, moved) <= (240.0f / 50.0f) * (240.0f / 50.0f)) return car.pad_frame_pos_bt;
    }
    const Vec3 p0 = PadRbPos(car.tick_start_pos_bt);
    const Vec3 p1 = PadRbPos(car.pos_bt);
    return car_bt - (p1 - p0) * (1.0f / 50.0f);
}

struct PadLayout {
    const Vec3* pos;  // uu
    int count;
};

// The pads the car's frame sweep enters this tick (`hits`) and the pads its hitbox box overlaps among those and the
// one it is touching (`overlap`) (BoostPadContactsForCar, env_step.hpp:2412-2504, custom-layout branch).
RG_HD uint64_t PadContacts(const CarState& car, const CarConfig& cfg, const BoostPadState* pads, const PadLayout& L,
                           const CarParams& m, uint64_t& overlap, Vec3& frame_bt) {
    overlap = 0u;
    uint64_t hits = 0u;
    const float ff = m.frame_frac;
    if (car.is_demoed || ff < 0.0f) return hits;
    const Vec3 car_uu = PadFramePos(car, ff);
    const Vec3 car_bt = car_uu * (1.0f / 50.0f);
    frame_bt = car_bt;
    const Vec3 prev_bt = PadSweepFrom(car, car_bt);
    const Vec3 lo = Min3(car_bt, prev_bt), hi = Max3(car_bt, prev_bt);
    constexpr float k240 = 240.0f * (1.0f / 50.0f), k135 = 135.0f * (1.0f / 50.0f);
    for (int p = 0; p < L.count; ++p) {
        const Vec3 pad_bt = L.pos[p] * (1.0f / 50.0f);
        if (pad_bt.x < lo.x - k240 || pad_bt.x > hi.x + k240 || pad_bt.y < lo.y - k240 || pad_bt.y > hi.y + k240 ||
            pad_bt.z < lo.z - k135 || pad_bt.z > hi.z + k135) {
            continue;
        }
        float enter = 0.0f;
        if (PadSweepEnter(L.pos[p], pads[p], prev_bt, car_bt, enter)) hits |= uint64_t{1} << p;
    }
    uint64_t check = hits;
    if (car.boost_pad_touching != 0) check |= uint64_t{1} << (car.boost_pad_touching - 1);
    if (check != 0u) {
        Vec3 axes[3];
        PadFrameAxes(car, m, ff, axes);
        for (; check != 0u; check &= check - 1u) {
            const int p = Ctz64(check);
            if (PadBoxOverlap(L.pos[p], pads[p].is_big != 0, car_uu, axes, cfg)) overlap |= uint64_t{1} << p;
        }
    }
    return hits;
}

RG_HD bool PadEligible(const CarState& car, const CarParams& m) {
    return !car.is_demoed && !car.driving_disabled && car.boost < m.boost_max;
}

RG_HD void PadTouchStep(BoostPadState& pad, int p, CarState& car, uint32_t car_id, bool in_cyl, bool overlapping,
                        bool eligible) {
    if (!pad.is_active || pad.cur_locked_car_id != 0) return;
    const uint8_t touching = static_cast<uint8_t>(p + 1);
    if (car.boost_pad_touching == touching) {
        if (!overlapping) car.boost_pad_touching = 0;
    } else if (in_cyl) {
        if (eligible) {
            car.boost_pad_touching = touching;
            pad.cur_locked_car_id = car_id;
        } else if (overlapping) {
            car.boost_pad_touching = touching;
        }
    }
}

// Give the pad's boost to car `car_id` (a batch car id, base + slot + 1).
RG_HD void PadGive(CarArenaRefs& a, const CarParams& m, BoostPadState& pad, int p, uint32_t car_id) {
    CarState& car = a.cars[static_cast<int>(car_id) - 1 - a.car_base];
    SetCarBoostFraction(car, Min(m.boost_max / 100.0f, CarBoostFraction(car) + (pad.is_big ? 1.0f : 0.12f)));
    pad.is_active = 0;
    pad.cooldown = (pad.is_big ? m.boost_pad_cooldown_big : m.boost_pad_cooldown_small) - m.tick_dt;
    pad.pickup_frac_q = PadFracQ(m.frame_frac);
    for (int i = 0; i < a.count; ++i) {
        if (a.cars[i].boost_pad_touching == p + 1) a.cars[i].boost_pad_touching = 0;
    }
}

// A car whose sweep entered two or more pads takes them in entry order (BoostPadSweepTouchSteps, :2592-2638).
RG_HD void PadSweepTouchSteps(CarArenaRefs& a, const PadLayout& L, const CarParams& m, int slot, uint64_t hits,
                              uint64_t overlap) {
    CarState& car = a.cars[slot];
    const uint32_t car_id = static_cast<uint32_t>(a.car_base + slot + 1);
    const Vec3 to = PadFramePos(car, m.frame_frac) * (1.0f / 50.0f);
    const Vec3 from = PadSweepFrom(car, to);
    if (car.boost_pad_touching != 0 && ((hits >> (car.boost_pad_touching - 1)) & 1u) == 0u) {
        const int p = car.boost_pad_touching - 1;
        PadTouchStep(a.pads[p], p, car, car_id, false, ((overlap >> p) & 1u) != 0u, false);
    }
    for (uint64_t left = hits; left != 0u;) {
        int p = -1;
        float best = 0.0f;
        for (uint64_t mm = left; mm != 0u; mm &= mm - 1u) {
            const int q = Ctz64(mm);
            float s = 0.0f;
            PadSweepEnter(L.pos[q], a.pads[q], from, to, s);
            if (p < 0 || s < best) {
                p = q;
                best = s;
            }
        }
        if (p < 0) break;
        left &= ~(uint64_t{1} << p);
        BoostPadState& pad = a.pads[p];
        if (!pad.is_active || pad.cur_locked_car_id != 0) continue;
        if (car.boost_pad_touching == p + 1) {
            if (((overlap >> p) & 1u) == 0u) car.boost_pad_touching = 0;
            continue;
        }
        if (PadEligible(car, m)) {
            car.boost_pad_touching = static_cast<uint8_t>(p + 1);
            pad.cur_locked_car_id = car_id;
            PadGive(a, m, pad, p, car_id);
        } else if (((overlap >> p) & 1u) != 0u) {
            car.boost_pad_touching = static_cast<uint8_t>(p + 1);
        }
    }
}

// The tail's pad pass (BoostPadTouchPassForEnv, :2706-2834): per car in visit order its sweep contacts, then
// respawns of inactive pads (a car on a respawned pad starts touching it), then the touch list in reverse visit
// order locks pads for the post-tick.
RG_HD void PadTouchPass(CarArenaRefs& a, const PadLayout& L, const CarParams& m) {
    if (m.frame_frac < 0.0f) return;
    const EnvState& env = *a.env;
    for (int k = 0; k < a.count; ++k) {
        const int i = VisitSlot(env, k, a.count);
        CarState& car = a.cars[i];
        uint64_t overlap = 0u;
        Vec3 frame_bt(0.0f, 0.0f, 0.0f);
        const uint64_t hits = PadContacts(car, a.cfgs[i], a.pads, L, m, overlap, frame_bt);
        const auto store_frame = [&]() {
            if (car.is_demoed) {
                car.pad_frame_pos_bt = PadFramePos(car, m.frame_frac) * (1.0f / 50.0f);
                car.pad_frame_pos_valid = 0;
            } else {
                car.pad_frame_pos_bt = frame_bt;
                car.pad_frame_pos_valid = 1;
            }
        };
        if ((hits & (hits - 1u)) != 0u) {
            PadSweepTouchSteps(a, L, m, i, hits, overlap);
            store_frame();
            continue;
        }
        const bool eligible = PadEligible(car, m);
        uint64_t near = hits;
        if (car.boost_pad_touching != 0) near |= uint64_t{1} << (car.boost_pad_touching - 1);
        for (; near != 0u; near &= near - 1u) {
            const int p = Ctz64(near);
            PadTouchStep(a.pads[p], p, car, static_cast<uint32_t>(a.car_base + i + 1), ((hits >> p) & 1u) != 0u,
                         ((overlap >> p) & 1u) != 0u, eligible);
        }
        store_frame();
    }
    for (int p = 0; p < L.count; ++p) {
        BoostPadState& pad = a.pads[p];
        if (pad.is_active || !PadRespawnDue(pad, m.tick_dt, m.frame_frac)) continue;
        pad.cooldown = 0.0f;
        pad.is_active = 1;
        for (int k = a.count - 1; k >= 0; --k) {
            const int i = VisitSlot(env, k, a.count);
            CarState& car = a.cars[i];
            if (car.is_demoed) continue;
            const Vec3 car_uu = PadFramePos(car, m.frame_frac);
            Vec3 axes[3];
            PadFrameAxes(car, m, m.frame_frac, axes);
            if (PadBoxOverlap(L.pos[p], pad.is_big != 0, car_uu, axes, a.cfgs[i])) {
                car.boost_pad_touching = static_cast<uint8_t>(p + 1);
            }
        }
    }
    for (int k = a.count - 1; k >= 0; --k) {
        const int i = VisitSlot(env, k, a.count);
        CarState& car = a.cars[i];
        if (car.boost_pad_touching == 0) continue;
        const int p = car.boost_pad_touching - 1;
        BoostPadState& pad = a.pads[p];
        if (pad.is_active && pad.cur_locked_car_id == 0 && car.boost_pad_touching == p + 1 && PadEligible(car, m)) {
            pad.cur_locked_car_id = static_cast<uint32_t>(a.car_base + i + 1);
            const int whole = static_cast<int>(car.boost);
            car.rlbotv5_boost_hold =
                static_cast<uint8_t>(MinI(254, whole + (static_cast<float>(whole) < car.boost ? 1 : 0)) + 1);
        }
    }
}

// The pads' post-tick (PostTickBoostPad, :2836-2849): a locked active pad gives its boost.
RG_HD void PadPostTick(CarArenaRefs& a, const PadLayout& L, const CarParams& m) {
    for (int p = 0; p < L.count; ++p) {
        BoostPadState& pad = a.pads[p];
        const uint32_t locked = pad.cur_locked_car_id;
        if (locked != 0 && pad.is_active) PadGive(a, m, pad, p, locked);
        if (pad.prev_locked_car_id != locked) pad.prev_locked_car_id = locked;
    }
}

// ------------------------------------------------------------------------------------------------------------
// Broadphase with every body (BuildBulletManifoldList, arena_geometry.hpp:4162-5033).

struct BroadphaseBodies {
    const CarState* cars;
    const CarConfig* cfgs;
    const CarProxyPose* pose;  // cars' hitbox poses now
    int body_count;            // 1 + cars
};

// A new car-mesh manifold needs the mesh's box to reach the car's hitbox in the car's frame.
RG_HD bool CompoundTreeOverlap(const BodyProxy& px, const CarProxyPose& pose, const CarState& car,
                               const CarConfig& cfg, const MeshInfo& info) {
    const Vec3 child = cfg.hitbox_offset * (1.0f / 50.0f);
    const Vec3 lmin = child - pose.half, lmax = child + pose.half;
    const Vec3 no(-car.pos_bt.x, -car.pos_bt.y, -car.pos_bt.z);
    const Vec3 c = (info.aabb_max + info.aabb_min) * 0.5f;
    const Vec3 hh = (info.aabb_max - info.aabb_min) * 0.5f;
    (void)px;
    for (int k = 0; k < 3; ++k) {
        const Vec3& r = pose.axis[k];
        const float io = (r.x * no.x + r.y * no.y) + r.z * no.z;
        const float ctr = (r.x * c.x + r.y * c.y) + r.z * c.z + io;
        const float ext = (hh.x * Abs(r.x) + hh.y * Abs(r.y)) + hh.z * Abs(r.z);
        const float lo = k == 0 ? lmin.x : (k == 1 ? lmin.y : lmin.z);
        const float hi = k == 0 ? lmax.x : (k == 1 ? lmax.y : lmax.z);
        if (!(ctr - ext <= hi && ctr + ext >= lo)) return false;
    }
    return true;
}

// A new body-body manifold needs the hitboxes (or hitbox and ball box) to overlap on the first body's axes.
RG_HD bool LocalOverlap(const SweepState& h, const CarArenaTick& t, uint32_t key) {
    const int kind = KeyKind(key);
    int a = KeyBodyA(key) + 1;
    int b = kind == 1 ? 0 : KeyBodyB(key) + 1;
    if (kind == 3 && h.body_handle[b] < h.body_handle[a]) {
        const int s = a;
        a = b;
        b = s;
    }
    if (!t.proxy[a].valid || !t.proxy[b].valid || a == 0) return false;
    const CarProxyPose& pa = t.pose[a - 1];
    Vec3 cb, hb_v;
    const Vec3* bax = nullptr;
    if (b == 0) {
        cb = (t.proxy[0].child_min + t.proxy[0].child_max) * 0.5f;
        hb_v = (t.proxy[0].child_max - t.proxy[0].child_min) * 0.5f;
    } else {
        cb = t.pose[b - 1].center;
        hb_v = t.pose[b - 1].half;
        bax = t.pose[b - 1].axis;
    }
    const Vec3 d = cb - pa.center;
    const float ha[3] = {pa.half.x, pa.half.y, pa.half.z};
    const float hb[3] = {hb_v.x, hb_v.y, hb_v.z};
    for (int k = 0; k < 3; ++k) {
        const Vec3& ak = pa.axis[k];
        float e = hb[k];
        if (bax != nullptr) {
            e = Abs(Dot(ak, bax[0])) * hb[0] + Abs(Dot(ak, bax[1])) * hb[1] + Abs(Dot(ak, bax[2])) * hb[2];
        }
        if (Abs(Dot(d, ak)) > ha[k] + e) return false;
    }
    return true;
}

RG_HD void SweepQuantizeBox(const Vec3& mn, const Vec3& mx, uint16_t qmin[3], uint16_t qmax[3]) {
    const float a[3] = {mn.x, mn.y, mn.z};
    const float b[3] = {mx.x, mx.y, mx.z};
    for (int k = 0; k < 3; ++k) {
        qmin[k] = SweepQuantize(a[k], 0);
        qmax[k] = SweepQuantize(b[k], 1);
    }
}

// Update the sweep with every body's proxy this tick and rebuild the island-sorted manifold list into t.entries.
// `predictive`: bit b set when body b's predictive ray hit.
RG_HD void ArenaBroadphase(ArenaInternal& st, const ArenaMeshView& mesh, CarArenaTick& t, const BroadphaseBodies& B,
                           uint32_t predictive, bool static_pairs_skipped) {
    BroadphaseState& bp = st.bp;
    SweepState& h = bp.sweep;
    const int bodies = B.body_count;
    if (h.handle_count == 0) {
        SweepInitStatics(bp, mesh);
        bp.pair_count = 0;
        bp.manifold_count = 0;
        bp.predictive_count = 0;
    }
    Sweep s(bp, bodies);
    for (int body = 1; body < bodies; ++body) {
        if (!t.proxy[body].valid && h.body_handle[body] != 0) s.RemoveBody(body);
    }
    for (int body = 0; body < bodies; ++body) {
        const BodyProxy& px = t.proxy[body];
        if (!px.valid || h.body_handle[body] != 0 || h.handle_count >= kSweepMaxHandles - 1) continue;
        uint16_t qmin[3], qmax[3];
        SweepQuantizeBox(px.child_min, px.child_max, qmin, qmax);
        h.body_handle[body] = h.free_count > 0 ? h.free_list[h.free_count - 1] : h.next_fresh;
        s.AddHandle(qmin, qmax, static_cast<int16_t>(-1 - body), true);
        h.obj[h.obj_count++] = static_cast<uint8_t>(body);
    }
    for (int i = 0; i < bp.predictive_count; ++i) {
        SwapRemove(bp.manifolds, bp.manifold_count, FindKey(bp.manifolds, bp.manifold_count, bp.predictive[i]));
    }
    bp.predictive_count = 0;
    for (int obj = 0; obj < h.obj_count && predictive != 0u; ++obj) {
        const int body = h.obj[obj];
        if (body >= bodies || ((predictive >> body) & 1u) == 0u || bp.manifold_count >= kMaxManifoldEntries) continue;
        const uint32_t key = PackPairKey(body == 0 ? 0 : 2, body - 1, -1, kPredictiveStaticId);
        bp.manifolds[bp.manifold_count++] = key;
        bp.predictive[bp.predictive_count++] = key;
    }
    for (int obj = 0; obj < h.obj_count; ++obj) {
        const int body = h.obj[obj];
        if (body >= bodies) continue;
        const BodyProxy& px = t.proxy[body];
        const int handle = h.body_handle[body];
        if (!px.valid || handle == 0) continue;
        float* box = h.body_aabb[body];
        box[0] = px.aabb_min.x;
        box[1] = px.aabb_min.y;
        box[2] = px.aabb_min.z;
        box[3] = px.aabb_max.x;
        box[4] = px.aabb_max.y;
        box[5] = px.aabb_max.z;
        h.body_aabb_valid = static_cast<uint16_t>(h.body_aabb_valid | (1u << body));
        uint16_t qmin[3], qmax[3];
        SweepQuantizeBox(px.aabb_min, px.aabb_max, qmin, qmax);
        for (int axis = 0; axis < 3; ++axis) {
            const int emin = h.min_edge[handle][axis];
            const int emax = h.max_edge[handle][axis];
            const int dmin = static_cast<int>(qmin[axis]) - static_cast<int>(h.edge_pos[axis][emin]);
            const int dmax = static_cast<int>(qmax[axis]) - static_cast<int>(h.edge_pos[axis][emax]);
            if (dmin == 0 && dmax == 0) continue;
            if (dmin != 0) h.edge_pos[axis][emin] = qmin[axis];
            if (dmax != 0) h.edge_pos[axis][emax] = qmax[axis];
            if (dmin < 0) s.SortDown(axis, emin, false, true);
            if (dmax > 0) s.SortUp(axis, emax, true, true);
            if (dmin > 0) s.SortUp(axis, emin, false, true);
            if (dmax < 0) s.SortDown(axis, emax, true, true);
        }
    }
    // Pairs -> manifolds.
    const auto mesh_info_for = [&](int mesh_index) -> const MeshInfo* {
        if (mesh_index >= 0 && mesh_index < mesh.mesh_count && mesh.meshes[mesh_index].mesh_index == mesh_index) {
            return mesh.meshes[mesh_index].tri_count > 0 ? &mesh.meshes[mesh_index] : nullptr;
        }
        for (int k = 0; k < mesh.mesh_count; ++k) {
            if (mesh.meshes[k].mesh_index == mesh_index && mesh.meshes[k].tri_count > 0) return &mesh.meshes[k];
        }
        return nullptr;
    };
    for (int i = 0; i < bp.pair_count; ++i) {
        const uint32_t key = bp.pairs[i];
        const int kind = KeyKind(key);
        const int at = FindKey(bp.manifolds, bp.manifold_count, key);
        bool want;
        if (kind == 0) {
            want = t.proxy[0].valid != 0;
        } else {
            const int body_a = kind <= 1 ? 0 : KeyBodyA(key) + 1;
            const BodyProxy* pa = body_a < bodies && t.proxy[body_a].valid ? &t.proxy[body_a] : nullptr;
            if (pa == nullptr) {
                want = false;
            } else if (kind == 1 || kind == 3) {
                const int body_b = kind == 1 ? KeyBodyA(key) + 1 : KeyBodyB(key) + 1;
                const BodyProxy* pb = body_b < bodies && t.proxy[body_b].valid ? &t.proxy[body_b] : nullptr;
                want = pb != nullptr && AabbOverlap(pa->child_min, pa->child_max, pb->child_min, pb->child_max);
            } else {
                const MeshInfo* info = mesh_info_for(KeyStatic(key));
                want = info != nullptr && AabbOverlap(pa->child_min, pa->child_max, info->aabb_min, info->aabb_max);
            }
        }
        if (want && at < 0 && (kind == 1 || kind == 3) && !LocalOverlap(h, t, key)) continue;
        if (want && at < 0 && kind == 2 && KeyStatic(key) < 100) {
            const int car = KeyBodyA(key);
            const MeshInfo* info = mesh_info_for(KeyStatic(key));
            if (info != nullptr && !CompoundTreeOverlap(t.proxy[car + 1], B.pose[car], B.cars[car], B.cfgs[car], *info)) {
                continue;
            }
        }
        if (want && at < 0 && kind == 0 && static_pairs_skipped) continue;
        if (want && at < 0) {
            bp.manifolds[bp.manifold_count++] = key;
        } else if (!want) {
            SwapRemove(bp.manifolds, bp.manifold_count, at);
        }
    }
    // Islands: union of bodies joined by a body-body pair, by object index.
    int8_t tag[kMaxBodies];
    for (int b = 0; b < kMaxBodies; ++b) tag[b] = -1;
    for (int obj = 0; obj < h.obj_count; ++obj) tag[h.obj[obj]] = static_cast<int8_t>(obj);
    int parent[kMaxBodies];
    for (int k = 0; k < kMaxBodies; ++k) parent[k] = k;
    const auto root_of = [&](int x) {
        while (parent[x] != x) x = parent[x];
        return x;
    };
    for (int i = 0; i < bp.pair_count; ++i) {
        const uint32_t key = bp.pairs[i];
        const int kind = KeyKind(key);
        if (kind != 1 && kind != 3) continue;
        int b0 = kind == 1 ? 0 : KeyBodyA(key) + 1;
        int b1 = kind == 1 ? KeyBodyA(key) + 1 : KeyBodyB(key) + 1;
        if (h.body_handle[b0] > h.body_handle[b1]) {
            const int s2 = b0;
            b0 = b1;
            b1 = s2;
        }
        if (tag[b0] < 0 || tag[b1] < 0) continue;
        const int ra = root_of(tag[b0]), rb = root_of(tag[b1]);
        if (ra != rb) parent[ra] = rb;
    }
    int8_t island_of[kMaxBodies];
    for (int b = 0; b < kMaxBodies; ++b) island_of[b] = static_cast<int8_t>(tag[b] >= 0 ? root_of(tag[b]) : -1);
    const int n = bp.manifold_count;
    int8_t island[kMaxManifoldEntries];
    int8_t order[kMaxManifoldEntries];
    for (int i = 0; i < n; ++i) {
        const uint32_t key = bp.manifolds[i];
        island[i] = island_of[KeyKind(key) <= 1 ? 0 : KeyBodyA(key) + 1];
        order[i] = static_cast<int8_t>(i);
    }
    if (n > 1) QuickSortByIsland(order, island, 0, n - 1);
    for (int i = 0; i < n; ++i) {
        t.entries[i].key = bp.manifolds[order[i]];
        t.entries[i].island = island[order[i]];
    }
    t.entry_count = n;
    // The ball's view of the list: its own manifolds in list order.
    int nb = 0;
    for (int i = 0; i < n && nb < kMaxBallManifoldEntries; ++i) {
        if (KeyKind(t.entries[i].key) == 0) t.ball.entries[nb++] = t.entries[i];
    }
    t.ball.entry_count = nb;
}

// ------------------------------------------------------------------------------------------------------------
// Body-contact screens: whether a body contact this build does not compute could arise this tick.

// Could the car's hitbox come within `reach` bt of a triangle of the mesh?  A separating-axis screen over the
// triangles whose boxes the hitbox's box (grown by reach) overlaps: box axes, triangle normal and the nine edge
// cross products; a triangle separated by more than reach on some axis cannot touch.
RG_HD bool CarNearMesh(const CarProxyPose& pose, const ArenaMeshView& mesh, float reach) {
    Vec3 ext(0.0f, 0.0f, 0.0f);
    for (int k = 0; k < 3; ++k) {
        const float hk = k == 0 ? pose.half.x : (k == 1 ? pose.half.y : pose.half.z);
        ext = ext + Vec3(Abs(pose.axis[k].x), Abs(pose.axis[k].y), Abs(pose.axis[k].z)) * hk;
    }
    const Vec3 lo = pose.center - ext - Vec3(reach, reach, reach);
    const Vec3 hi = pose.center + ext + Vec3(reach, reach, reach);
    bool near = false;
    for (int mi = 0; mi < mesh.mesh_count && !near; ++mi) {
        const MeshInfo& info = mesh.meshes[mi];
        if (info.tri_count <= 0 || !AabbOverlap(lo, hi, info.aabb_min, info.aabb_max)) continue;
        uint16_t qmin[3], qmax[3];
        QuantizePoint(info.q, lo, 0, qmin);
        QuantizePoint(info.q, hi, 1, qmax);
        WalkMeshBvh(
            mesh.nodes, info.node_start, info.node_start + info.node_count, qmin, qmax,
            [&](int32_t local) {
                if (near) return;
                const MeshTriangle& tri = mesh.tris[info.tri_start + local];
                const Vec3 v[3] = {tri.a - pose.center, tri.b - pose.center, tri.c - pose.center};
                const Vec3 e[3] = {tri.b - tri.a, tri.c - tri.b, tri.a - tri.c};
                const auto separated = [&](const Vec3& axis) {
                    const float len = Length(axis);
                    if (!(len > 1e-6f)) return false;
                    const Vec3 u = axis * (1.0f / len);
                    float tmin = Dot(v[0], u), tmax = tmin;
                    for (int q = 1; q < 3; ++q) {
                        const float d = Dot(v[q], u);
                        tmin = Min(tmin, d);
                        tmax = Max(tmax, d);
                    }
                    const float r = Abs(Dot(pose.axis[0], u)) * pose.half.x + Abs(Dot(pose.axis[1], u)) * pose.half.y +
                                    Abs(Dot(pose.axis[2], u)) * pose.half.z;
                    return tmin > r + reach || tmax < -r - reach;
                };
                if (separated(Cross(e[0], e[1]))) return;
                for (int k = 0; k < 3; ++k) {
                    if (separated(pose.axis[k])) return;
                }
                for (int k = 0; k < 3; ++k) {
                    for (int q = 0; q < 3; ++q) {
                        if (separated(Cross(pose.axis[k], e[q]))) return;
                    }
                }
                near = true;
            },
            AcceptAll{});
    }
    return near;
}

// ------------------------------------------------------------------------------------------------------------
// The tick.

RG_HD void StepCarArena(CarArenaRefs& a, const ArenaMeshView& mesh, const StepParams& p, const CarParams& m,
                        const PadLayout& pads, const CarRespawnContext& respawn_base, bool report_fields,
                        CarArenaTick& t) {
    BallState& ball = *a.ball;
    EnvState& env = *a.env;
    ArenaInternal& st = *a.st;
    const int count = a.count;
    const float dt = p.tick_dt;
    uint32_t unsupported = st.unsupported;

    // 1. env pre-tick.
    UpdateVisitOrder(env, a.cars, count);
    {
        uint32_t demoed = 0u;
        if (env.car_visit_order_mode == 3 && count <= 8) {
            demoed = ~static_cast<uint32_t>(env.car_visit_present_mask) & ((1u << count) - 1u);
        } else {
            for (int i = 0; i < count; ++i) demoed |= static_cast<uint32_t>(a.cars[i].is_demoed != 0) << i;
        }
        for (int i = 0; i < count; ++i) {
            if (((demoed >> i) & 1u) == 0u) continue;
            // A due car with a chosen spot (a kickoff re-creation or an override) keeps it; one without draws a
            // free respawn spot here (PickRespawnSpotsForEnv, env_step.hpp:6615-6650), not built.
            const float left = Max(0.0f, a.cars[i].demo_respawn_timer - dt);
            if (!(Max(0.0f, left - dt) > 0.0f) && a.cars[i].respawn_slot_override_plus1 == 0) {
                unsupported |= kUnsupportedRespawnSpot;
            }
        }
    }
    env.goal_scored_this_tick = 0;

    // 2. car pre-ticks, ball, pads.
    for (int k = 0; k < count; ++k) {
        const int i = VisitSlot(env, k, count);
        CarRespawnContext rc = respawn_base;
        rc.team = a.teams[i];
        rc.car_id = static_cast<uint32_t>(a.car_base + i + 1);
        rc.env_tick = env.tick;
        CarPreTick(a.cars[i], a.controls[i], a.cfgs[i], m, &a.wheels[i], rc, &t.start_vel[i]);
        t.predict_vel[i] = a.cars[i].vel_bt;
        t.predict_ang[i] = a.cars[i].ang_vel;
    }
    BallPreTick(ball, p, t.ball.predict_vel_bt);
    if (count > 0) {
        for (int q = 0; q < pads.count; ++q) PadPreTick(a.pads[q], dt);
    }

    // 3. vehicle updates.  Each car's rays are cast against the bodies whose last swept boxes they reach.
    for (int k = 0; k < count; ++k) {
        const int i = VisitSlot(env, k, count);
        CarState& car = a.cars[i];
        uint32_t near = 0u;
        if (!car.is_demoed) {
            const SweepState& h = st.bp.sweep;
            const int bodies = MinI(count + 1, kMaxBodies);
            const uint32_t mask = h.body_aabb_valid & ((1u << bodies) - 1u) & ~(1u << (i + 1));
            const uint32_t all_dyn = ((1u << (MinI(count, kMaxBodies - 1) + 1)) - 1u) & ~(1u << (i + 1));
            if (mask == 0u) {
                near = all_dyn;
            } else {
                // Box of the four wheel rays at full length.
                Vec3 f, r, u;
                if (car.basis_authoritative) {
                    f = car.basis_forward;
                    r = car.basis_right;
                    u = car.basis_up;
                } else {
                    QuatColumns(car.rot, f, r, u);
                }
                const CarConfig& cfg = a.cfgs[i];
                const float oz = cfg.hitbox_offset.z;
                const float fb = WheelBoneZ(cfg, true);
                const float bb = cfg.back_suspension_rest == cfg.front_suspension_rest ? fb : WheelBoneZ(cfg, false);
                const float flen = ((oz - fb) * (1.0f / 50.0f) + cfg.front_wheel_radius * (1.0f / 50.0f)) +
                                   ((kWheelMaxTravel - kWheelTravelTrim) * (1.0f / 50.0f));
                const float blen = ((oz - bb) * (1.0f / 50.0f) + cfg.back_wheel_radius * (1.0f / 50.0f)) +
                                   ((kWheelMaxTravel - kWheelTravelTrim) * (1.0f / 50.0f));
                Vec3 mn(0.0f, 0.0f, 0.0f), mx(0.0f, 0.0f, 0.0f);
                for (int wi = 0; wi < 4; ++wi) {
                    Vec3 off = wi < 2 ? cfg.front_wheels_offset : cfg.back_wheels_offset;
                    if (wi & 1) off.y = -off.y;
                    off.z = oz;
                    const Vec3 ob = off * (1.0f / 50.0f);
                    const Vec3 from = car.pos_bt + ((f * ob.x + r * ob.y) + u * ob.z);
                    const Vec3 to = from + (-u) * (wi < 2 ? flen : blen);
                    const Vec3 lo = Min3(from, to), hi = Max3(from, to);
                    mn = wi == 0 ? lo : Min3(mn, lo);
                    mx = wi == 0 ? hi : Max3(mx, hi);
                }
                uint32_t excluded = 0u;
                for (int body = 0; body < bodies; ++body) {
                    if (!((mask >> body) & 1u)) continue;
                    const float* b = h.body_aabb[body];
                    if (b[3] < mn.x || mx.x < b[0] || b[4] < mn.y || mx.y < b[1] || b[5] < mn.z || mx.z < b[2]) {
                        excluded |= 1u << body;
                    }
                }
                near = all_dyn & ~excluded;
            }
        }
        VehicleWorld world;
        world.mesh = &mesh;
        world.ball = &ball;
        world.ball_radius = p.ball_radius;
        world.cars = a.cars;
        world.cfgs = a.cfgs;
        world.count = count;
        world.self = i;
        world.others_near = near;
        world.ray_lanes = a.ray_lanes;
        world.ray_sub = a.ray_sub;
        CarVehicleUpdate(car, a.controls[i], a.cfgs[i], m, a.wheels[i], world, unsupported);
    }

    // 4. world step.
    const bool vehicle_woke = (ball.sleeping & 4u) != 0;
    ball.sleeping &= 3u;
    if (ball.sleeping == 2u) ball.sleeping = 0u;
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
    for (int i = 0; i < count; ++i) CarPrestep(a.cars[i], a.cfgs[i], a.wheels[i], m, dt, t.pre[i]);
    const Vec3 solver_vel_bt = ball.vel_bt;
    t.ball.rows.n_split = t.ball.rows.n_vel = t.ball.rows.n_fric = 0;
    // Predictive rays: the ball's, then each car's swept hitbox center.
    t.ball.predictive_hit = 0;
    if (!(sleeping && !vehicle_woke)) {
        t.ball.predictive_hit =
            BallPredictiveRay<SerialWalk>(ball, dt, mesh, t.ball.predictive_normal, t.ball.predictive_distance_bt) ? 1
                                                                                                                   : 0;
    }
    uint32_t predictive = t.ball.predictive_hit ? 1u : 0u;
    for (int i = 0; i < count; ++i) {
        const CarState& car = a.cars[i];
        if (t.pre[i].is_demoed) continue;
        const Vec3 child = a.cfgs[i].hitbox_offset * (1.0f / 50.0f);
        const Vec3 f = car.basis_forward, r = car.basis_right, u = car.basis_up;
        const Vec3 cur(car.pos_bt.x + (f.x * child.x + r.x * child.y + u.x * child.z),
                       car.pos_bt.y + (f.y * child.x + r.y * child.y + u.y * child.z),
                       car.pos_bt.z + (f.z * child.x + r.z * child.y + u.z * child.z));
        Vec3 pf = f, pr = r, pu = u;
        Quat unused;
        IntegrateBasis(pf, pr, pu, unused, car.ang_vel, dt);
        const Vec3 po = car.pos_bt + car.vel_bt * dt;
        const Vec3 to(po.x + (pf.x * child.x + pr.x * child.y + pu.x * child.z),
                      po.y + (pf.y * child.x + pr.y * child.y + pu.y * child.z),
                      po.z + (pf.z * child.x + pr.z * child.y + pu.z * child.z));
        const Vec3 mv = to - cur;
        const float m2 = mv.x * mv.x + mv.y * mv.y + mv.z * mv.z;
        if (!(m2 > 0.02f * 0.02f)) continue;
        const float inv_len = 1.0f / sqrtf(m2);
        const Vec3 from(cur.x - inv_len * mv.x * 0.04f, cur.y - inv_len * mv.y * 0.04f, cur.z - inv_len * mv.z * 0.04f);
        const RayHit hit = CastArenaRay(mesh, from, to - from);
        if (hit.hit && hit.fraction < 1.0f) {
            unsupported |= kUnsupportedCarPredictive;
            predictive |= 1u << (i + 1);
        }
    }
    // Broadphase proxies.
    {
        constexpr float kContactThresholdBt = 0.02f;
        constexpr float kSphereAabbExtraBt = 0.08f;
        const Vec3 pos = ball.pos_bt;
        const Vec3 pred = pos + t.ball.predict_vel_bt * dt;
        const float extent = p.ball_radius * (1.0f / 50.0f) + kSphereAabbExtraBt;
        const float pe = (p.ball_radius * (1.0f / 50.0f) + 0.5f * kSphereAabbExtraBt) + 0.5f * kSphereAabbExtraBt;
        BodyProxy& px = t.proxy[0];
        px.aabb_min = Vec3(Min(pos.x, pred.x) - pe - kContactThresholdBt, Min(pos.y, pred.y) - pe - kContactThresholdBt,
                           Min(pos.z, pred.z) - pe - kContactThresholdBt);
        px.aabb_max = Vec3(Max(pos.x, pred.x) + pe + kContactThresholdBt, Max(pos.y, pred.y) + pe + kContactThresholdBt,
                           Max(pos.z, pred.z) + pe + kContactThresholdBt);
        px.child_min = Vec3(pos.x - extent, pos.y - extent, pos.z - extent);
        px.child_max = Vec3(pos.x + extent, pos.y + extent, pos.z + extent);
        px.valid = 1;
        t.ball.ball_proxy = px;
    }
    for (int i = 0; i < count; ++i) {
        if (t.pre[i].is_demoed) {
            t.proxy[i + 1].valid = 0;
            continue;
        }
        CarBroadphaseProxy(a.cars[i], a.cfgs[i], t.predict_vel[i], t.predict_ang[i], dt, t.proxy[i + 1], t.pose[i]);
    }
    BroadphaseBodies B;
    B.cars = a.cars;
    B.cfgs = a.cfgs;
    B.pose = t.pose;
    B.body_count = MinI(count + 1, kMaxBodies);
    ArenaBroadphase(st, mesh, t, B, predictive, sleeping && !vehicle_woke);
    // Ball-car pairs the tick would have to resolve.
    for (int i = 0; i < count; ++i) {
        const CarState& ci = a.cars[i];
        if (t.proxy[i + 1].valid && !ci.just_respawned &&
            AabbOverlap(t.proxy[i + 1].child_min, t.proxy[i + 1].child_max, t.proxy[0].child_min,
                        t.proxy[0].child_max)) {
            unsupported |= kUnsupportedBallCar;
        }
    }
    // Car-car contacts: the pairs in the contact-pair list's order, then the other pairs in slot order, each skipped
    // while its manifold is empty and its cars' bounding spheres are apart (or a car is out of play).
    CarPairRows& pr = t.pair_rows;
    pr.nn = pr.nf = pr.ns = 0;
    pr.cars = 0u;
    if (m.enable_car_car && count >= 2) {
        uint32_t skip = 0u;
        for (int hi = 1; hi < count; ++hi) {
            for (int lo = 0; lo < hi; ++lo) {
                const CarState& cl = a.cars[lo];
                const CarState& ch = a.cars[hi];
                const bool live = !cl.is_demoed && !cl.just_respawned && !ch.is_demoed && !ch.just_respawned;
                const float reach = CarBoundingRadius(a.cfgs[hi]) + CarBoundingRadius(a.cfgs[lo]);
                const bool apart = !(LengthSq(cl.pos - ch.pos) < reach * reach);
                if (st.car_pair[CarPairSlot(lo, hi)].count == 0 && (apart || !live)) skip |= 1u << CarPairSlot(lo, hi);
            }
        }
        int8_t order_lo[kCarPairSlots], order_hi[kCarPairSlots];
        int n_order = 0;
        uint32_t listed = 0u;
        // The contact-pair list as the broadphase keeps it (its island sort comes after the car-car stage).
        for (int e = 0; e < st.bp.manifold_count; ++e) {
            const uint32_t key = st.bp.manifolds[e];
            if (KeyKind(key) != 3) continue;
            const int lo = KeyBodyA(key), hi = KeyBodyB(key);
            if (lo < 0 || hi >= count || lo >= hi) continue;
            const uint32_t bit = 1u << CarPairSlot(lo, hi);
            if (listed & bit) continue;
            listed |= bit;
            if (skip & bit) continue;
            order_lo[n_order] = static_cast<int8_t>(lo);
            order_hi[n_order++] = static_cast<int8_t>(hi);
        }
        for (int lo = 0; lo < count; ++lo) {
            for (int hi = lo + 1; hi < count; ++hi) {
                const uint32_t bit = 1u << CarPairSlot(lo, hi);
                if ((listed | skip) & bit) continue;
                order_lo[n_order] = static_cast<int8_t>(lo);
                order_hi[n_order++] = static_cast<int8_t>(hi);
            }
        }
        for (int k = 0; k < n_order; ++k) {
            const int i = order_lo[k], j = order_hi[k];
            // Car a is the one whose broadphase handle is the larger (j on a tie).
            const bool i_first = st.bp.sweep.body_handle[i + 1] <= st.bp.sweep.body_handle[j + 1];
            const int ia = i_first ? j : i, ib = i_first ? i : j;
            CarPairSide sa{&a.cars[ia], &a.cfgs[ia], &t.pre[ia], t.start_vel[ia],
                           static_cast<uint32_t>(a.car_base + ia + 1), a.teams[ia], ia};
            CarPairSide sb{&a.cars[ib], &a.cfgs[ib], &t.pre[ib], t.start_vel[ib],
                           static_cast<uint32_t>(a.car_base + ib + 1), a.teams[ib], ib};
            CarPairContact(sa, sb, st.car_pair[CarPairSlot(i, j)], m, dt, pr, unsupported);
        }
    }
    // Ball-mesh contacts.
    const bool awake_rows = !sleeping || vehicle_woke;
    bool any_retained = false;
    for (int mi = 0; mi < mesh.mesh_count; ++mi) {
        any_retained = any_retained || (((st.ball_mesh_retained >> mi) & 1u) != 0u && st.ball_mesh[mi].count != 0);
    }
    const bool dispatched = !sleeping || vehicle_woke || any_retained;
    st.walk_rounds = 0;
    if (awake_rows && dispatched) {
        BallMeshContacts<SerialWalk, false>(ball, p, st, mesh, t.ball, ext_lin_uu, solver_vel_bt, dt, EventSink{});
    }
    // Car-mesh contacts: each car against the meshes its broadphase pairs list (and those holding retained points).
    // A car without car-car rows has rows that touch no other body: they are solved on their own.  The rows of the
    // cars with car-car rows are kept for the joint solve.
    bool car_vel_rows = false, car_split_rows = false;
    CarWorldRows& wr = t.world_rows;
    wr.nn = wr.nf = wr.ns = 0;
    wr.full = 0;
    for (int i = 0; i < count; ++i) {
        t.world_vel[i] = t.world_split[i] = 0;
        CarWorldManifold* retained = st.car_world[i];
        if (t.pre[i].is_demoed || a.cars[i].just_respawned) {
            for (int k = 0; k < kCarWorldManifoldsPerCar; ++k) retained[k].count = 0;
            continue;
        }
        uint32_t mesh_bits = 0u;
        for (int e = 0; e < t.entry_count; ++e) {
            const uint32_t key = t.entries[e].key;
            const int sid = KeyStatic(key);
            if (KeyKind(key) == 2 && KeyBodyA(key) == i && sid != kPredictiveStaticId && sid >= 0 && sid < 32) {
                mesh_bits |= 1u << sid;
            }
        }
        bool held = false;
        for (int k = 0; k < kCarWorldManifoldsPerCar; ++k) held = held || retained[k].count > 0;
        if (mesh_bits == 0u && !held) continue;
        CarState& car = a.cars[i];
        if (!car.basis_authoritative) {
            unsupported |= kUnsupportedCarWorld;
            continue;
        }
        const int n0 = wr.nn, f0 = wr.nf, s0 = wr.ns;
        CarWorldContacts(car, a.cfgs[i], m, mesh, mesh_bits, retained, dt, wr);
        for (int r = n0; r < wr.nn; ++r) wr.n_car[r] = static_cast<int8_t>(i);
        if (wr.hit) {
            car.world_contact.has_contact = 1;
            car.world_contact.contact_normal = wr.latest_normal;
        }
        // Rows of a car that touches another car stay for the joint solve.
        if ((pr.cars >> i) & 1u) continue;
        const float inv_mass = 1.0f / m.car_mass;
        if (wr.nn > n0 || wr.nf > f0) {
            SolveCarWorldVelocity(wr, n0, f0, inv_mass, m.arena_friction, t.world_dlin[i], t.world_dang[i]);
            t.world_vel[i] = 1;
            car_vel_rows = true;
        }
        if (wr.ns > s0) {
            SolveCarWorldSplit(wr, s0, inv_mass, t.world_push[i], t.world_turn[i]);
            t.world_split[i] = 1;
            car_split_rows = true;
        }
        wr.nn = n0;
        wr.nf = f0;
        wr.ns = s0;
    }
    if (wr.full) unsupported |= (pr.cars != 0u ? kUnsupportedCarCar : kUnsupportedCarWorld);
    // The joint solve of the cars with car-car rows.
    if (pr.cars != 0u) {
        Vec3 dl[kMaxArenaCars], da[kMaxArenaCars];
        for (int i = 0; i < count; ++i) dl[i] = da[i] = Vec3(0.0f, 0.0f, 0.0f);
        const float inv_mass = 1.0f / m.car_mass;
        if (wr.nn + wr.nf + pr.nn + pr.nf > 0) {
            SolveGroupVelocity(t.entries, t.entry_count, wr, pr, inv_mass, m.arena_friction, dl, da);
            car_vel_rows = true;
            for (int i = 0; i < count; ++i) {
                if (!((pr.cars >> i) & 1u)) continue;
                t.world_dlin[i] = dl[i];
                t.world_dang[i] = da[i];
                t.world_vel[i] = 1;
            }
        }
        if (wr.ns + pr.ns > 0) {
            for (int i = 0; i < count; ++i) dl[i] = da[i] = Vec3(0.0f, 0.0f, 0.0f);
            SolveGroupSplit(t.entries, t.entry_count, wr, pr, inv_mass, dl, da);
            car_split_rows = true;
            uint32_t touched = 0u;
            for (int r = 0; r < wr.ns; ++r) touched |= 1u << wr.n_car[wr.s[r].normal_row];
            for (int r = 0; r < pr.ns; ++r) {
                const CarPairRow& nr = pr.n[pr.s[r].normal_row];
                touched |= (1u << nr.a) | (1u << nr.b);
            }
            for (int i = 0; i < count; ++i) {
                if (!((touched >> i) & 1u)) continue;
                t.world_push[i] = dl[i];
                t.world_turn[i] = da[i];
                t.world_split[i] = 1;
            }
        }
    }
    if (t.ball.predictive_hit) BallPredictiveRows(ball, p, st, t.ball, ext_lin_uu, solver_vel_bt, dt);
    // Velocity solve.  With rows anywhere in the arena every body takes its accumulated change (zero for bodies
    // without rows), as the arena-wide solve writes them all back.
    const bool ball_vel_rows = t.ball.rows.n_vel > 0 || t.ball.rows.n_fric > 0;
    if (ball_vel_rows) {
        SolveBallVelocity(ball, t.ball, solver_vel_bt);
    } else if (car_vel_rows) {
        SetBallVelBt(ball, solver_vel_bt + Vec3(0.0f, 0.0f, 0.0f));
        ball.ang_vel += Vec3(0.0f, 0.0f, 0.0f);
    }
    if (ball_vel_rows || car_vel_rows) {
        for (int i = 0; i < count; ++i) {
            const Vec3 dl = t.world_vel[i] ? t.world_dlin[i] : Vec3(0.0f, 0.0f, 0.0f);
            const Vec3 da = t.world_vel[i] ? t.world_dang[i] : Vec3(0.0f, 0.0f, 0.0f);
            SetCarVelBt(a.cars[i], a.cars[i].vel_bt + dl);
            a.cars[i].ang_vel += da;
        }
    }
    // Split impulses.
    SolveBallSplit(ball, t.ball, dt);
    if (t.ball.rows.n_split > 0 || car_split_rows) {
        for (int i = 0; i < count; ++i) {
            CarState& car = a.cars[i];
            if (!t.world_split[i]) {
                SetCarPosBt(car, car.pos_bt + Vec3(0.0f, 0.0f, 0.0f) * dt);
                continue;
            }
            const Vec3 push = t.world_push[i], turn = t.world_turn[i];
            const bool nonzero = push.x != 0.0f || push.y != 0.0f || push.z != 0.0f || turn.x != 0.0f ||
                                 turn.y != 0.0f || turn.z != 0.0f;
            if (!nonzero) continue;
            SetCarPosBt(car, car.pos_bt + push * dt);
            IntegrateBasis(car.basis_forward, car.basis_right, car.basis_up, car.rot, turn * 0.1f, dt);
        }
    }
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
    for (int i = 0; i < count; ++i) CarFinish(a.cars[i], t.pre[i], dt);

    // 5. car post-ticks.
    for (int k = 0; k < count; ++k) {
        const int i = VisitSlot(env, k, count);
        CarState& car = a.cars[i];
        if (!t.pre[i].is_demoed && !car.just_respawned && car.is_demoed) car.rot = t.pre[i].rot;
        CarPostTick(car, a.controls[i], m, &a.wheels[i], report_fields);
    }

    // 6. env tail.
    if (count > 0) {
        PadTouchPass(a, pads, m);
        PadPostTick(a, pads, m);
    }
    for (int k = 0; k < count; ++k) {
        CarState& car = a.cars[VisitSlot(env, k, count)];
        if (!car.is_demoed && car.driving_disabled) --car.driving_disabled;
    }
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
    env.frame_ticks += 1;
    if (p.frame_frac >= 0.0f) {
        const float frame_dt = (static_cast<float>(env.frame_ticks) + p.frame_frac - env.frame_frac_prev) * p.tick_dt;
        env.frame_ticks = 0;
        env.frame_frac_prev = p.frame_frac;
        const bool counting = env.countdown_phase != 0;
        CountdownUpdateTimers(env, ball, frame_dt);
        // The countdown's end hands driving back to every car (env_step.hpp:6516-6538).
        if (counting && env.countdown_phase == 0) {
            for (int i = 0; i < count; ++i) a.cars[i].driving_disabled = 0;
        }
    }
    env.tick += 1;
    st.unsupported = unsupported;
}

}  // namespace core
}  // namespace rg
