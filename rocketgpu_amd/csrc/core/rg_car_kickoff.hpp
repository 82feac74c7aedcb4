// Kickoff resets and env commands of arenas with cars: which kickoff spot each car takes, the spawn itself (placed
// at once, or pending for the countdown's game spawn), and the settled kickoff rows that replace a placed car by the
// state it reaches after a countdown.  Follows ResetEnvToKickoff and its helpers (include/rocketgpu/detail/step/
// arena_ball.hpp:2698-3083), ApplyEnvCommandBodies / ApplyEnvCommand (env_step.hpp:5674-5755) and the host path of
// apply_env_commands (src/rocketgpu.cpp:2136-2205).  The ball-only pieces (env state, ball, pads) are those of
// rg_env_commands.hpp.
#pragma once

#include "rg_arena_step.hpp"

namespace rg {
namespace core {

// The settled kickoff rows of a batch (src/rocketgpu.cpp:1439-1612).  Cars are grouped by (CarConfig, team); for
// each group `spawns` rows hold the car and wheel state a car of the group reaches at spawn index k (kickoff spot k
// for k < 5) when a countdown kickoff ends, and one ball row the settled ball.
struct KickoffRows {
    const CarState* cars;             // [group * spawns + k]
    const WheelRuntimeState* wheels;  // [group * spawns + k]
    const int32_t* group;             // per batch car slot
    const BallState* ball;
    int spawns;
};

// Where a kickoff puts car slot i of the env: its team, its rank among its team's cars in visit order, and the spot
// (the plan's spot for the first five, else the rank).  False when the slot is not visited.
RG_HD bool KickoffSpotOf(const EnvState& env, const uint8_t* teams, int count, int i, const KickoffPlan& plan,
                         int& team, int& nth, int& spot) {
    team = teams[i] == 0 ? 0 : 1;
    nth = 0;
    for (int k = 0; k < count; ++k) {
        const int s = VisitSlot(env, k, count);
        if (s == i) {
            spot = nth < 5 ? static_cast<int>((plan.order_bits[team] >> (3 * nth)) & 7u) : nth;
            return true;
        }
        if ((teams[s] == 0 ? 0 : 1) == team) ++nth;
    }
    return false;
}

// The spawn of a kickoff car (PlaceKickoffCar, arena_ball.hpp:3006-3039, the arena-rules branch): kickoff spot
// `spot` for the team's first five cars, respawn spot (nth - 5) % 4 for the rest, every further four of them 250 uu
// further back along the spot's facing.  A game spawn (countdown) leaves the car demolished for 1.5 ticks with the
// kickoff spot as its respawn override, so the countdown's first pre-tick re-creates it there.
RG_HD void PlaceKickoffCar(CarState& car, const CarSpawnRules& rules, int team, int nth, int spot, float spawn_z,
                           float boost, float tick_dt, bool game_spawn) {
    Vec3 pos;
    float yaw;
    if (nth < 5) {
        pos = rules.kickoff_pos[team][spot];
        yaw = rules.kickoff_yaw[team][spot];
    } else {
        const int r = (nth - 5) % 4;
        const float back = 250.0f * static_cast<float>((nth - 5) / 4);
        pos = rules.respawn_pos[team][r];
        yaw = rules.respawn_yaw[team][r];
        pos = pos - Vec3(GameCosf(yaw), GameSinf(yaw), 0.0f) * back;
    }
    PlaceCarAtWorldSpawn(car, pos, yaw, spawn_z, boost);
    if (game_spawn && nth < 5) {
        constexpr int kKickoffSpawnOverridePlus1 = 5;
        car.is_demoed = 1;
        car.demo_respawn_timer = 1.5f * tick_dt;
        car.respawn_slot_override_plus1 = static_cast<uint8_t>(kKickoffSpawnOverridePlus1 + spot);
    }
}

// A placed kickoff car replaced by its settled row, keeping the fields that outlive a kickoff (touch, bump and
// flip history; SettleKickoffCar, arena_ball.hpp:2757-2767); the car's wheel state becomes the row's.
RG_HD void SettleKickoffCar(CarState& car, WheelRuntimeState& wheels, const CarState& row,
                            const WheelRuntimeState& row_wheels) {
    const CarState kept = car;
    car = row;
    for (int k = 0; k < 7; ++k) {
        car.bump_event_bumps[k] = kept.bump_event_bumps[k];
        car.bump_event_demos[k] = kept.bump_event_demos[k];
    }
    car.previous_ball_contact_start = kept.previous_ball_contact_start;
    car.last_ball_contact_start = kept.last_ball_contact_start;
    car.last_world_surface_contact = kept.last_world_surface_contact;
    car.first_world_after_previous_ball = kept.first_world_after_previous_ball;
    car.first_world_after_last_ball = kept.first_world_after_last_ball;
    car.last_flip_or_jump_start = kept.last_flip_or_jump_start;
    car.last_directional_flip_start = kept.last_directional_flip_start;
    car.last_ball_started_with_world = kept.last_ball_started_with_world;
    car.flip_or_jump_last_tick = kept.flip_or_jump_last_tick;
    car.latest_touch_tick = kept.latest_touch_tick;
    car.latest_touch_location = kept.latest_touch_location;
    car.latest_touch_normal = kept.latest_touch_normal;
    wheels = row_wheels;
}

// One env's kickoff (ResetEnvToKickoff): env state and spot plan, ball, pads unless kept, every control slot of the
// env cleared, each car spawned; without a countdown, a batch with settled rows gets the settled ball and cars.
// `a.controls` is not written; `controls` is the env's max_cars control slots.
RG_HD void ResetEnvToKickoffCars(CarArenaRefs& a, CarControls* controls, int max_cars, uint64_t seed,
                                 bool preserve_score, uint8_t seed_mode, bool keep_pads, bool game_spawn,
                                 const KickoffRows* rows, const CarSpawnRules& rules, const StepParams& p,
                                 const CarParams& m) {
    Vec3 spots[5];
    for (int k = 0; k < 5; ++k) spots[k] = Vec3(rules.kickoff_pos[0][k].x, rules.kickoff_pos[0][k].y, 0.0f);
    const KickoffPlan plan = ResetKickoffEnvState(*a.env, a.env_idx, seed, preserve_score, seed_mode, a.teams,
                                                  a.count, spots);
    ResetKickoffBall(*a.ball, p, game_spawn);
    if (!game_spawn && rows != nullptr) *a.ball = *rows->ball;
    if (!keep_pads) ResetBoostPads(a.pads, p);
    for (int i = 0; i < max_cars; ++i) controls[i] = CarControls{};
    for (int i = 0; i < a.count; ++i) {
        int team, nth, spot;
        if (!KickoffSpotOf(*a.env, a.teams, a.count, i, plan, team, nth, spot)) continue;
        PlaceKickoffCar(a.cars[i], rules, team, nth, spot, a.cfgs[i].spawn_z, m.boost_spawn_amount, m.tick_dt,
                        game_spawn);
        if (!game_spawn && rows != nullptr) {
            const int r = rows->group[a.car_base + i] * rows->spawns + spot;
            SettleKickoffCar(a.cars[i], a.wheels[i], rows->cars[r], rows->wheels[r]);
        }
    }
}

// The bodies' side of a command with cars (ApplyEnvCommandBodies, BulletSweepNewRound): a new match starts a fresh
// sweep; a reset or a new round takes demolished cars, then the ball and every car out of the sweep, and empties
// the ball's manifolds.
RG_HD void ApplyCommandBodiesCars(ArenaInternal& st, const EnvCommand& c, const CarState* cars, int count) {
    if (c.reset_to_kickoff && c.reset_scoreboard) {
        st.bp = BroadphaseState{};
        return;
    }
    if (!c.reset_to_kickoff && !c.new_round_bodies) return;
    if (st.bp.sweep.handle_count != 0) {
        const int n = count < kMaxArenaCars ? count : kMaxArenaCars;
        Sweep s(st.bp, n + 1);
        for (int i = 0; i < n; ++i) {
            if (cars[i].is_demoed) s.RemoveBody(i + 1);
        }
        s.RemoveBody(0);
        for (int i = 0; i < n; ++i) s.RemoveBody(i + 1);
    }
    ResetBallManifolds(st);
}

// The env side of a command with cars (ApplyEnvCommand): after a kickoff reset every visited car and the ball
// carry the countdown as their driving lock and 2-D constraint.
RG_HD void ApplyCommandEnvCars(CarArenaRefs& a, CarControls* controls, int max_cars, const EnvCommand& c,
                               const KickoffRows* rows, const CarSpawnRules& rules, const StepParams& p,
                               const CarParams& m) {
    if (!c.reset_to_kickoff) return;
    ResetEnvToKickoffCars(a, controls, max_cars, c.kickoff_seed, c.reset_scoreboard == 0, c.kickoff_seed_mode,
                          c.keep_boost_pads != 0, c.kickoff_countdown_ticks != 0, rows, rules, p, m);
    for (int k = 0; k < a.count; ++k) a.cars[VisitSlot(*a.env, k, a.count)].driving_disabled = c.kickoff_countdown_ticks;
    a.ball->constrained_2d = c.kickoff_countdown_ticks;
    ApplyCommandEnvFields(*a.env, c);
}

// The per-arena state a kickoff reset leaves: fresh except the sweep.  The unsupported word is cleared only when
// every car took a settled row (its wheel state then comes from the row, not from the ticks before).
RG_HD void ClearAfterKickoff(ArenaInternal& st, bool settled) {
    const BroadphaseState bp = st.bp;
    const uint32_t unsupported = st.unsupported;
    ClearArenaInternal(st);
    st.bp = bp;
    if (!settled) st.unsupported = unsupported;
}

// apply_env_commands on one env with cars (src/rocketgpu.cpp:2173-2205).
RG_HD void ApplyEnvCommandCars(CarArenaRefs& a, CarControls* controls, int max_cars, const EnvCommand& c,
                               const KickoffRows* rows, const CarSpawnRules& rules, const StepParams& p,
                               const CarParams& m) {
    if (!CommandActive(c)) return;
    ApplyCommandBodiesCars(*a.st, c, a.cars, a.count);
    ApplyCommandEnvCars(a, controls, max_cars, c, rows, rules, p, m);
    if (c.reset_to_kickoff) ClearAfterKickoff(*a.st, rows != nullptr && c.kickoff_countdown_ticks == 0);
}

// BatchSim::reset_to_kickoff on one env with cars (src/rocketgpu.cpp:2085-2133): fresh per-arena state, a
// new-match reset without a countdown.
RG_HD void ResetToKickoffCars(CarArenaRefs& a, CarControls* controls, int max_cars, uint64_t seed,
                              const KickoffRows* rows, const CarSpawnRules& rules, const StepParams& p,
                              const CarParams& m) {
    const uint32_t unsupported = a.st->unsupported;
    ClearArenaInternal(*a.st);
    if (rows == nullptr) a.st->unsupported = unsupported;
    ResetEnvToKickoffCars(a, controls, max_cars, seed, false, 0, false, false, rows, rules, p, m);
}

}  // namespace core
}  // namespace rg
