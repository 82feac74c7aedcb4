// Env commands and kickoff resets of arenas without cars: the round reset of a ball arena (ball back to the kickoff
// spot, scoreboard / tick / countdown fields of the env, pads) and the broadphase's new round.  Follows the
// reference's ApplyEnvCommand / ApplyEnvCommandBodies / ApplyEnvCommandEnvState (include/rocketgpu/detail/step/
// env_step.hpp:5674-5755), ResetEnvToKickoffState / ResetKickoffBall / ResetEnvToKickoff (detail/step/
// arena_ball.hpp:2845-3083) and the per-tick command path StepEnvPreTickEnv (env_step.hpp:6700-6740), with the car
// loops empty.  The kickoff and respawn spot orders are drawn from the same integer generator so that an env's
// EnvState after a reset is byte-identical to the reference's.
#pragma once

#include "rg_ball_step.hpp"

namespace rg {
namespace core {

constexpr uint16_t kKickoffCountdownGameTimers = 0xFFFFu;  // rocketgpu.h:635
constexpr uint16_t kKickoffCountdownTicks = 480;           // rocketgpu.h:632

// Seed mixing of one env's draw stream (arena_ball.hpp:2833-2838).
RG_HD uint32_t MixEnvSeed(uint64_t seed, int env_idx, int extra) {
    uint64_t x = seed;
    x ^= static_cast<uint64_t>(static_cast<uint32_t>(env_idx) + 0x9e3779b9u) + (x << 6u) + (x >> 2u);
    x ^= static_cast<uint64_t>(static_cast<uint32_t>(extra) + 0x85ebca6bu) + (x << 13u) + (x >> 7u);
    return static_cast<uint32_t>(x & 0xffffffffu);
}

// 32-bit linear congruential step (Numerical Recipes' constants; arena_ball.hpp:2600-2602).
RG_HD uint32_t NextLcg(uint32_t state) { return state * 1664525u + 1013904223u; }

// Index in [0, n) from the top 15 bits of the draw, in float as the reference computes it.
RG_HD int LcgPick(uint32_t state, int n) {
    return static_cast<int>(static_cast<float>((state >> 16) & 0x7fffu) * (1.0f / 32768.0f) * static_cast<float>(n));
}

// The five kickoff spots in draw order: pick one of the spots left, five times (arena_ball.hpp:2606-2616).
RG_HD void DrawKickoffOrder(uint64_t seed, int env_idx, int* order) {
    int left[5] = {0, 1, 2, 3, 4};
    uint32_t state = MixEnvSeed(seed, env_idx, 0x6b1d4f2a);
    for (int n = 5; n > 0; --n) {
        state = NextLcg(state);
        const int idx = LcgPick(state, n);
        order[5 - n] = left[idx];
        for (int i = idx; i + 1 < n; ++i) left[i] = left[i + 1];
    }
}

// The four respawn spots in draw order, 2 bits each (arena_ball.hpp:2618-2630).
RG_HD uint8_t DrawRespawnOrder(uint64_t seed, int env_idx) {
    uint32_t left = 0x3210u;  // the spots not drawn yet, 4 bits each
    uint32_t state = MixEnvSeed(seed, env_idx, 0x3c6ef372);
    uint32_t packed = 0;
    for (int n = 4; n > 0; --n) {
        state = NextLcg(state);
        const int idx = LcgPick(state, n);
        packed |= ((left >> (4 * idx)) & 0xFu) << (2 * (4 - n));
        left = (left & ((1u << (4 * idx)) - 1u)) | ((left >> (4 * (idx + 1))) << (4 * idx));
    }
    return static_cast<uint8_t>(packed);
}

// The car slot visited k-th of `count` (rocketgpu.h:575-626): slot order (mode 0), reversed (2), or the packed
// 4-bit order (1, and 3, the game's actor list).
RG_HD int VisitSlot(const EnvState& env, int k, int count) {
    if (env.car_visit_order_mode == 1 || env.car_visit_order_mode == 3) {
        return static_cast<int>((env.car_visit_order_packed >> (4 * k)) & 0xFu);
    }
    if (env.car_visit_order_mode == 2) return count - 1 - k;
    return k;
}

// The spot order a kickoff reset hands out per team: spot of the team's n-th car = (bits[team] >> 3n) & 7.
struct KickoffPlan {
    uint32_t order_bits[2];
};

// Of two kickoff spots (x, y), the one a short team's last car takes (KickoffSpotPreferred, arena_ball.hpp:2632):
// nearer the centre; then the one with x y < 0; then the larger |y|.
RG_HD bool KickoffSpotNearer(float ax, float ay, float bx, float by) {
    const float da = ax * ax + ay * ay, db = bx * bx + by * by;
    if (da != db) return da < db;
    const bool la = ax * ay < 0.0f, lb = bx * by < 0.0f;
    if (la != lb) return la;
    return Abs(ay) > Abs(by);
}

// EnvState of a kickoff reset (ResetEnvToKickoffState with no respawn generator and no order override,
// arena_ball.hpp:2845-2925) and the spots the env's cars take.  `teams` (0 Blue, 1 Orange) holds the env's `count`
// cars by slot and `spots` the Blue kickoff spots' ground positions; both may be null when count is 0.
RG_HD KickoffPlan ResetKickoffEnvState(EnvState& env, int env_idx, uint64_t seed, bool preserve_score,
                                       uint8_t seed_mode, const uint8_t* teams = nullptr, int count = 0,
                                       const Vec3* spots = nullptr) {
    const EnvState saved = env;
    EnvState next{};
    next.kickoff_resets = static_cast<uint8_t>(saved.kickoff_resets + 1);
    next.kickoff_match_order = saved.kickoff_match_order;
    next.kickoff_match_index = saved.kickoff_match_index;
    next.respawn_match_order = saved.respawn_match_order;
    const bool no_seed = seed_mode == 2 || (seed_mode == 0 && seed == 0);
    const bool next_kickoff = no_seed && preserve_score && (saved.kickoff_match_order & 0x8000u) != 0;
    if (no_seed) seed = saved.kickoff_resets * 0x9e3779b97f4a7c15ull;
    next.car_visit_order_mode = saved.car_visit_order_mode;
    next.car_visit_order_packed = saved.car_visit_order_packed;
    if (saved.car_visit_order_mode == 3) {
        next.car_visit_order_packed = 0x76543210u;
        next.car_visit_present_mask = 0xFFu;
    }
    next.frame_ticks = saved.frame_ticks;
    next.frame_frac_prev = saved.frame_frac_prev;
    if (preserve_score) {
        next.blue_score = saved.blue_score;
        next.orange_score = saved.orange_score;
        next.tick = saved.tick;
        next.last_scored_team = saved.last_scored_team;
    }
    // Spot order: identity for mode 3, the stored match order's next rotation for a follow-up kickoff, else drawn.
    int order[5] = {0, 1, 2, 3, 4};
    if (seed_mode != 3) {
        if (next_kickoff) {
            for (int j = 0; j < 5; ++j) {
                order[j] = (saved.kickoff_match_order >> (3 * ((saved.kickoff_match_index + j) % 5))) & 7;
            }
            next.kickoff_match_index = static_cast<uint8_t>((saved.kickoff_match_index + 1) % 5);
        } else {
            DrawKickoffOrder(seed, env_idx, order);
            uint32_t packed = 0x8000u;
            for (int j = 0; j < 5; ++j) packed |= static_cast<uint32_t>(order[j]) << (3 * j);
            next.kickoff_match_order = static_cast<uint16_t>(packed);
            next.kickoff_match_index = 1;
            next.respawn_match_order = DrawRespawnOrder(seed, env_idx);
        }
    }
    int team_order[2][5];
    for (int j = 0; j < 5; ++j) team_order[0][j] = team_order[1][j] = order[j];
    int blues = 0;
    for (int k = 0; k < count; ++k) blues += teams[VisitSlot(next, k, count)] == 0 ? 1 : 0;
    const int oranges = count - blues;
    // A follow-up kickoff with uneven teams moves the short team's last car to the best of its first four spots
    // (PreferKickoffSpotForShortTeam, arena_ball.hpp:2642-2653).
    if (next_kickoff && blues != oranges) {
        const int t = blues < oranges ? 0 : 1;
        const int n = t == 0 ? blues : oranges;
        if (n > 0 && n < 4) {
            int* o = team_order[t];
            int best = 0;
            for (int i = 1; i < 4; ++i) {
                if (KickoffSpotNearer(spots[o[i]].x, spots[o[i]].y, spots[o[best]].x, spots[o[best]].y)) best = i;
            }
            if (best >= n) {
                const int keep = o[best];
                o[best] = o[n - 1];
                o[n - 1] = keep;
            }
        }
    }
    env = next;
    KickoffPlan plan;
    for (int t = 0; t < 2; ++t) {
        plan.order_bits[t] = 0;
        for (int j = 0; j < 5; ++j) plan.order_bits[t] |= static_cast<uint32_t>(team_order[t][j]) << (3 * j);
    }
    return plan;
}

// The ball of a kickoff on an arena with rules (ResetBall with ArenaRules, arena_ball.hpp:2775-2810): at the spawn,
// asleep; a game spawn (the countdown's) starts 7.737 uu higher and awake.  Fields the reset does not name keep
// their values.
RG_HD void ResetKickoffBall(BallState& ball, const StepParams& p, bool game_spawn) {
    constexpr float kUuToBt = 1.0f / 50.0f;
    ball.rot = Quat(1.0f, 0.0f, 0.0f, 0.0f);
    ball.basis_forward = Vec3(1.0f, 0.0f, 0.0f);
    ball.basis_right = Vec3(0.0f, 1.0f, 0.0f);
    ball.basis_up = Vec3(0.0f, 0.0f, 1.0f);
    ball.basis_authoritative = 1;
    SetBallVelBt(ball, Vec3(0.0f, 0.0f, 0.0f));
    ball.ang_vel = Vec3(0.0f, 0.0f, 0.0f);
    ball.vel_impulse_cache = Vec3(0.0f, 0.0f, 0.0f);
    ball.vel_impulse_cache_bt = Vec3(0.0f, 0.0f, 0.0f);
    ball.tick_count_when_extra_impulse_applied = 0;
    ball.tick_count_since_update = 0;
    ball.hs_info = BallState::HeatseekerInfo{};
    ball.ds_info = BallState::DropshotInfo{};
    ball.deactivation_time = 0.0f;
    ball.constrained_2d = 0;
    const Vec3 s = p.ball_spawn;
    SetBallPosBt(ball, Vec3(s.x, s.y, game_spawn ? s.z + (100.487184524536f - 92.75f) : s.z) * kUuToBt);
    ball.sleeping = game_spawn ? 0u : 1u;
}

RG_HD void ResetBoostPads(BoostPadState* pads, const StepParams& p) {
    for (int k = 0; k < p.pad_count; ++k) {
        BoostPadState& pad = pads[k];
        pad.cooldown = 0.0f;
        pad.is_active = 1;
        pad.is_big = static_cast<uint8_t>((p.pad_big_mask >> k) & 1u);
        pad.prev_locked_car_id = 0;
        pad.cur_locked_car_id = 0;
        pad.pickup_frac_q = 32768;
    }
}

// Empty every persistent ball-mesh manifold (ResetBallMeshPersistentManifold over the env's slots).
RG_HD void ResetBallManifolds(ArenaInternal& st) {
    for (int m = 0; m < kMaxArenaMeshes; ++m) {
        BallMeshManifold& mf = st.ball_mesh[m];
        mf.count = 0;
        for (int k = 0; k < 4; ++k) mf.pts[k] = BallManifoldPoint{};
    }
    st.ball_mesh_retained = 0;
}

// The broadphase's round reset (BulletSweepNewRound with no cars): the ball leaves the sweep with its pairs and
// manifolds, and is inserted again at its next tick.
RG_HD void SweepNewRound(ArenaInternal& st) {
    if (st.bp.sweep.handle_count == 0) return;
    Sweep s(st.bp, 1);
    s.RemoveBody(0);
}

// Every piece of persistent per-arena state back to a fresh world (the reference's WorldSolverScratch{}).
RG_HD void ClearArenaInternal(ArenaInternal& st) {
    unsigned char* b = reinterpret_cast<unsigned char*>(&st);
    for (size_t i = 0; i < sizeof(ArenaInternal); ++i) b[i] = 0;
}

// Kickoff reset of one env (ResetEnvToKickoff): env state, ball (copied whole from `settled` when a settled kickoff
// ball exists and the reset has no countdown), pads unless kept.  The env's controls are cleared by the caller.
RG_HD void ResetEnvToKickoffNoCars(BallState& ball, EnvState& env, BoostPadState* pads, int env_idx, uint64_t seed,
                                   bool preserve_score, uint8_t seed_mode, bool keep_pads, bool game_spawn,
                                   const BallState* settled, const StepParams& p) {
    ResetKickoffEnvState(env, env_idx, seed, preserve_score, seed_mode);
    ResetKickoffBall(ball, p, game_spawn);
    if (!game_spawn && settled != nullptr) ball = *settled;
    if (!keep_pads && pads != nullptr) ResetBoostPads(pads, p);
}

RG_HD void ApplyCommandEnvFields(EnvState& env, const EnvCommand& c) {
    if (c.reset_scoreboard) {
        env.blue_score = 0;
        env.orange_score = 0;
        env.last_scored_team = -1;
    }
    if (c.reset_tick_count) env.tick = 0;
    env.goal_scored_this_tick = 0;
    env.countdown_phase = c.kickoff_countdown_ticks == kKickoffCountdownGameTimers ? 1 : 0;
    env.countdown_fires = 0;
    env.countdown_count = 0.0f;
}

RG_HD bool CommandActive(const EnvCommand& c) {
    return (c.reset_to_kickoff | c.reset_scoreboard | c.reset_tick_count | c.new_round_bodies) != 0;
}

// The bodies' side of a command (ApplyEnvCommandBodies): a new match starts a fresh sweep; a reset or a new round
// takes the ball out of the sweep and empties its manifolds.
RG_HD void ApplyCommandBodies(ArenaInternal& st, const EnvCommand& c) {
    if (c.reset_to_kickoff && c.reset_scoreboard) {
        st.bp = BroadphaseState{};
        return;
    }
    if (!c.reset_to_kickoff && !c.new_round_bodies) return;
    SweepNewRound(st);
    ResetBallManifolds(st);
}

// The env side of a command (ApplyEnvCommand): only a kickoff reset changes anything.
RG_HD void ApplyCommandEnv(BallState& ball, EnvState& env, BoostPadState* pads, int env_idx, const EnvCommand& c,
                           const BallState* settled, const StepParams& p) {
    if (!c.reset_to_kickoff) return;
    ResetEnvToKickoffNoCars(ball, env, pads, env_idx, c.kickoff_seed, c.reset_scoreboard == 0, c.kickoff_seed_mode,
                            c.keep_boost_pads != 0, c.kickoff_countdown_ticks != 0, settled, p);
    ball.constrained_2d = c.kickoff_countdown_ticks;
    ApplyCommandEnvFields(env, c);
}

// apply_env_commands on one env (the reference's host path, src/rocketgpu.cpp:2173-2205): bodies, env, and after a
// kickoff reset every piece of per-arena state but the sweep back to fresh.
RG_HD void ApplyEnvCommandNoCars(BallState& ball, EnvState& env, ArenaInternal& st, BoostPadState* pads, int env_idx,
                                 const EnvCommand& c, const BallState* settled, const StepParams& p) {
    if (!CommandActive(c)) return;
    ApplyCommandBodies(st, c);
    ApplyCommandEnv(ball, env, pads, env_idx, c, settled, p);
    if (c.reset_to_kickoff) {
        const BroadphaseState bp = st.bp;
        ClearArenaInternal(st);
        st.bp = bp;
    }
}

// A command at the start of a tick of step_with_device_commands (StepEnvPreTickEnv): bodies, env, then the ball's
// manifolds emptied again after a reset.  The tick itself follows (StepBallArena clears goal_scored_this_tick).
RG_HD void ApplyTickCommandNoCars(BallState& ball, EnvState& env, ArenaInternal& st, BoostPadState* pads,
                                  int env_idx, const EnvCommand& c, const BallState* settled, const StepParams& p) {
    ApplyCommandBodies(st, c);
    ApplyCommandEnv(ball, env, pads, env_idx, c, settled, p);
    if (c.reset_to_kickoff) ResetBallManifolds(st);
}

// BatchSim::reset_to_kickoff on one env (src/rocketgpu.cpp:2085-2133): fresh per-arena state, a new-match reset.
RG_HD void ResetToKickoffNoCars(BallState& ball, EnvState& env, ArenaInternal& st, BoostPadState* pads, int env_idx,
                                uint64_t seed, const BallState* settled, const StepParams& p) {
    ClearArenaInternal(st);
    ResetEnvToKickoffNoCars(ball, env, pads, env_idx, seed, false, 0, false, false, settled, p);
}

}  // namespace core
}  // namespace rg
