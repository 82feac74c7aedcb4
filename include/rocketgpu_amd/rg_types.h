// Public state records of the batched arena step.  Every record here is byte-compatible with the reference's
// public structs (include/rocketgpu/rocketgpu.h:36-706): same field order, types, defaults and therefore the same
// sizes (CarState 880 B, BallState 288 B, EnvState 44 B, BoostPadState 16 B, CarControls 24 B, CarConfig 120 B,
// EnvCommand 16 B, PhysicsEvent 104 B, PhysicsEventWindow 16 B), so arrays downloaded from either implementation can
// be compared byte for byte and fed to the other.  Field names follow the reference because state_layouts() and
// download_state_array() expose them.
//
// The records are plain aggregates usable from host C++ and from gfx950 device code.
#pragma once

#include <cstddef>
#include <cstdint>

#if defined(__HIPCC__)
#define RG_HD __host__ __device__ inline
#else
#define RG_HD inline
#endif

namespace rg {

struct Vec3 {
    float x = 0.0f;
    float y = 0.0f;
    float z = 0.0f;
    RG_HD Vec3() = default;
    RG_HD constexpr Vec3(float x_, float y_, float z_) : x(x_), y(y_), z(z_) {}
};

struct Quat {
    float w = 1.0f;
    float x = 0.0f;
    float y = 0.0f;
    float z = 0.0f;
    RG_HD Quat() = default;
    RG_HD constexpr Quat(float w_, float x_, float y_, float z_) : w(w_), x(x_), y(y_), z(z_) {}
};

enum class Team : uint8_t { Blue = 0, Orange = 1 };

struct CarControls {
    float throttle = 0.0f;
    float steer = 0.0f;
    float pitch = 0.0f;
    float yaw = 0.0f;
    float roll = 0.0f;
    uint8_t jump = 0;
    uint8_t boost = 0;
    uint8_t handbrake = 0;
};

// Car body and handling preset (Octane defaults).  See the reference's CarConfig (rocketgpu.h:132-173) for the
// meaning of each number; the values are the game's published hitboxes plus Bullet's 2.5 uu margin.
struct CarConfig {
    Vec3 hitbox_size = Vec3(118.00738f + 2.5f, 84.19941f + 2.5f, 36.159073f + 2.5f);
    Vec3 hitbox_published = Vec3(118.00738f, 84.19941f, 36.159073f);
    Vec3 hitbox_offset = Vec3(13.87566f, 0.0f, 20.754988f);
    Vec3 front_wheels_offset = Vec3(51.25f, 25.90f, 20.754988f);
    Vec3 back_wheels_offset = Vec3(-33.75f, 29.50f, 20.754988f);
    float front_wheel_radius = 12.50f;
    float back_wheel_radius = 15.00f;
    float front_suspension_rest = 38.75499f;
    float back_suspension_rest = 37.054989f;
    float dodge_deadzone = 0.5f;
    uint8_t three_wheels = 0;
    uint16_t body = 147;
    float steer_angle_curve_deg[6] = {30.570510864257812f, 18.294261932373047f, 10.429570198059082f,
                                      6.0558881759643555f, 4.874049186706543f, 1.9791786670684814f};
    float powerslide_steer_angle_curve_deg[2] = {22.478458404541016f, 7.224311351776123f};
    float spawn_z = 36.02f;
};

struct BallHitInfo {
    uint8_t is_valid = 0;
    uint8_t extra_pending = 0;
    uint8_t extra_tick_is_previous = 0;
    uint8_t reserved2 = 0;
    uint64_t tick_count_when_hit = 0;
    uint64_t tick_count_when_extra_impulse_applied = 0;
    uint32_t hit_car_id = 0;
    Vec3 relative_pos_on_ball = Vec3(0.0f, 0.0f, 0.0f);
    Vec3 ball_pos = Vec3(0.0f, 0.0f, 0.0f);
    Vec3 extra_hit_vel = Vec3(0.0f, 0.0f, 0.0f);
};

struct CarState {
    Vec3 pos = Vec3(0.0f, 0.0f, 17.0f);
    Quat rot = Quat(1.0f, 0.0f, 0.0f, 0.0f);
    Vec3 vel = Vec3(0.0f, 0.0f, 0.0f);
    Vec3 ang_vel = Vec3(0.0f, 0.0f, 0.0f);
    Vec3 flip_rel_torque = Vec3(0.0f, 0.0f, 0.0f);
    float jump_time = 0.0f;
    float flip_time = 0.0f;
    float air_time = 0.0f;
    float air_time_since_jump = 0.0f;
    float boost = 100.0f;
    float time_since_boosted = 0.0f;
    float boosting_time = 0.0f;
    float supersonic_time = 0.0f;
    float handbrake_val = 0.0f;
    float auto_flip_timer = 0.0f;
    float boost_amount = 1.0f;
    float demo_respawn_timer = 0.0f;
    uint64_t tick_count_since_update = 0;
    uint8_t ball_contact_last_tick = 0;
    uint8_t ball_contact_this_tick = 0;
    uint8_t is_on_ground = 1;
    uint8_t wheels_with_contact[4] = {1, 1, 1, 1};
    uint8_t has_jumped = 0;
    uint8_t has_double_jumped = 0;
    uint8_t has_flipped = 0;
    uint8_t is_flipping = 0;
    uint8_t is_jumping = 0;
    uint8_t is_boosting = 0;
    uint8_t is_supersonic = 0;
    uint8_t is_auto_flipping = 0;
    uint8_t is_demoed = 0;
    uint8_t just_respawned = 0;
    uint8_t demo_no_contact_response = 0;
    uint8_t respawn_slot_override_plus1 = 0;
    uint8_t rlbotv5_boost_hold = 0;
    struct {
        uint8_t has_contact = 0;
        uint8_t wheel_ball_mask = 0;
        uint8_t reserved1 = 0;
        uint8_t reserved2 = 0;
        Vec3 contact_normal = Vec3(0.0f, 0.0f, 0.0f);
    } world_contact;
    struct {
        uint32_t other_car_id = 0;
        float cooldown_timer = 0.0f;
    } car_contact;
    Vec3 vel_impulse_cache = Vec3(0.0f, 0.0f, 0.0f);
    BallHitInfo ball_hit_info = {};
    CarControls last_controls = {};
    Vec3 pos_bt = Vec3(0.0f, 0.0f, 17.0f * (1.0f / 50.0f));
    Vec3 vel_bt = Vec3(0.0f, 0.0f, 0.0f);
    Vec3 tick_start_pos_bt = Vec3(0.0f, 0.0f, 17.0f * (1.0f / 50.0f));
    Vec3 pad_frame_pos_bt = Vec3(0.0f, 0.0f, 17.0f * (1.0f / 50.0f));
    uint8_t vel_touched_solve_stage = 1;
    uint8_t basis_authoritative = 0;
    uint8_t ext_force_impulse_bt_valid = 0;
    uint8_t pad_frame_pos_valid = 0;
    Vec3 basis_forward = Vec3(1.0f, 0.0f, 0.0f);
    Vec3 basis_right = Vec3(0.0f, 1.0f, 0.0f);
    Vec3 basis_up = Vec3(0.0f, 0.0f, 1.0f);
    Vec3 ext_force_impulse_bt = Vec3(0.0f, 0.0f, 0.0f);
    Vec3 ext_torque_impulse = Vec3(0.0f, 0.0f, 0.0f);
    Vec3 solve_delta_linear_bt = Vec3(0.0f, 0.0f, 0.0f);
    Vec3 vel_impulse_cache_bt = Vec3(0.0f, 0.0f, 0.0f);
    Vec3 solve_delta_angular = Vec3(0.0f, 0.0f, 0.0f);
    uint8_t bump_event_bumps[7] = {};
    uint8_t bump_event_demos[7] = {};
    uint8_t boost_pad_touching = 0;
    uint8_t bump_event_reserved1 = 0;
    uint64_t previous_ball_contact_start = ~0ULL;
    uint64_t last_ball_contact_start = ~0ULL;
    uint64_t last_world_surface_contact = ~0ULL;
    uint64_t first_world_after_previous_ball = ~0ULL;
    uint64_t first_world_after_last_ball = ~0ULL;
    uint64_t last_flip_or_jump_start = ~0ULL;
    uint64_t last_directional_flip_start = ~0ULL;
    uint8_t last_ball_started_with_world = 0;
    uint8_t flip_or_jump_last_tick = 1;
    float internal_jump_time = 0.0f;
    float internal_flip_time = 0.0f;
    uint8_t flip_timer_live = 0;
    uint8_t air_state = 0;
    uint16_t driving_disabled = 0;
    float dodge_timeout = -1.0f;
    float dodge_elapsed = 0.0f;
    float dodge_dir_x = 0.0f;
    float dodge_dir_y = 0.0f;
    uint64_t latest_touch_tick = ~0ULL;
    Vec3 latest_touch_location = Vec3(0.0f, 0.0f, 0.0f);
    Vec3 latest_touch_normal = Vec3(0.0f, 0.0f, 0.0f);
    Vec3 internal_vel = Vec3(0.0f, 0.0f, 0.0f);
    Vec3 internal_ang_vel = Vec3(0.0f, 0.0f, 0.0f);
    Vec3 reported_vel = Vec3(0.0f, 0.0f, 0.0f);
    Vec3 reported_ang_vel = Vec3(0.0f, 0.0f, 0.0f);
    uint8_t internal_has_jumped = 0;
    uint8_t internal_has_double_jumped = 0;
    uint8_t internal_has_flipped = 0;
    uint8_t internal_is_supersonic = 0;
    uint8_t reported_flags = 0;
    uint8_t game_has_jumped = 0;
    uint8_t double_jump_ticks = 255;
    uint8_t jump_hold_broken = 0;
    uint8_t supersonic_set_v3 = 0;
    uint8_t rlbotv5_dodged = 0;
    Vec3 set_wheel_normals[4] = {};
    uint8_t set_wheel_normals_mask = 0;
    uint8_t cal_pre_flags = 0;
    float cal_pre_forward_speed_uu = 0.0f;
    float cal_pre_pitchlock_restart_time = -1.0f;
    Vec3 cal_tick_start_ang_vel = Vec3(0.0f, 0.0f, 0.0f);
    Vec3 tick_start_basis_forward = Vec3(1.0f, 0.0f, 0.0f);
    Vec3 tick_start_basis_right = Vec3(0.0f, 1.0f, 0.0f);
    Vec3 tick_start_basis_up = Vec3(0.0f, 0.0f, 1.0f);
    Vec3 restored_inv_inertia_bt[3] = {};
    uint8_t restored_inv_inertia_valid = 0;
    static constexpr int kMaxPendingBumps = 7;
    Vec3 pending_bump_bt[kMaxPendingBumps] = {};
    uint32_t pending_bump_hitter[kMaxPendingBumps] = {};
    uint8_t pending_bump_count = 0;
    float auto_flip_torque_scale = 0.0f;
};

struct BallState {
    struct HeatseekerInfo {
        float y_target_dir = 0.0f;
        float cur_target_speed = 2900.0f;
        float time_since_hit = 0.0f;
    } hs_info = {};
    struct DropshotInfo {
        int32_t charge_level = 1;
        float accumulated_hit_force = 0.0f;
        float y_target_dir = 0.0f;
        uint8_t has_damaged = 0;
        uint8_t reserved0 = 0;
        uint8_t reserved1 = 0;
        uint8_t reserved2 = 0;
        uint64_t last_damage_tick = 0;
    } ds_info = {};
    Vec3 pos = Vec3(0.0f, 0.0f, 93.15f);
    Quat rot = Quat(1.0f, 0.0f, 0.0f, 0.0f);
    Vec3 vel = Vec3(0.0f, 0.0f, 0.0f);
    Vec3 ang_vel = Vec3(0.0f, 0.0f, 0.0f);
    Vec3 vel_impulse_cache = Vec3(0.0f, 0.0f, 0.0f);
    uint64_t tick_count_when_extra_impulse_applied = 0;
    uint64_t tick_count_since_update = 0;
    Vec3 pos_bt = Vec3(0.0f, 0.0f, 93.15f * (1.0f / 50.0f));
    Vec3 vel_bt = Vec3(0.0f, 0.0f, 0.0f);
    Vec3 basis_forward = Vec3(1.0f, 0.0f, 0.0f);
    Vec3 basis_right = Vec3(0.0f, 1.0f, 0.0f);
    Vec3 basis_up = Vec3(0.0f, 0.0f, 1.0f);
    uint32_t basis_authoritative = 0;
    Vec3 inertia_basis_forward = Vec3(1.0f, 0.0f, 0.0f);
    Vec3 inertia_basis_right = Vec3(0.0f, 1.0f, 0.0f);
    Vec3 inertia_basis_up = Vec3(0.0f, 0.0f, 1.0f);
    uint32_t inertia_basis_valid = 0;
    Vec3 restored_inv_inertia_bt[3] = {};
    uint32_t restored_inv_inertia_valid = 0;
    Vec3 vel_impulse_cache_bt = Vec3(0.0f, 0.0f, 0.0f);
    // 0 awake, 1 asleep, 2 just uploaded (resolved at the next tick start); bit 4: woken by a wheel this step.
    uint32_t sleeping = 0;
    float deactivation_time = 0.0f;
    uint16_t constrained_2d = 0;
    uint16_t reserved_c2d = 0;
};

struct BoostPadState {
    float cooldown = 0.0f;
    uint8_t is_active = 1;
    uint8_t is_big = 0;
    uint16_t pickup_frac_q = 32768;
    uint32_t prev_locked_car_id = 0;
    uint32_t cur_locked_car_id = 0;
};

struct EnvState {
    uint32_t tick = 0;
    uint32_t blue_score = 0;
    uint32_t orange_score = 0;
    int32_t last_scored_team = -1;
    uint8_t goal_scored_this_tick = 0;
    // Car visit order: 0 slot order, 1 packed order as set, 2 reverse slot order, 3 the game's vehicle order.
    uint8_t car_visit_order_mode = 3;
    uint8_t car_visit_present_mask = 0xFFu;
    uint8_t kickoff_resets = 0;
    uint32_t car_visit_order_packed = 0x76543210u;
    uint8_t countdown_phase = 0;
    uint8_t countdown_fires = 0;
    uint16_t frame_ticks = 0;
    float countdown_count = 0.0f;
    float frame_frac_prev = 0.5f;
    uint16_t kickoff_match_order = 0;
    uint8_t kickoff_match_index = 0;
    uint8_t goal_latched = 0;
    uint8_t respawn_match_order = 0;
};

struct EnvCommand {
    uint8_t reset_to_kickoff = 0;
    uint8_t reset_scoreboard = 0;
    uint8_t reset_tick_count = 0;
    uint8_t kickoff_seed_mode = 0;
    uint8_t new_round_bodies = 0;
    uint8_t keep_boost_pads = 0;
    uint16_t kickoff_countdown_ticks = 0;
    uint64_t kickoff_seed = 0;
};

// Per-car-slot wheel and vehicle runtime state carried across ticks (the reference's detail::WheelRuntimeState,
// include/rocketgpu/detail/step/arena_geometry.hpp:1545-1578; BatchSim::download_wheel_runtime_states,
// rocketgpu.h:1037), 292 bytes.
struct WheelRuntimeState {
    int32_t event_world_ids[4] = {-1, -1, -1, -1};
    Vec3 event_world_normals[4] = {};
    float steer_angle[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    float engine_force[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    float brake[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    float lat_friction[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    float long_friction[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    float extra_pushback[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    float solver_time_step = 1.0f / 60.0f;
    Vec3 accumulated_force = Vec3(0.0f, 0.0f, 0.0f);
    Vec3 accumulated_ang_accel = Vec3(0.0f, 0.0f, 0.0f);
    Vec3 accumulated_torque_bt = Vec3(0.0f, 0.0f, 0.0f);
    Quat pre_solve_rot = Quat(1.0f, 0.0f, 0.0f, 0.0f);
    Vec3 pre_solve_basis_forward = Vec3(1.0f, 0.0f, 0.0f);
    Vec3 pre_solve_basis_right = Vec3(0.0f, 1.0f, 0.0f);
    Vec3 pre_solve_basis_up = Vec3(0.0f, 0.0f, 1.0f);
    uint8_t pre_solve_basis_authoritative = 0;
    Vec3 cal_prev_sticky_dir = Vec3(0.0f, 0.0f, 0.0f);
    Vec3 cal_prev_wheel_normal_sum = Vec3(0.0f, 0.0f, 0.0f);
    uint8_t cal_prev_world_wheel_contacts = 0;
    uint8_t cal_game_jumped = 0;
    uint8_t cal_game_double_jumped = 0;
    uint8_t cal_cached_ground = 0;
    float cal_jump_inactive_time = 0.0f;
    uint8_t cal_frame_open = 0;
    uint8_t cal_restored = 0;
};

// Physics event stream (rocketgpu.h:45-98).  The caller owns both buffers: records holds capacity_per_env records
// per env, windows one PhysicsEventWindow per env.  The step sets each env's window.tick to the env's tick at the
// start of every tick and appends that tick's events after window.count; the caller clears count and overflow
// between reads.  A full window sets overflow and drops the event.
enum class PhysicsEventType : uint8_t { BallTouchStart, BallWorldContact, CarWorldContact, BoostPickup, CarBump, Goal };

struct PhysicsEvent {
    uint64_t tick = 0;
    PhysicsEventType type = PhysicsEventType::BallTouchStart;
    uint32_t car_id = 0;
    uint32_t other_car_id = 0;
    int32_t team = -1;
    uint32_t pad_index = 0;
    uint8_t car_on_ground = 0;
    uint8_t is_big_pad = 0;
    uint8_t is_demo = 0;
    Vec3 position = {};
    Vec3 normal = {};
    Vec3 car_velocity = {};
    Vec3 ball_position = {};
    Vec3 ball_velocity = {};
    float boost_before = 0;
    float boost_after = 0;
};

struct PhysicsEventWindow {
    uint64_t tick = 0;
    uint32_t count = 0;
    uint32_t overflow = 0;
};

static_assert(sizeof(CarControls) == 24, "CarControls layout");
static_assert(sizeof(CarConfig) == 120, "CarConfig layout");
static_assert(sizeof(CarState) == 880, "CarState layout");
static_assert(sizeof(BallState) == 288, "BallState layout");
static_assert(sizeof(EnvState) == 44, "EnvState layout");
static_assert(sizeof(BoostPadState) == 16, "BoostPadState layout");
static_assert(sizeof(EnvCommand) == 16, "EnvCommand layout");
static_assert(sizeof(PhysicsEvent) == 104, "PhysicsEvent layout");
static_assert(sizeof(WheelRuntimeState) == 292, "WheelRuntimeState layout");
static_assert(sizeof(PhysicsEventWindow) == 16, "PhysicsEventWindow layout");

}  // namespace rg
