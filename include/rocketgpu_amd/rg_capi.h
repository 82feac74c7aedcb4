/* C-ABI of the MI355X batched arena step (librocketgpu_amd.so).
 *
 * The entry points replace the reference's BatchSim methods (include/rocketgpu/rocketgpu.h:988-1089) one for one,
 * in the flat form the reference's own C shim uses (rust/shim/shim.h:57-78: int status, out-pointers, a
 * thread-local last error).  Plain pointers and sizes only; streams are hipStream_t passed as void*.  State arrays
 * are the byte-compatible records of rocketgpu_amd/rg_types.h.
 *
 * Status: 0 success, nonzero failure (the message is in rg_last_error()).  Where the reference throws
 * std::invalid_argument / std::out_of_range / std::runtime_error, the C-ABI returns RG_EINVAL / RG_ERANGE /
 * RG_ERUNTIME.
 */
#ifndef ROCKETGPU_AMD_RG_CAPI_H
#define ROCKETGPU_AMD_RG_CAPI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RG_OK 0
#define RG_EINVAL 1
#define RG_ERANGE 2
#define RG_ERUNTIME 3

/* State array ids for rg_batch_upload / rg_batch_download / rg_batch_device_array. */
#define RG_ARRAY_CAR_STATES 0
#define RG_ARRAY_BALL_STATES 1
#define RG_ARRAY_ENV_STATES 2
#define RG_ARRAY_BOOST_PAD_STATES 3
#define RG_ARRAY_CONTROLS 4
#define RG_ARRAY_ENV_CAR_COUNT 5
/* WheelRuntimeState records (292 B) per car slot: upload_wheel_runtime_states / download_wheel_runtime_states
 * (rocketgpu.h:1037-1039). */
#define RG_ARRAY_WHEEL_STATES 6
/* Download only: one uint32 per env, the physics an arena with cars met that this build does not step (bit 0 a
 * car's hitbox near a mesh triangle, 1 ball-car, 2 car-car, 3 a wheel ray near another car, 4 a car's predictive
 * contact, 5 a respawn spot pick, 6 set-state wheel normals, 7 a three-wheel body); sticky until the arena is reset.
 * An arena whose word is 0 has been stepped exactly as the reference steps it.  No counterpart in the reference. */
#define RG_ARRAY_UNSUPPORTED 7

typedef struct rg_batch rg_batch;

/* Simulation config: the SimConfig fields (rocketgpu.h:842-856) this implementation reads. */
typedef struct rg_sim_config {
    int32_t env_count;
    int32_t max_cars_per_env;
    float tick_dt;
    uint8_t enable_boost_pads;
    uint8_t enable_car_car_collisions;
    uint8_t enable_demos;
    uint8_t enable_scoring;
    uint64_t respawn_rng_seed;
} rg_sim_config;

/* MutatorConfig (rocketgpu.h:708-840) as a flat record: the reference's physics and rules fields in its order
 * (its pointers, caches and calibration fields omitted), floats then flags.  rg_mutator_config_defaults fills the
 * reference's defaults.  This build honours the fields its step reads; a config that changes any other field from
 * its default is refused (RG_EINVAL, the field named by rg_last_error). */
typedef struct rg_mutator_config {
    float gravity_z;
    float ball_radius;
    float ball_mass;
    float car_mass;
    float ball_drag;
    float ball_friction;
    float ball_restitution;
    float ball_max_speed;
    float ball_max_ang_speed;
    float car_max_speed;
    float car_max_ang_speed;
    float car_air_throttle_accel;
    float boost_max;
    float boost_used_per_second;
    float boost_min_time;
    float boost_accel_ground;
    float boost_accel_air;
    float boost_spawn_amount;
    float recharge_boost_per_second;
    float recharge_boost_delay;
    float jump_accel;
    float jump_immediate_speed;
    float jump_min_time;
    float jump_reset_time_pad;
    float jump_max_time;
    float doublejump_max_delay;
    float flip_torque_time;
    float flip_torque_min_time;
    float flip_pitchlock_time;
    float flip_pitchlock_extra_time;
    float flip_initial_vel_scale;
    float flip_torque_x;
    float flip_torque_y;
    float flip_forward_impulse_max_speed_scale;
    float flip_side_impulse_max_speed_scale;
    float flip_backward_impulse_max_speed_scale;
    float flip_backward_impulse_scale_x;
    float supersonic_start_speed;
    float supersonic_maintain_min_speed;
    float supersonic_maintain_max_time;
    float powerslide_rise_rate;
    float powerslide_fall_rate;
    float arena_friction;
    float arena_restitution;
    float sticky_force;
    float ball_car_extra_impulse_z_scale;
    float ball_car_extra_impulse_forward_scale;
    float ball_car_extra_impulse_max_delta_vel;
    float ball_hit_extra_force_scale;
    float bump_force_scale;
    float goal_base_threshold_y;
    float bump_cooldown_time;
    float bump_min_forward_dist;
    float demo_respawn_time;
    float boost_pad_cooldown_small;
    float boost_pad_cooldown_big;
    float boost_pad_frame_frac;
    float frame_start_dt;
    float car_air_control_pitch;
    float car_air_control_yaw;
    float car_air_control_roll;
    float car_air_damp_pitch;
    float car_air_damp_yaw;
    float car_air_damp_roll;
    uint8_t recharge_boost_enabled;
    uint8_t unlimited_flips;
    uint8_t unlimited_double_jumps;
    uint8_t enable_team_demos;
    uint8_t countdown_on_frame_start;
    uint8_t game_mode;
    uint8_t ball_no_rot;
    uint8_t kickoff_settled;
    uint8_t reserved[8];
} rg_mutator_config;

/* A custom arena (ArenaLayout, rocketgpu.h:943-959) given as flat arrays: mesh_count meshes, mesh k owning
 * vertex_counts[k] vertices (xyz floats, uu) and index_counts[k] indices, concatenated in mesh order.  Goals: point,
 * normal (facing into the goal), min, max for team 0 then team 1.  Spawns: kickoff [team][5] and respawn [team][4]
 * as (x, y, z, yaw).  Pads: pad_count positions (xyz) and big flags. */
typedef struct rg_arena_layout {
    int32_t mesh_count;
    const int32_t* vertex_counts;
    const int32_t* index_counts;
    const float* vertices_uu;
    const int32_t* indices;
    float goals[2][12];
    float kickoff_spawns[2][5][4];
    float respawn_spawns[2][4][4];
    float ball_spawn[3];
    int32_t pad_count;
    const float* pad_positions;
    const uint8_t* pad_is_big;
} rg_arena_layout;

/* Device addresses of the state arrays (DeviceView, rocketgpu.h:961-986).  Valid until rg_batch_destroy. */
typedef struct rg_device_view {
    void* car_states;
    void* controls;
    void* car_configs;
    void* teams;
    void* ball_states;
    void* env_states;
    void* boost_pad_states;
    void* env_car_count;
} rg_device_view;

const char* rg_last_error(void);
const char* rg_version(void);

/* BatchSim(const SimConfig&, const ArenaLayout&)  (rocketgpu.h:992; src/rocketgpu.cpp:2786-2790). */
int rg_batch_create(const rg_sim_config* config, const rg_arena_layout* layout, rg_batch** out);
/* Convenience: the procedural arena PA(scale, fillet_radius) of rocketgpu_amd/workload.h as the layout. */
int rg_batch_create_procedural(const rg_sim_config* config, double scale, double fillet_radius, rg_batch** out);
/* ~BatchSim()  (rocketgpu.h:993). */
void rg_batch_destroy(rg_batch* batch);

/* car_capacity() / env_count()  (rocketgpu.h:1017-1018). */
int64_t rg_batch_env_count(const rg_batch* batch);
int64_t rg_batch_car_capacity(const rg_batch* batch);

/* The reference's MutatorConfig defaults. */
void rg_mutator_config_defaults(rg_mutator_config* out);
/* set_mutator_config(config)  (rocketgpu.h:1003): the batch's physics from the next call on; settled kickoff rows are
 * recorded again for the new config.  Also sets the settled-kickoff switch (kickoff_settled). */
int rg_batch_set_mutator_config(rg_batch* batch, const rg_mutator_config* config);
/* The batch's current MutatorConfig. */
int rg_batch_get_mutator_config(const rg_batch* batch, rg_mutator_config* out);

/* set_env_car_count(host_counts, count, stream)  (rocketgpu.h:1023). */
int rg_batch_set_env_car_count(rg_batch* batch, const int32_t* host_counts, size_t count, void* stream);

/* upload_car_states / upload_ball_states / upload_env_states / upload_boost_pad_states / upload_controls
 * (rocketgpu.h:1028-1034), public-state mode: uploaded ball states are canonicalized as the reference does. */
int rg_batch_upload(rg_batch* batch, int array_id, const void* host_src, size_t count, void* stream);
/* The same with the reference's StateUploadMode (rocketgpu.h:1027-1030): RG_UPLOAD_PUBLIC_STATE canonicalizes
 * uploaded car and ball records (rg_batch_upload's behaviour), RG_UPLOAD_PRESERVE_INTERNAL takes them as given (a
 * state downloaded from a batch resumes bit for bit). */
#define RG_UPLOAD_PUBLIC_STATE 0
#define RG_UPLOAD_PRESERVE_INTERNAL 1
int rg_batch_upload_mode(rg_batch* batch, int array_id, const void* host_src, size_t count, int32_t mode,
                         void* stream);
/* download_*_states (rocketgpu.h:1036-1053); synchronous like the reference's non-async downloads. */
int rg_batch_download(rg_batch* batch, int array_id, void* host_dst, size_t count, void* stream);
/* download_*_async (rocketgpu.h:1050-1053): the copy is queued on `stream` and not waited for; the caller
 * synchronizes the stream before reading host_dst. */
int rg_batch_download_async(rg_batch* batch, int array_id, void* host_dst, size_t count, void* stream);
/* copy_state_from(source, stream)  (rocketgpu.h:1045): every state array of `source`, a batch of the same env count,
 * car capacity, arena and physics, copied on the device (wheel runtime and internal state included). */
int rg_batch_copy_state_from(rg_batch* batch, const rg_batch* source, void* stream);

/* boost_pad_capacity()  (rocketgpu.h:1019): env_count x the layout's pad count. */
int64_t rg_batch_boost_pad_capacity(const rg_batch* batch);
/* set_teams(host_teams, count, stream)  (rocketgpu.h:1024): Team per car slot (0 Blue, 1 Orange), one byte each. */
int rg_batch_set_teams(rg_batch* batch, const uint8_t* host_teams, size_t count, void* stream);
/* set_car_configs(host_configs, count, stream)  (rocketgpu.h:1025): CarConfig records (120 B) per car slot. */
int rg_batch_set_car_configs(rg_batch* batch, const void* host_configs, size_t count, void* stream);
/* MutatorConfig::kickoff_settled (rocketgpu.h:827-832): nonzero (the default) makes a kickoff reset without a
 * countdown put the ball where the game's countdown leaves it (computed once per batch on the GPU); zero leaves it
 * asleep at the spawn, as RocketSim does. */
int rg_batch_set_kickoff_settled(rg_batch* batch, int32_t on);

/* set_physics_event_stream(stream)  (rocketgpu.h:1015; PhysicsEventStream rocketgpu.h:68-89): device buffers of
 * capacity_per_env PhysicsEvent records per env and one PhysicsEventWindow per env (rg_types.h).  All zero disables
 * the stream; otherwise env_count and cars_per_env must be the batch's and the buffers non-null (RG_EINVAL,
 * "physics event stream must match the complete batch").  Every step form (step, step_with_device_controls,
 * step_with_device_commands, step_arenas, step_graph, ball_path) then stamps window.tick with the env's tick at the
 * start of each tick and appends the tick's events: BallWorldContact for each mesh triangle the ball touches, Goal
 * for a goal (goal_only nonzero keeps only these).  The caller clears window.count and window.overflow between
 * reads; a full window sets overflow. */
int rg_batch_set_physics_event_stream(rg_batch* batch, void* device_records, void* device_windows,
                                      uint32_t capacity_per_env, uint32_t env_count, uint32_t cars_per_env,
                                      int32_t goal_only);

/* Full-state snapshot, the checkpoint form of the reference's Arena::GetFullState / SetFullState / Serialize
 * (src/rocketsim_compat_dev.cpp:369-447, include/rocketgpu/rocketsim_compat.h:293-325): one flat host blob with a
 * header naming the batch shape and arena geometry, the public records in the reference's layouts (car states,
 * controls, car configs, teams, ball states, env states, boost pads, car counts), then this build's internal
 * per-arena state (broadphase, retained ball manifolds) and the settled kickoff ball.  A loaded snapshot continues
 * bit for bit as the saved batch would have.  rg_batch_load_state rejects (RG_EINVAL) a blob of another size, shape,
 * arena or version, one whose digest does not match its data, and one holding an out-of-range index, before writing
 * anything.  Both calls synchronize `stream`. */
int rg_batch_state_snapshot_size(const rg_batch* batch, size_t* bytes);
int rg_batch_save_state(rg_batch* batch, void* host_dst, size_t bytes, void* stream);
int rg_batch_load_state(rg_batch* batch, const void* host_src, size_t bytes, void* stream);

/* reset_to_kickoff(seed, stream)  (rocketgpu.h:1065): every env to a new match's kickoff. */
int rg_batch_reset_to_kickoff(rg_batch* batch, uint64_t seed, void* stream);
/* apply_env_commands(host_commands, count, stream)  (rocketgpu.h:1066): EnvCommand i (16 B records,
 * rg_types.h) to env i for i < count. */
int rg_batch_apply_env_commands(rg_batch* batch, const void* host_commands, size_t count, void* stream);
/* apply_env_commands_device(device_commands, count, stream)  (rocketgpu.h:1067). */
int rg_batch_apply_env_commands_device(rg_batch* batch, const void* device_commands, size_t count, void* stream);
/* step_with_device_commands(controls, control_stride, env_commands, env_command_stride, ticks, stream)
 * (rocketgpu.h:1077-1083): each tick t starts with env e's command device_env_commands_by_tick[t * stride + e];
 * afterwards the last tick's controls are the batch's controls. */
int rg_batch_step_with_device_commands(rg_batch* batch, const void* device_controls_by_tick, int32_t control_stride,
                                       const void* device_env_commands_by_tick, int32_t env_command_stride,
                                       int32_t ticks, void* stream);

/* step(ticks, stream)  (rocketgpu.h:1068). */
int rg_batch_step(rg_batch* batch, int32_t ticks, void* stream);
/* step(ticks, stream) of a ball-only batch through the split schedule (launch stages 6 and 7 below) with the
 * schedule's arena lists in a fixed device array of the library instead of batch-owned memory: it allocates and
 * frees nothing.  No counterpart in the reference; it is for callers that cannot allocate device memory (code under
 * graph capture, tight serving loops).  At most 262,144 arenas (RG_EINVAL beyond, and for batches with cars).  The
 * array (2 MiB per GPU, held until the process ends) is shared by every call on one GPU: calls that overlap on one
 * GPU race. */
int rg_batch_step_split_no_alloc(rg_batch* batch, int32_t ticks, void* stream);
/* step_with_device_controls(device_controls_by_tick, ticks, control_stride, stream)  (rocketgpu.h:1075). */
int rg_batch_step_with_device_controls(rg_batch* batch, const void* device_controls_by_tick, int32_t ticks,
                                       int32_t control_stride, void* stream);
/* Steps `count` distinct arenas listed in device_arena_ids (device int32), arena i for device_ticks[i] ticks, or
 * for `ticks` ticks each when device_ticks is null.  No reference counterpart: it is the launch of the serving loop
 * below (the reference steps whole batches only, rocketgpu.h:1068). */
int rg_batch_step_arenas(rg_batch* batch, const int32_t* device_arena_ids, const int32_t* device_ticks,
                         int32_t count, int32_t ticks, void* stream);
/* launches × step(ticks_per_launch) captured once into a hipGraph and replayed on `stream` (the reference captures
 * its multi-kernel step into CUDA graphs keyed on pointers, src/rocketgpu_cuda.cu:3816-3819). */
int rg_batch_step_graph(rg_batch* batch, int32_t ticks_per_launch, int32_t launches, void* stream);
/* Launch shape of the step kernel: lanes per workgroup (64, 128 or 256); staging (0: arena records stepped in place
 * in HBM; 1: records copied into lane-private memory for the launch; 2: records in place, per-tick contact scratch
 * in LDS; 3: records and scratch in LDS; 4: 64/arenas_per_wave lanes share each arena and split its contact walks;
 * 5: as 4, with the per-tick scratch in LDS, one slice per arena; 6 and 7: as 5 with 8 lanes per arena, and 32 or
 * 16 for the arenas with the longest contact walks, see rg_batch_set_split); arenas per 64-lane wavefront (64, 32,
 * 16, 8 or 4; stages 4 and 5 take 16, 8 or 4, stages 6 and 7 take 8).  A batch with cars runs one lane per arena,
 * except at stage 2 (per-tick data in LDS, one slot per lane) and stage 4 (2 lanes per arena at 32 arenas per wave,
 * 4 at 16, 8 or 4, splitting each car's wheel-ray tree walks).  A shape whose LDS exceeds the device's
 * per-workgroup limit is rejected (RG_EINVAL) and the previous shape kept.  Results are identical for every shape;
 * the benchmark tunes it.  Defaults 64, 0, 64. */
int rg_batch_set_launch_shape(rg_batch* batch, int32_t block, int32_t stage, int32_t arenas_per_wave);
/* Stages 6 and 7 of the launch shape: before each whole-batch launch the arenas whose last contact walk loaded at
 * least `rounds` windows (default 26) go to a heavy list of at most env_count / cap_divisor arenas (default 16),
 * stepped with 32 (stage 6) or 16 (stage 7) lanes each; the rest get 8.  Results are identical for any split.  No
 * reference counterpart: a tuning knob, like the launch shape. */
int rg_batch_set_split(rg_batch* batch, int32_t rounds, int32_t cap_divisor);

/* ball_path(samples, ticks, goal_y, out, result, stream)  (rocketgpu.h:1069-1073): every arena's ball stepped in
 * place for samples x ticks ticks; pos and vel after every `ticks` ticks into device_out[env][sample][6] (f32), and
 * into device_result[env][2] the goal it enters (0 none, 1 the +y goal, 2 the -y goal) and the tick it crosses
 * |y| = goal_y on, after which it is not stepped.  The arenas must hold no cars. */
int rg_batch_ball_path(rg_batch* batch, int32_t samples, int32_t ticks, float goal_y, float* device_out,
                       int32_t* device_result, void* stream);

/* synchronize(stream)  (rocketgpu.h:1084). */
int rg_batch_synchronize(rg_batch* batch, void* stream);

/* device_view()  (rocketgpu.h:1013). */
int rg_batch_device_view(rg_batch* batch, rg_device_view* out);

/* Per-env body digests of the current state into a device array of env_count uint64 (FNV-1a over pos, vel,
 * ang_vel, rot of the ball, then each car in slot order; rocketgpu_amd/workload.h BodyDigest).  No reference
 * counterpart: it replaces downloading every state to hash it (tools/bench_cpp.cpp:505-525) for parity checks. */
int rg_batch_env_digests(rg_batch* batch, uint64_t* device_out, void* stream);

/* The CarConfig of a stock body: 0 Octane, 1 Dominus, 2 Plank, 3 Breakout, 4 Hybrid, 5 Merc, 6 Psyclops,
 * 7 Batmobile (OctaneConfig() ... BatmobileConfig(), rocketgpu.h:1098-1108); out receives one CarConfig. */
int rg_car_preset_config(int32_t preset, void* out_car_config);

/* FNV-1a 64 over the bytes of n per-env digests (host memory): the batch digest of one tick.  Parity utility
 * with no reference counterpart. */
uint64_t rg_fold_digests(const uint64_t* host_env_digests, int64_t n);

/* Synthetic workload: `count` BallState records (default-constructed, pos / vel / ang_vel from the seeded ball
 * launches of rocketgpu_amd/workload.h) for arenas first_arena.. on PA(scale, .).  No reference counterpart; the
 * reference's benchmark draws its own inputs (tools/bench_cpp.cpp:300-360). */
int rg_workload_ball_states(uint64_t seed, int64_t first_arena, int64_t count, double scale, void* out_ball_states);

/* Synthetic workload: seeded held controls (rocketgpu_amd/workload.h ControlStream: axes uniform in [-1, 1], jump /
 * boost / handbrake with probability 0.1 / p_boost / 0.05, each draw held 1..30 ticks) for cars_per_arena slots of
 * arenas first_arena..first_arena + arena_count - 1, written as CarControls [ticks][arena_count * cars_per_arena].
 * No reference counterpart. */
int rg_workload_controls(uint64_t seed, int64_t first_arena, int64_t arena_count, int32_t cars_per_arena,
                         int32_t ticks, float p_boost, void* out_controls);

/* Serving loop (no reference counterpart; the reference has no server).  `clients` simulated actors own disjoint
 * ranges of min_arenas..max_arenas resident arenas and send "step my arenas k ticks" requests (k drawn from
 * horizons) with Poisson arrivals at request_rate_hz in aggregate for duration_s.  Requests are admitted into the
 * next launch while its arena budget (max_arenas_per_launch) allows; up to max_in_flight launches overlap on
 * separate streams.  Latency is completion minus arrival. */
typedef struct rg_serve_config {
    int32_t clients;
    int32_t min_arenas;
    int32_t max_arenas;
    int32_t horizon_count;
    int32_t horizons[4];
    int32_t max_arenas_per_launch;
    int32_t max_in_flight;
    double request_rate_hz;
    double duration_s;
    uint64_t seed;
    /* Admission by HBM bytes: a launch takes requests while the arena records it will touch
     * (rg_batch_arena_bytes of each arena's car count) stay within this many bytes; 0: no byte budget. */
    int64_t launch_budget_bytes;
    /* Every reset_every-th request of a client (its requests counted from 1) first resets the client's arenas to a
     * settled kickoff with the spots in slot order (kickoff seed mode 3, a new match); 0: no resets. */
    int32_t reset_every;
    /* Nonzero: each request brings its own held controls for the client's cars (rg_workload_request_controls),
     * written into the batch's control rows on the launch's stream before the step. */
    int32_t fresh_controls;
    float p_boost;
    int32_t reserved;
} rg_serve_config;

typedef struct rg_serve_report {
    int64_t requests;
    int64_t arena_ticks;
    int64_t launches;
    int64_t arenas_resident;
    int64_t resident_bytes;
    double seconds;
    double arena_ticks_per_s;
    double p50_ms;
    double p99_ms;
    double mean_ms;
    double max_ms;
} rg_serve_report;

/* Runs the loop to completion; host_ticks_per_arena (nullable, env_count entries) receives the ticks each arena
 * was stepped. */
int rg_serve_run(rg_batch* batch, const rg_serve_config* config, rg_serve_report* report,
                 int32_t* host_ticks_per_arena);
/* The serving workload, for replaying a client's requests elsewhere: client c's arenas [*first, *first + *count),
 * and the ticks of its requests in the order they are served (returns how many; at most `capacity` are written). */
int rg_serve_client_arenas(const rg_serve_config* config, int32_t client, int32_t* first, int32_t* count);
int64_t rg_serve_client_requests(const rg_serve_config* config, int32_t client, int32_t* out_ticks, int64_t capacity);
/* The held controls of request `request` (counted from 0) for arenas [first_arena, first_arena + arena_count) with
 * cars_per_arena car slots each: out receives arena_count * cars_per_arena CarControls. */
int rg_workload_request_controls(uint64_t seed, int64_t request, int64_t first_arena, int64_t arena_count,
                                 int32_t cars_per_arena, float p_boost, void* out);
/* The HBM bytes of one arena's records a step touches with `cars` cars: ball, env, internal state, pads, car count,
 * and per car its state, wheel state, controls, config and team. */
int64_t rg_batch_arena_bytes(const rg_batch* batch, int32_t cars);

/* state_layouts()  (python/rocketgpu_bindings.cpp:541-806) as JSON into buf; returns the length needed. */
int64_t rg_state_layouts_json(char* buf, size_t capacity);

#ifdef __cplusplus
}
#endif

#endif /* ROCKETGPU_AMD_RG_CAPI_H */
