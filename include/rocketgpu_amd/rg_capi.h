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

/* set_env_car_count(host_counts, count, stream)  (rocketgpu.h:1023). */
int rg_batch_set_env_car_count(rg_batch* batch, const int32_t* host_counts, size_t count, void* stream);

/* upload_car_states / upload_ball_states / upload_env_states / upload_boost_pad_states / upload_controls
 * (rocketgpu.h:1028-1034), public-state mode: uploaded ball states are canonicalized as the reference does. */
int rg_batch_upload(rg_batch* batch, int array_id, const void* host_src, size_t count, void* stream);
/* download_*_states (rocketgpu.h:1036-1053); synchronous like the reference's non-async downloads. */
int rg_batch_download(rg_batch* batch, int array_id, void* host_dst, size_t count, void* stream);

/* step(ticks, stream)  (rocketgpu.h:1068). */
int rg_batch_step(rg_batch* batch, int32_t ticks, void* stream);
/* step_with_device_controls(device_controls_by_tick, ticks, control_stride, stream)  (rocketgpu.h:1075). */
int rg_batch_step_with_device_controls(rg_batch* batch, const void* device_controls_by_tick, int32_t ticks,
                                       int32_t control_stride, void* stream);
/* synchronize(stream)  (rocketgpu.h:1084). */
int rg_batch_synchronize(rg_batch* batch, void* stream);

/* device_view()  (rocketgpu.h:1013). */
int rg_batch_device_view(rg_batch* batch, rg_device_view* out);

/* Per-env body digests of the current state into a device array of env_count uint64 (FNV-1a over pos, vel,
 * ang_vel, rot of the ball, then each car in slot order; rocketgpu_amd/workload.h BodyDigest).  No reference
 * counterpart: it replaces downloading every state to hash it (tools/bench_cpp.cpp:505-525) for parity checks. */
int rg_batch_env_digests(rg_batch* batch, uint64_t* device_out, void* stream);

/* FNV-1a 64 over the bytes of n per-env digests (host memory): the batch digest of one tick.  Parity utility
 * with no reference counterpart. */
uint64_t rg_fold_digests(const uint64_t* host_env_digests, int64_t n);

/* Synthetic workload: `count` BallState records (default-constructed, pos / vel / ang_vel from the seeded ball
 * launches of rocketgpu_amd/workload.h) for arenas first_arena.. on PA(scale, .).  No reference counterpart; the
 * reference's benchmark draws its own inputs (tools/bench_cpp.cpp:300-360). */
int rg_workload_ball_states(uint64_t seed, int64_t first_arena, int64_t count, double scale, void* out_ball_states);

/* state_layouts()  (python/rocketgpu_bindings.cpp:541-806) as JSON into buf; returns the length needed. */
int64_t rg_state_layouts_json(char* buf, size_t capacity);

#ifdef __cplusplus
}
#endif

#endif /* ROCKETGPU_AMD_RG_CAPI_H */
