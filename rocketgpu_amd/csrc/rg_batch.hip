// BatchSim on MI355X: device buffers, the gfx950 step kernels and the C-ABI (include/rocketgpu_amd/rg_capi.h).
//
// Layout in HBM.  Arena geometry (triangles, BVH nodes, mesh infos) is uploaded once per BatchSim and read by every
// arena.  Per-arena state is three arrays indexed by arena: BallState (288 B, the public record), EnvState (44 B)
// and ArenaInternal (the broadphase sweep, pair/manifold lists and persistent ball-mesh manifolds that Bullet keeps
// across ticks).  One lane steps one arena; a launch covers every arena.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "batch_common.hpp"
#include "car_presets.hpp"
#include "mutator.hpp"
#include "core/rg_arena_step.hpp"
#include "core/rg_car_kickoff.hpp"
#include "core/rg_env_commands.hpp"
#include "group_walk.hpp"
#include "rocketgpu_amd/rg_capi.h"

namespace rg {
namespace {

thread_local std::string g_last_error;

struct InvalidArgument : std::invalid_argument {
    using std::invalid_argument::invalid_argument;
};

void Check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

template <typename T>
class DeviceArray {
public:
    DeviceArray() = default;
    DeviceArray(const DeviceArray&) = delete;
    DeviceArray& operator=(const DeviceArray&) = delete;
    ~DeviceArray() {
        if (ptr_) (void)hipFree(ptr_);
    }
    void Resize(size_t n) {
        if (ptr_) Check(hipFree(ptr_), "hipFree");
        ptr_ = nullptr;
        size_ = n;
        if (n) Check(hipMalloc(reinterpret_cast<void**>(&ptr_), n * sizeof(T)), "hipMalloc");
    }
    void Upload(const T* src, size_t n, hipStream_t s) {
        if (n > size_) throw std::out_of_range("upload larger than the device array");
        if (n) Check(hipMemcpyAsync(ptr_, src, n * sizeof(T), hipMemcpyHostToDevice, s), "hipMemcpyAsync H2D");
    }
    void Download(T* dst, size_t n, hipStream_t s) const {
        if (n > size_) throw std::out_of_range("download larger than the device array");
        if (n) Check(hipMemcpyAsync(dst, ptr_, n * sizeof(T), hipMemcpyDeviceToHost, s), "hipMemcpyAsync D2H");
    }
    T* data() const { return ptr_; }
    size_t size() const { return size_; }

private:
    T* ptr_ = nullptr;
    size_t size_ = 0;
};

// The caller's physics event buffers (PhysicsEventStream, rocketgpu.h:68-89), device addresses.
struct EventStream {
    PhysicsEvent* records = nullptr;
    PhysicsEventWindow* windows = nullptr;
    uint32_t capacity_per_env = 0;
    uint32_t env_count = 0;
    uint32_t cars_per_env = 0;
    bool goal_only = false;
};

struct StepArgs {
    BallState* balls;
    EnvState* envs;
    core::ArenaInternal* internal;
    const int32_t* env_car_count;
    core::ArenaMeshView mesh;
    core::StepParams params;
    int32_t env_count;
    // Work list: lane i steps arena arena_ids[i] (or arena i when null) for arena_ticks[i] ticks (or `ticks` when
    // null).  The host guarantees distinct ids within one launch.
    const int32_t* arena_ids;
    const int32_t* arena_ticks;
    int32_t count;
    // Arenas per 64-lane wavefront (64, 32, 16, 8 or 4): fewer arenas per wave means more waves per CU to hide the
    // latency of the dependent BVH and state loads, and less divergence within a wave.
    int32_t arenas_per_wave;
    // Env commands and kickoff resets: pads (pad_count per env), controls (max_cars per env, cleared by a reset),
    // the settled kickoff ball (null: resets leave the ball at its spawn), and per-tick commands
    // (commands[t * command_stride + env]) for step_with_device_commands.
    BoostPadState* pads;
    CarControls* controls;
    int32_t max_cars;
    const BallState* settled;
    const EnvCommand* commands;
    int32_t command_stride;
    // Physics event stream (null windows: off): env e's window is event_windows[e], its records start at
    // event_records + e * event_capacity.
    PhysicsEventWindow* event_windows;
    PhysicsEvent* event_records;
    uint32_t event_capacity;
    uint8_t event_goal_only;
};

// Env e's event sink; `writer` is false on the lanes that mirror an arena's writing lane.
__device__ inline core::EventSink EnvEvents(const StepArgs& a, int e, bool writer = true) {
    core::EventSink s;
    if (a.event_windows != nullptr) {
        s.window = a.event_windows + e;
        s.records = a.event_records + static_cast<size_t>(e) * a.event_capacity;
        s.capacity = a.event_capacity;
        s.goal_only = a.event_goal_only != 0;
        s.writer = writer;
    }
    return s;
}

// One tick of a one-lane kernel (serial mesh walks).
template <bool kEvents>
__device__ inline void Tick(BallState& ball, EnvState& env, core::ArenaInternal& st, const StepArgs& a,
                            core::BallTickScratch& scratch, const core::EventSink& events) {
    core::StepBallArena<core::SerialWalk, core::SerialWalk, kEvents>(ball, env, st, a.mesh, a.params, scratch, events);
}

constexpr int kMaxStepBlock = 256;

// Where a launch keeps an arena's records while it steps them (the launch shape's `stage`):
//   0  in place in HBM, per-tick contact scratch in lane-private memory;
//   1  records and scratch in lane-private memory (the hardware interleaves private memory across a wave's lanes,
//      so a wave's accesses to one field coalesce);
//   2  records in place, per-tick contact scratch in LDS (the hottest data: contact rows, manifold candidates);
//   3  records and scratch in LDS.
// Records are copied in before the first tick and back after the last.  LDS holds one slice per active lane of
// the workgroup: blockDim / 64 waves x arenas_per_wave lanes.
//   4  records in place, and 64 / arenas_per_wave lanes per arena: the lanes of an arena run its tick together
//      and split its mesh-tree walks (GroupWalk), which shortens the chain of dependent node loads;
//   5  as 4, with the arena's per-tick contact scratch in LDS, one slice shared by the arena's lanes (instead of a
//      private copy on every lane).
//   6  as 5 with 8 lanes per arena, except that the arenas whose last contact walk loaded the most windows (the
//      ball in a corner or along a fillet, which set the launch's duration) get 32 lanes, in the launch's first
//      workgroups; 7  the same with 16 lanes for them.  A whole-batch launch only: a subset launch runs as 5.
enum StageMode : int {
    kStageInPlace = 0,
    kStagePrivate = 1,
    kStageLdsScratch = 2,
    kStageLds = 3,
    kStageGroup = 4,
    kStageGroupLds = 5,
    kStageSplit32 = 6,
    kStageSplit16 = 7
};

constexpr bool IsSplitStage(int stage) { return stage == kStageSplit32 || stage == kStageSplit16; }
constexpr bool IsGroupStage(int stage) { return stage == kStageGroup || stage == kStageGroupLds || IsSplitStage(stage); }

// Stages 6 and 7: by default an arena whose last tick's contact walk loaded at least 26 windows (the slowest 0.5 %
// or so of the benchmark's workload, where the median is 7) is heavy, and at most env_count / 16 of them get the
// wider groups (rg_batch_set_split tunes both; a sweep of 12-40 windows and env_count / 64 to / 8 on 16,384 arenas
// put the optimum at 24-28 and / 16).
constexpr int kSplitRoundsDefault = 26;
constexpr int kSplitCapDivisorDefault = 16;

struct LdsSlice {
    core::BallTickScratch scratch;
    core::ArenaInternal st;
    BallState ball;
    EnvState env;
};

// LDS per arena of a launch (a workgroup holds blockDim / 64 x arenas_per_wave arenas).
__host__ __device__ constexpr size_t LdsBytesPerLane(int stage) {
    return stage == kStageLds ? sizeof(LdsSlice)
                              : (stage == kStageLdsScratch || stage == kStageGroupLds || IsSplitStage(stage)
                                     ? sizeof(core::BallTickScratch)
                                     : 0);
}

// One lane per arena; its ticks back to back.  kEvents: the batch has a physics event stream.
template <int kStage, bool kEvents>
__global__ void __launch_bounds__(kMaxStepBlock) StepBallArenasKernel(StepArgs a, int32_t ticks) {
    extern __shared__ __align__(16) unsigned char lds[];
    const int g = static_cast<int>(blockIdx.x) * static_cast<int>(blockDim.x) + static_cast<int>(threadIdx.x);
    const int lane = g & 63;
    if (lane >= a.arenas_per_wave) return;
    const int i = (g >> 6) * a.arenas_per_wave + lane;
    if (i >= a.count) return;
    const int e = a.arena_ids ? a.arena_ids[i] : i;
    if (e < 0 || e >= a.env_count) return;
    const int n = a.arena_ticks ? a.arena_ticks[i] : ticks;
    const int slot = (static_cast<int>(threadIdx.x) >> 6) * a.arenas_per_wave + lane;
    const core::EventSink events = EnvEvents(a, e);
    if constexpr (kStage == kStageInPlace || kStage == kStagePrivate) {
        core::BallTickScratch scratch;
        if constexpr (kStage == kStagePrivate) {
            BallState ball = a.balls[e];
            EnvState env = a.envs[e];
            core::ArenaInternal st = a.internal[e];
            for (int t = 0; t < n; ++t) Tick<kEvents>(ball, env, st, a, scratch, events);
            a.internal[e] = st;
            a.envs[e] = env;
            a.balls[e] = ball;
        } else {
            for (int t = 0; t < n; ++t) Tick<kEvents>(a.balls[e], a.envs[e], a.internal[e], a, scratch, events);
        }
    } else if constexpr (kStage == kStageLdsScratch) {
        core::BallTickScratch& scratch = reinterpret_cast<core::BallTickScratch*>(lds)[slot];
        for (int t = 0; t < n; ++t) Tick<kEvents>(a.balls[e], a.envs[e], a.internal[e], a, scratch, events);
    } else {
        LdsSlice& sl = reinterpret_cast<LdsSlice*>(lds)[slot];
        sl.ball = a.balls[e];
        sl.env = a.envs[e];
        sl.st = a.internal[e];
        for (int t = 0; t < n; ++t) Tick<kEvents>(sl.ball, sl.env, sl.st, a, sl.scratch, events);
        a.internal[e] = sl.st;
        a.envs[e] = sl.env;
        a.balls[e] = sl.ball;
    }
}

__device__ inline BoostPadState* EnvPads(const StepArgs& a, int e) {
    return a.pads ? a.pads + static_cast<size_t>(e) * a.params.pad_count : nullptr;
}

__device__ inline void ClearEnvControls(const StepArgs& a, int e) {
    for (int k = 0; k < a.max_cars; ++k) a.controls[static_cast<size_t>(e) * a.max_cars + k] = CarControls{};
}

// step_with_device_commands: each tick starts with the env's command for that tick (StepEnvPreTickEnv), then the
// tick.  Records stay in place; the work mapping is the step kernel's.
template <bool kEvents>
__global__ void __launch_bounds__(kMaxStepBlock) StepBallArenasCommandsKernel(StepArgs a, int32_t ticks) {
    const int g = static_cast<int>(blockIdx.x) * static_cast<int>(blockDim.x) + static_cast<int>(threadIdx.x);
    const int lane = g & 63;
    if (lane >= a.arenas_per_wave) return;
    const int e = (g >> 6) * a.arenas_per_wave + lane;
    if (e >= a.env_count) return;
    core::BallTickScratch scratch;
    BoostPadState* pads = EnvPads(a, e);
    const core::EventSink events = EnvEvents(a, e);
    for (int t = 0; t < ticks; ++t) {
        const EnvCommand c = a.commands[static_cast<size_t>(t) * a.command_stride + e];
        core::ApplyTickCommandNoCars(a.balls[e], a.envs[e], a.internal[e], pads, e, c, a.settled, a.params);
        if (c.reset_to_kickoff) ClearEnvControls(a, e);
        Tick<kEvents>(a.balls[e], a.envs[e], a.internal[e], a, scratch, events);
    }
}

// apply_env_commands: command e to env e, for e < count.
__global__ void __launch_bounds__(64) ApplyEnvCommandsKernel(StepArgs a, const EnvCommand* commands, int32_t count) {
    const int e = static_cast<int>(blockIdx.x) * static_cast<int>(blockDim.x) + static_cast<int>(threadIdx.x);
    if (e >= count) return;
    const EnvCommand c = commands[e];
    if (!core::CommandActive(c)) return;
    core::ApplyEnvCommandNoCars(a.balls[e], a.envs[e], a.internal[e], EnvPads(a, e), e, c, a.settled, a.params);
    if (c.reset_to_kickoff) ClearEnvControls(a, e);
}

// reset_to_kickoff: every env to a new match's kickoff.
__global__ void __launch_bounds__(64) ResetToKickoffKernel(StepArgs a, uint64_t seed) {
    const int e = static_cast<int>(blockIdx.x) * static_cast<int>(blockDim.x) + static_cast<int>(threadIdx.x);
    if (e >= a.env_count) return;
    core::ResetToKickoffNoCars(a.balls[e], a.envs[e], a.internal[e], EnvPads(a, e), e, seed, a.settled, a.params);
    ClearEnvControls(a, e);
}

// What a car arena's tick reads beyond the ball arena's: the car records, the car constants, the arena's spawn
// spots and pads, the batch seed of the respawn draws, and where the tick's controls come from.
struct CarStepArgs {
    CarState* cars;
    WheelRuntimeState* wheels;
    const CarConfig* cfgs;
    const uint8_t* teams;
    // Controls of tick t for car slot k: controls[t * control_stride + k] (stride 0: the same row every tick).
    const CarControls* controls;
    int64_t control_stride;
    core::CarParams m;
    core::CarSpawnRules rules;
    Vec3 pad_pos[rgw::kMaxPads];
    uint64_t seed;
};

// One lane per arena, its ticks back to back: arenas without cars take the ball tick, arenas with cars the full
// car tick (one lane walks all of its queries).  Workgroups are one wavefront; its first a.arenas_per_wave lanes
// take one arena each (fewer arenas per wave: more waves per CU to hide the dependent loads).  Fields the game
// reports are refreshed on the call's last tick.
// Arena e's n car ticks with `scratch` as its per-tick working data.
__device__ inline void RunCarArenaTicks(const StepArgs& a, const CarStepArgs& c, int e, int n, int count,
                                        core::CarArenaTick& scratch, int lanes = 1, int sub = 0) {
    const int base = e * a.max_cars;
    const core::PadLayout pads{c.pad_pos, a.params.pad_count};
    for (int t = 0; t < n; ++t) {
        core::CarArenaRefs r;
        r.ball = &a.balls[e];
        r.env = &a.envs[e];
        r.st = &a.internal[e];
        r.pads = a.pads + static_cast<size_t>(e) * a.params.pad_count;
        r.cars = c.cars + base;
        r.wheels = c.wheels + base;
        r.cfgs = c.cfgs + base;
        r.teams = c.teams + base;
        r.controls = c.controls + static_cast<int64_t>(t) * c.control_stride + base;
        r.count = count;
        r.env_idx = e;
        r.car_base = base;
        r.ray_lanes = lanes;
        r.ray_sub = sub;
        const core::CarRespawnContext rc{&c.rules, c.seed, e, 0u, 0u, 0};
        core::StepCarArena(r, a.mesh, a.params, c.m, pads, rc, t + 1 == n, scratch);
    }
}

__device__ inline void RunBallArenaTicks(const StepArgs& a, int e, int n, core::BallTickScratch& scratch) {
    for (int t = 0; t < n; ++t) Tick<false>(a.balls[e], a.envs[e], a.internal[e], a, scratch, core::EventSink{});
}

// kLds: each arena's per-tick working data (CarArenaTick) lives in LDS, one slot per lane, instead of lane-private
// memory.  G (1, 2 or 4): lanes per arena.  The G lanes of an arena step it together on the same data (their
// stores write equal values) and split its wheel rays' tree walks between them; arenas_per_wave * G <= 64.
template <bool kLds, int G>
__global__ void __launch_bounds__(64) StepCarArenasKernel(StepArgs a, CarStepArgs c, int32_t ticks) {
    static_assert(!kLds || G == 1, "LDS slots are one per lane");
    extern __shared__ __align__(16) unsigned char lds[];
    if (static_cast<int>(threadIdx.x) >= a.arenas_per_wave * G) return;
    const int i = static_cast<int>(blockIdx.x) * a.arenas_per_wave + static_cast<int>(threadIdx.x) / G;
    const int sub = static_cast<int>(threadIdx.x) % G;
    if (i >= a.count) return;
    const int e = a.arena_ids ? a.arena_ids[i] : i;
    if (e < 0 || e >= a.env_count) return;
    const int n = a.arena_ticks ? a.arena_ticks[i] : ticks;
    const int count = a.env_car_count[e];
    if (count == 0 && sub != 0) return;  // a ball-only arena runs on one lane
    if constexpr (kLds) {
        core::CarArenaTick& slot = reinterpret_cast<core::CarArenaTick*>(lds)[threadIdx.x];
        if (count == 0) {
            RunBallArenaTicks(a, e, n, slot.ball);
        } else {
            RunCarArenaTicks(a, c, e, n, count, slot);
        }
    } else {
        if (count == 0) {
            core::BallTickScratch scratch;
            RunBallArenaTicks(a, e, n, scratch);
            return;
        }
        core::CarArenaTick scratch;
        RunCarArenaTicks(a, c, e, n, count, scratch, G, sub);
    }
}

// Env e's records as the car cores see them (the controls are the batch's own row).
__device__ inline core::CarArenaRefs CarRefsOf(const StepArgs& a, const CarStepArgs& c, int e) {
    const int base = e * a.max_cars;
    core::CarArenaRefs r;
    r.ball = &a.balls[e];
    r.env = &a.envs[e];
    r.st = &a.internal[e];
    r.pads = EnvPads(a, e);
    r.cars = c.cars + base;
    r.wheels = c.wheels + base;
    r.cfgs = c.cfgs + base;
    r.teams = c.teams + base;
    r.controls = a.controls + base;
    r.count = a.env_car_count[e];
    r.env_idx = e;
    r.car_base = base;
    return r;
}

// reset_to_kickoff of a batch with cars: every env to a new match's kickoff; `rows.cars` null: no settled rows.
__global__ void __launch_bounds__(64) ResetToKickoffCarsKernel(StepArgs a, CarStepArgs c, core::KickoffRows rows,
                                                               uint64_t seed) {
    const int e = static_cast<int>(blockIdx.x) * static_cast<int>(blockDim.x) + static_cast<int>(threadIdx.x);
    if (e >= a.env_count) return;
    core::CarArenaRefs r = CarRefsOf(a, c, e);
    core::ResetToKickoffCars(r, a.controls + r.car_base, a.max_cars, seed, rows.cars ? &rows : nullptr, c.rules,
                             a.params, c.m);
}

// apply_env_commands of a batch with cars: command e to env e, for e < count.
__global__ void __launch_bounds__(64) ApplyEnvCommandsCarsKernel(StepArgs a, CarStepArgs c, core::KickoffRows rows,
                                                                 const EnvCommand* commands, int32_t count) {
    const int e = static_cast<int>(blockIdx.x) * static_cast<int>(blockDim.x) + static_cast<int>(threadIdx.x);
    if (e >= count) return;
    const EnvCommand cmd = commands[e];
    if (!core::CommandActive(cmd)) return;
    core::CarArenaRefs r = CarRefsOf(a, c, e);
    core::ApplyEnvCommandCars(r, a.controls + r.car_base, a.max_cars, cmd, rows.cars ? &rows : nullptr, c.rules,
                              a.params, c.m);
}

// The unsupported words of the arenas (0 for every arena stepped in full).
__global__ void __launch_bounds__(256) UnsupportedKernel(const core::ArenaInternal* st, int32_t n, uint32_t* out) {
    const int e = static_cast<int>(blockIdx.x) * static_cast<int>(blockDim.x) + static_cast<int>(threadIdx.x);
    if (e < n) out[e] = st[e].unsupported;
}

template <int G, bool kEvents>
__device__ inline void StepGroupArenaWith(const StepArgs& a, int e, int n, core::BallTickScratch& scratch) {
    const core::EventSink events = EnvEvents(a, e, (static_cast<int>(threadIdx.x) % G) == 0);
    for (int t = 0; t < n; ++t) {
        // The predictive ray's walk is short and stays serial (every lane walks it, with the same loads).
        core::StepBallArena<GroupWalk<G>, core::SerialWalk, kEvents>(a.balls[e], a.envs[e], a.internal[e], a.mesh,
                                                                      a.params, scratch, events);
    }
}

// Arena e's n ticks on the G lanes of one group (stages 4-7): every lane steps it with the same data, so their
// in-place writes agree; the group's first lane writes its events.  kLds: the tick scratch is `slot`, the group's
// LDS slice; otherwise lane-private.
template <int G, bool kLds, bool kEvents>
__device__ inline void StepGroupArena(const StepArgs& a, int e, int n, core::BallTickScratch* slot) {
    if constexpr (kLds) {
        StepGroupArenaWith<G, kEvents>(a, e, n, *slot);
    } else {
        core::BallTickScratch private_scratch;
        StepGroupArenaWith<G, kEvents>(a, e, n, private_scratch);
    }
}

// G lanes per arena (stages 4 and 5).  Work item i = thread / G.
// At most 256 registers per lane, so that two waves fit each SIMD (8 arenas per wave fill 8 waves per CU at 16,384
// arenas).
template <int G, bool kLdsScratch, bool kEvents>
__global__ void __launch_bounds__(kMaxStepBlock) __attribute__((amdgpu_waves_per_eu(2)))
StepBallArenasGroupKernel(StepArgs a, int32_t ticks) {
    extern __shared__ __align__(16) unsigned char lds[];
    const int g = static_cast<int>(blockIdx.x) * static_cast<int>(blockDim.x) + static_cast<int>(threadIdx.x);
    const int i = g / G;
    if (i >= a.count) return;
    const int e = a.arena_ids ? a.arena_ids[i] : i;
    if (e < 0 || e >= a.env_count) return;
    const int n = a.arena_ticks ? a.arena_ticks[i] : ticks;
    core::BallTickScratch* slot =
        kLdsScratch ? reinterpret_cast<core::BallTickScratch*>(lds) + static_cast<int>(threadIdx.x) / G : nullptr;
    StepGroupArena<G, kLdsScratch, kEvents>(a, e, n, slot);
}

// Stages 6 and 7, whole-batch launches: workgroups [0, heavy_blocks) step the heavy list with GH lanes per arena,
// the rest the light list with 8.  lists[0..cap) heavy arenas, lists[cap..cap+n) light ones; counts[0] heavy (may
// exceed cap: those went to the light list), counts[1] light.
template <int GH, bool kEvents>
__global__ void __launch_bounds__(kMaxStepBlock) __attribute__((amdgpu_waves_per_eu(2)))
StepBallArenasSplitKernel(StepArgs a, int32_t ticks, const int32_t* lists, const int32_t* counts, int32_t cap,
                          int32_t heavy_blocks) {
    extern __shared__ __align__(16) unsigned char lds[];
    core::BallTickScratch* const slices = reinterpret_cast<core::BallTickScratch*>(lds);
    const int b = static_cast<int>(blockIdx.x);
    if (b < heavy_blocks) {
        const int i = (b * static_cast<int>(blockDim.x) + static_cast<int>(threadIdx.x)) / GH;
        const int heavy = counts[0] < cap ? counts[0] : cap;
        if (i >= heavy) return;
        StepGroupArena<GH, true, kEvents>(a, lists[i], ticks, slices + static_cast<int>(threadIdx.x) / GH);
    } else {
        const int i = ((b - heavy_blocks) * static_cast<int>(blockDim.x) + static_cast<int>(threadIdx.x)) / 8;
        if (i >= counts[1]) return;
        StepGroupArena<8, true, kEvents>(a, lists[cap + i], ticks, slices + static_cast<int>(threadIdx.x) / 8);
    }
}

__global__ void __launch_bounds__(64) ClearSplitCountsKernel(int32_t* counts) {
    if (threadIdx.x < 2) counts[threadIdx.x] = 0;
}

// Sorts the arenas into the heavy and light lists (any order: arenas step independently).
__global__ void __launch_bounds__(256) PartitionArenasKernel(const core::ArenaInternal* internal, int32_t n,
                                                             int32_t rounds, int32_t cap, int32_t* lists,
                                                             int32_t* counts) {
    const int e = static_cast<int>(blockIdx.x) * static_cast<int>(blockDim.x) + static_cast<int>(threadIdx.x);
    if (e >= n) return;
    if (internal[e].walk_rounds >= rounds) {
        const int h = atomicAdd(&counts[0], 1);
        if (h < cap) {
            lists[h] = e;
            return;
        }
    }
    lists[cap + atomicAdd(&counts[1], 1)] = e;
}

// The split schedule's lists for rg_batch_step_split_no_alloc, in a fixed array of the library instead of
// batch-owned memory: words [0, 2) are the two list counts, [64, 64 + 2 kNoAllocMaxArenas) the heavy then light
// arena ids.  Each GPU holds its own copy (2 MiB) until the process ends; calls on one GPU share it.
constexpr int32_t kNoAllocMaxArenas = 262144;
constexpr int32_t kNoAllocListOffset = 64;
static __device__ __align__(256) int32_t scratch[kNoAllocListOffset + 2 * kNoAllocMaxArenas];

// Device address of `scratch` on `device`, looked up once per device with that device current (the host-side name
// of a __device__ array is a stand-in, not a device address).
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
int32_t* g_scratch_ptr[kMaxDevices] = {};

int32_t* NoAllocScratch(int device) {
    if (device < 0 || device >= kMaxDevices) throw std::runtime_error("device ordinal outside 0..63");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (g_scratch_ptr[device] == nullptr) {
        void* p = nullptr;
        Check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        if (p == nullptr) throw std::runtime_error("hipGetSymbolAddress(scratch) returned null");
        g_scratch_ptr[device] = static_cast<int32_t*>(p);
    }
    return g_scratch_ptr[device];
}

__device__ inline uint64_t Fnv(uint64_t h, const void* data, int n) {
    const unsigned char* p = static_cast<const unsigned char*>(data);
    for (int i = 0; i < n; ++i) {
        h ^= p[i];
        h *= rgw::kFnvPrime;
    }
    return h;
}

// Ball path (BallPathEnv, env_step.hpp:7217-7259): every arena's ball stepped in place for samples x ticks ticks;
// after each `ticks` ticks its pos and vel go to out[arena][sample][6].  Once |pos.y| reaches goal_y the arena is
// not stepped any more (the remaining samples repeat that state); result[arena] = {goal side (1 +y, 2 -y, 0 none),
// the tick it crossed on (0 if it started there)}.
template <bool kEvents>
__global__ void __launch_bounds__(64) BallPathKernel(StepArgs a, int32_t samples, int32_t ticks, float goal_y,
                                                      float* out, int32_t* result) {
    const int e = static_cast<int>(blockIdx.x) * static_cast<int>(blockDim.x) + static_cast<int>(threadIdx.x);
    if (e >= a.env_count) return;
    core::BallTickScratch scratch;
    BallState& ball = a.balls[e];
    bool scored = fabsf(ball.pos.y) >= goal_y;
    int goal_tick = 0;
    float* o = out + static_cast<size_t>(e) * samples * 6;
    // The ticks are ordinary steps, events included.
    const core::EventSink events = EnvEvents(a, e);
    for (int s = 0; s < samples; ++s) {
        for (int k = 0; k < ticks && !scored; ++k) {
            Tick<kEvents>(ball, a.envs[e], a.internal[e], a, scratch, events);
            if (fabsf(ball.pos.y) >= goal_y) {
                scored = true;
                goal_tick = s * ticks + k + 1;
            }
        }
        o[6 * s + 0] = ball.pos.x;
        o[6 * s + 1] = ball.pos.y;
        o[6 * s + 2] = ball.pos.z;
        o[6 * s + 3] = ball.vel.x;
        o[6 * s + 4] = ball.vel.y;
        o[6 * s + 5] = ball.vel.z;
    }
    result[2 * e] = scored ? (ball.pos.y > 0.0f ? 1 : 2) : 0;
    result[2 * e + 1] = goal_tick;
}

__global__ void EnvDigestKernel(const BallState* balls, const CarState* cars, const int32_t* env_car_count,
                                int32_t env_count, int32_t max_cars, uint64_t* out) {
    const int e = static_cast<int>(blockIdx.x) * static_cast<int>(blockDim.x) + static_cast<int>(threadIdx.x);
    if (e >= env_count) return;
    const BallState& b = balls[e];
    uint64_t h = rgw::kFnvOffset;
    const float brot[4] = {b.rot.w, b.rot.x, b.rot.y, b.rot.z};
    h = Fnv(h, &b.pos, 12);
    h = Fnv(h, &b.vel, 12);
    h = Fnv(h, &b.ang_vel, 12);
    h = Fnv(h, brot, 16);
    const int n = env_car_count[e];
    for (int i = 0; i < n; ++i) {
        const CarState& c = cars[static_cast<size_t>(e) * max_cars + i];
        const float rot[4] = {c.rot.w, c.rot.x, c.rot.y, c.rot.z};
        h = Fnv(h, &c.pos, 12);
        h = Fnv(h, &c.vel, 12);
        h = Fnv(h, &c.ang_vel, 12);
        h = Fnv(h, rot, 16);
    }
    out[e] = h;
}

}  // namespace

class BatchSim {
public:
    BatchSim(const rg_sim_config& cfg, const std::vector<MeshSource>& meshes, const rgw::ArenaRulesSpec& rules)
        : cfg_(cfg) {
        if (cfg.env_count <= 0) throw InvalidArgument("env_count must be positive");
        if (cfg.max_cars_per_env < 0 || cfg.max_cars_per_env > 8) {
            throw InvalidArgument("max_cars_per_env must be 0..8");
        }
        if (!(cfg.tick_dt > 0.0f)) throw InvalidArgument("tick_dt must be positive");
        mesh_host_ = BuildArenaMesh(meshes);
        params_ = DefaultStepParams(rules);
        params_.tick_dt = cfg.tick_dt;
        params_.ball_damping_scale =
            params_.ball_drag <= 0.0f ? 1.0f
                                      : core::UcrtPowf(core::Max(0.0f, 1.0f - params_.ball_drag), params_.tick_dt);
        params_.enable_scoring = cfg.enable_scoring;
        car_args_ = CarStepArgs{};
        car_args_.m.tick_dt = cfg.tick_dt;
        car_args_.m.enable_car_car = cfg.enable_car_car_collisions ? 1 : 0;
        car_args_.m.enable_demos = cfg.enable_demos ? 1 : 0;
        car_args_.m.frame_frac = params_.frame_frac;
        car_args_.rules = SpawnRules(rules);
        for (int k = 0; k < rules.pad_count && k < rgw::kMaxPads; ++k) {
            car_args_.pad_pos[k] = Vec3(rules.pad_pos[k].x, rules.pad_pos[k].y, rules.pad_pos[k].z);
        }
        Allocate();
    }

    // A batch on another batch's arena (mesh, spawn rules, pads and physics) with its own arena and car counts and
    // settled-kickoff switch: the run that records settled kickoff rows.
    BatchSim(const BatchSim& arena, int env_count, int max_cars, bool kickoff_settled) : cfg_(arena.cfg_) {
        cfg_.env_count = env_count;
        cfg_.max_cars_per_env = max_cars;
        mesh_host_ = arena.mesh_host_;
        params_ = arena.params_;
        car_args_ = arena.car_args_;
        mutator_ = arena.mutator_;
        kickoff_settled_ = kickoff_settled;
        Allocate();
    }

    // Device buffers, default records and the constructor's kickoff reset with seed 0.
    void Allocate() {
        const rg_sim_config& cfg = cfg_;
        int device = 0;
        Check(hipGetDevice(&device), "hipGetDevice");
        no_alloc_scratch_ = NoAllocScratch(device);
        tris_.Resize(mesh_host_.tris.size());
        nodes_.Resize(mesh_host_.nodes.size());
        infos_.Resize(mesh_host_.meshes.size());
        tris_.Upload(mesh_host_.tris.data(), mesh_host_.tris.size(), nullptr);
        nodes_.Upload(mesh_host_.nodes.data(), mesh_host_.nodes.size(), nullptr);
        infos_.Upload(mesh_host_.meshes.data(), mesh_host_.meshes.size(), nullptr);
        const size_t n = static_cast<size_t>(cfg.env_count);
        const size_t cap = n * static_cast<size_t>(cfg.max_cars_per_env);
        balls_.Resize(n);
        envs_.Resize(n);
        internal_.Resize(n);
        car_count_.Resize(n);
        cars_.Resize(cap);
        controls_.Resize(cap);
        wheels_.Resize(cap);
        teams_.Resize(cap);
        configs_.Resize(cap);
        pads_.Resize(n * static_cast<size_t>(params_.pad_count));
        settled_.Resize(1);
        digests_.Resize(n);
        // Records start as the reference's constructor leaves them (src/rocketgpu.cpp:1111-1130): default-valued
        // records with zeroed padding, Blue teams, default car configs; then a kickoff reset with seed 0 (:1281).
        std::vector<BallState> balls(n);
        std::vector<EnvState> envs(n);
        ZeroedDefaults(balls);
        ZeroedDefaults(envs);
        const std::vector<int32_t> counts(n, 0);
        std::vector<CarState> cars(cap);
        std::vector<CarControls> controls(cap);
        std::vector<CarConfig> configs(cap);
        std::vector<WheelRuntimeState> wheels(cap);
        ZeroedDefaults(wheels);
        ZeroedDefaults(cars);
        ZeroedDefaults(controls);
        ZeroedDefaults(configs);
        const std::vector<uint8_t> teams(cap, 0);
        balls_.Upload(balls.data(), n, nullptr);
        envs_.Upload(envs.data(), n, nullptr);
        car_count_.Upload(counts.data(), n, nullptr);
        cars_.Upload(cars.data(), cap, nullptr);
        controls_.Upload(controls.data(), cap, nullptr);
        configs_.Upload(configs.data(), cap, nullptr);
        wheels_.Upload(wheels.data(), cap, nullptr);
        teams_.Upload(teams.data(), cap, nullptr);
        host_counts_.assign(n, 0);
        unsupported_.Resize(n);
        Check(hipMemset(internal_.data(), 0, n * sizeof(core::ArenaInternal)), "hipMemset");
        if (pads_.size()) Check(hipMemset(pads_.data(), 0, pads_.size() * sizeof(BoostPadState)), "hipMemset");
        Check(hipDeviceSynchronize(), "hipDeviceSynchronize");
        ResetToKickoff(0, nullptr);
        Check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    }

    int64_t env_count() const { return cfg_.env_count; }
    int64_t car_capacity() const { return static_cast<int64_t>(cfg_.env_count) * cfg_.max_cars_per_env; }
    int64_t boost_pad_capacity() const { return static_cast<int64_t>(cfg_.env_count) * params_.pad_count; }

    // set_teams / set_car_configs (src/rocketgpu.cpp:2934-2950): stored for the slots; arenas here carry no cars.
    void SetTeams(const uint8_t* teams, size_t n, hipStream_t s) {
        if (n > static_cast<size_t>(car_capacity())) throw std::out_of_range("set_teams count exceeds car_capacity");
        if (n && !teams) throw InvalidArgument("host_teams must not be null");
        teams_.Upload(teams, n, s);
        kick_dirty_ = true;
        Check(hipStreamSynchronize(s), "hipStreamSynchronize");
    }
    void SetCarConfigs(const CarConfig* configs, size_t n, hipStream_t s) {
        if (n > static_cast<size_t>(car_capacity())) {
            throw std::out_of_range("set_car_configs count exceeds car_capacity");
        }
        if (n && !configs) throw InvalidArgument("host_configs must not be null");
        configs_.Upload(configs, n, s);
        kick_dirty_ = true;
        Check(hipStreamSynchronize(s), "hipStreamSynchronize");
    }

    // MutatorConfig::kickoff_settled (rocketgpu.h:827-832): resets without a countdown put the ball where the
    // game's countdown leaves it.  On by default, as in the reference.
    void SetKickoffSettled(bool on) {
        kickoff_settled_ = on;
        mutator_.kickoff_settled = on ? 1 : 0;
    }

    // set_mutator_config (rocketgpu.h:1003): checked, applied to the step's parameters; the settled kickoff ball
    // and rows are recorded again under the new physics.
    void SetMutatorConfig(const rg_mutator_config& m) {
        CheckMutatorConfig(m);
        ApplyMutatorConfig(m, params_, car_args_.m);
        mutator_ = m;
        kickoff_settled_ = m.kickoff_settled != 0;
        settled_ready_ = false;
        kick_groups_.clear();
        kick_rows_car_.clear();
        kick_rows_wheels_.clear();
        kick_dirty_ = true;
        DestroyGraph();
    }
    const rg_mutator_config& MutatorConfig() const { return mutator_; }

    // set_physics_event_stream (src/rocketgpu.cpp:2896-2905): all zero disables the stream; otherwise it must cover
    // the whole batch.  Every step form writes to it from then on (a captured graph is captured again).  Arenas
    // here are Soccar, the mode the reference requires for detailed events.
    void SetPhysicsEventStream(const EventStream& st) {
        const bool disabled = !st.records && !st.windows && !st.capacity_per_env && !st.env_count && !st.cars_per_env;
        if (!disabled && (!st.records || !st.windows || !st.capacity_per_env ||
                          st.env_count != static_cast<uint32_t>(cfg_.env_count) ||
                          st.cars_per_env != static_cast<uint32_t>(cfg_.max_cars_per_env))) {
            throw InvalidArgument("physics event stream must match the complete batch");
        }
        events_ = disabled ? EventStream{} : st;
        DestroyGraph();
    }

    // reset_to_kickoff (src/rocketgpu.cpp:2085-2133): every env to a new match's kickoff.
    void ResetToKickoff(uint64_t seed, hipStream_t s) {
        if (has_cars_) {
            seed_ = seed;
            const bool rows = EnsureKickoffRows();
            const StepArgs a = BaseArgs();
            const int grid = (cfg_.env_count + 63) / 64;
            hipLaunchKernelGGL(ResetToKickoffCarsKernel, dim3(grid), dim3(64), 0, s, a, CarArgs(controls_.data(), 0),
                               KickoffRowsArg(rows), seed);
            Check(hipGetLastError(), "ResetToKickoffCarsKernel launch");
            return;
        }
        seed_ = seed;
        EnsureSettledBall();
        StepArgs a = BaseArgs();
        const int grid = (cfg_.env_count + 63) / 64;
        hipLaunchKernelGGL(ResetToKickoffKernel, dim3(grid), dim3(64), 0, s, a, seed);
        Check(hipGetLastError(), "ResetToKickoffKernel launch");
    }

    // apply_env_commands (src/rocketgpu.cpp:2136-2218): command i to env i for i < count.  Same checks and messages.
    void ApplyEnvCommands(const EnvCommand* host_commands, size_t count, hipStream_t s) {
        if (host_commands == nullptr) throw InvalidArgument("host_commands must not be null");
        if (count > static_cast<size_t>(cfg_.env_count)) {
            throw std::out_of_range("apply_env_commands_host count exceeds env_count");
        }
        size_t active = 0;
        bool settled_reset = false;
        for (size_t i = 0; i < count; ++i) {
            active += core::CommandActive(host_commands[i]) ? 1 : 0;
            settled_reset = settled_reset || (host_commands[i].reset_to_kickoff && host_commands[i].kickoff_countdown_ticks == 0);
        }
        if (active == 0) return;
        if (commands_.size() < count) commands_.Resize(static_cast<size_t>(cfg_.env_count));
        if (has_cars_) {
            const bool rows = settled_reset && EnsureKickoffRows();
            commands_.Upload(host_commands, count, s);
            LaunchApplyCommandsCars(commands_.data(), static_cast<int32_t>(count), rows, s);
            Check(hipStreamSynchronize(s), "hipStreamSynchronize");
            return;
        }
        if (settled_reset) EnsureSettledBall();
        commands_.Upload(host_commands, count, s);
        LaunchApplyCommands(commands_.data(), static_cast<int32_t>(count), s);
        Check(hipStreamSynchronize(s), "hipStreamSynchronize");
    }

    // apply_env_commands_device (src/rocketgpu.cpp:2221-2248): the commands are read on the device.
    void ApplyEnvCommandsDevice(const EnvCommand* device_commands, size_t count, hipStream_t s) {
        if (device_commands == nullptr) throw InvalidArgument("device_commands must not be null");
        if (count > static_cast<size_t>(cfg_.env_count)) {
            throw std::out_of_range("apply_env_commands_device count exceeds env_count");
        }
        if (has_cars_) {
            const bool rows = EnsureKickoffRows();
            if (count > 0) LaunchApplyCommandsCars(device_commands, static_cast<int32_t>(count), rows, s);
            return;
        }
        EnsureSettledBall();
        if (count == 0) return;
        LaunchApplyCommands(device_commands, static_cast<int32_t>(count), s);
    }

    // step_with_device_commands (src/rocketgpu.cpp:2554-2627): per tick, the env's command, then the tick; the
    // last tick's controls become the batch's controls.  Same checks and messages as the reference.
    void StepWithDeviceCommands(const CarControls* controls, int32_t control_stride, const EnvCommand* commands,
                                int32_t command_stride, int32_t ticks, hipStream_t s) {
        if (ticks <= 0) return;
        if (controls == nullptr) throw InvalidArgument("device_controls_by_tick must not be null");
        if (commands == nullptr) throw InvalidArgument("device_env_commands_by_tick must not be null");
        RejectCarCommands();
        EnsureSettledBall();
        if (control_stride != 0 && control_stride < car_capacity()) {
            throw InvalidArgument("control_stride must be zero or at least the total car capacity");
        }
        if (command_stride < cfg_.env_count) throw InvalidArgument("env_command_stride must be at least env_count");
        StepArgs a = BaseArgs();
        a.commands = commands;
        a.command_stride = command_stride;
        a.count = cfg_.env_count;
        a.arenas_per_wave = per_wave_;
        const int64_t threads = (static_cast<int64_t>(cfg_.env_count) + per_wave_ - 1) / per_wave_ * 64;
        const int grid = static_cast<int>((threads + block_ - 1) / block_);
        if (a.event_windows) {
            hipLaunchKernelGGL(StepBallArenasCommandsKernel<true>, dim3(grid), dim3(block_), 0, s, a, ticks);
        } else {
            hipLaunchKernelGGL(StepBallArenasCommandsKernel<false>, dim3(grid), dim3(block_), 0, s, a, ticks);
        }
        Check(hipGetLastError(), "StepBallArenasCommandsKernel launch");
        if (car_capacity() > 0) {
            Check(hipMemcpyAsync(controls_.data(), controls + static_cast<size_t>(ticks - 1) * control_stride,
                                 static_cast<size_t>(car_capacity()) * sizeof(CarControls), hipMemcpyDeviceToDevice, s),
                  "hipMemcpyAsync controls");
        }
    }

    void SetEnvCarCount(const int32_t* counts, size_t n, hipStream_t s) {
        if (n != static_cast<size_t>(cfg_.env_count)) throw InvalidArgument("set_env_car_count needs env_count values");
        for (size_t i = 0; i < n; ++i) {
            if (counts[i] < 0 || counts[i] > cfg_.max_cars_per_env) {
                throw std::out_of_range("env car count outside 0..max_cars_per_env");
            }
        }
        car_count_.Upload(counts, n, s);
        host_counts_.assign(counts, counts + n);
        has_cars_ = std::any_of(host_counts_.begin(), host_counts_.end(), [](int32_t c) { return c > 0; });
        RebuildCarOrder(s);
        DestroyGraph();
        Check(hipStreamSynchronize(s), "hipStreamSynchronize");
    }

    // preserve: StateUploadMode::PreserveInternal (rocketgpu.h:1027-1030), the records are taken as they are;
    // otherwise PublicState: car and ball records are canonicalized as the reference's upload does.
    void Upload(int id, const void* src, size_t n, hipStream_t s, bool preserve = false) {
        switch (id) {
        case RG_ARRAY_BALL_STATES: {
            Expect(n, balls_.size(), "ball states");
            std::vector<BallState> tmp(n);
            std::memcpy(static_cast<void*>(tmp.data()), src, n * sizeof(BallState));
            if (!preserve) {
                for (BallState& b : tmp) CanonicalizeBall(b);
            }
            balls_.Upload(tmp.data(), n, s);
            Check(hipStreamSynchronize(s), "hipStreamSynchronize");
            break;
        }
        case RG_ARRAY_ENV_STATES:
            Expect(n, envs_.size(), "env states");
            envs_.Upload(static_cast<const EnvState*>(src), n, s);
            Check(hipStreamSynchronize(s), "hipStreamSynchronize");
            break;
        case RG_ARRAY_CONTROLS:
            Expect(n, controls_.size(), "controls");
            controls_.Upload(static_cast<const CarControls*>(src), n, s);
            Check(hipStreamSynchronize(s), "hipStreamSynchronize");
            break;
        case RG_ARRAY_CAR_STATES: {
            Expect(n, cars_.size(), "car states");
            std::vector<CarState> tmp(n);
            std::memcpy(static_cast<void*>(tmp.data()), src, n * sizeof(CarState));
            if (!preserve) {
                for (CarState& c : tmp) CanonicalizeCar(c);
            }
            cars_.Upload(tmp.data(), n, s);
            Check(hipStreamSynchronize(s), "hipStreamSynchronize");
            break;
        }
        case RG_ARRAY_BOOST_PAD_STATES:
            Expect(n, pads_.size(), "boost pad states");
            pads_.Upload(static_cast<const BoostPadState*>(src), n, s);
            Check(hipStreamSynchronize(s), "hipStreamSynchronize");
            break;
        case RG_ARRAY_WHEEL_STATES:
            Expect(n, wheels_.size(), "wheel runtime states");
            wheels_.Upload(static_cast<const WheelRuntimeState*>(src), n, s);
            Check(hipStreamSynchronize(s), "hipStreamSynchronize");
            break;
        default:
            throw InvalidArgument("unknown or read-only state array id");
        }
    }

    // `sync` false: the copy is only queued on `s` (download_*_async, rocketgpu.h:1050-1053); the caller waits on s.
    void Download(int id, void* dst, size_t n, hipStream_t s, bool sync = true) const {
        switch (id) {
        case RG_ARRAY_BALL_STATES: balls_.Download(static_cast<BallState*>(dst), n, s); break;
        case RG_ARRAY_ENV_STATES: envs_.Download(static_cast<EnvState*>(dst), n, s); break;
        case RG_ARRAY_CAR_STATES: cars_.Download(static_cast<CarState*>(dst), n, s); break;
        case RG_ARRAY_CONTROLS: controls_.Download(static_cast<CarControls*>(dst), n, s); break;
        case RG_ARRAY_BOOST_PAD_STATES: pads_.Download(static_cast<BoostPadState*>(dst), n, s); break;
        case RG_ARRAY_ENV_CAR_COUNT: car_count_.Download(static_cast<int32_t*>(dst), n, s); break;
        case RG_ARRAY_WHEEL_STATES: wheels_.Download(static_cast<WheelRuntimeState*>(dst), n, s); break;
        case RG_ARRAY_UNSUPPORTED: {
            const int grid = (cfg_.env_count + 255) / 256;
            hipLaunchKernelGGL(UnsupportedKernel, dim3(grid), dim3(256), 0, s, internal_.data(), cfg_.env_count,
                               unsupported_.data());
            Check(hipGetLastError(), "UnsupportedKernel launch");
            unsupported_.Download(static_cast<uint32_t*>(dst), n, s);
            break;
        }
        default: throw InvalidArgument("unknown state array id");
        }
        if (sync) Check(hipStreamSynchronize(s), "hipStreamSynchronize");
    }

    // copy_state_from (rocketgpu.h:1045): every state array of `src` (a batch of the same shape on the same arena
    // and physics) copied on the device, with its car counts, seed and settled-kickoff state.
    void CopyStateFrom(const BatchSim& src, hipStream_t s) {
        if (&src == this) return;
        if (src.cfg_.env_count != cfg_.env_count || src.cfg_.max_cars_per_env != cfg_.max_cars_per_env ||
            src.params_.pad_count != params_.pad_count || src.GeometryDigest() != GeometryDigest()) {
            throw InvalidArgument("copy_state_from needs a batch of the same shape, arena and physics");
        }
        const auto copy = [&](auto& dst, const auto& from) {
            using T = typename std::remove_pointer<decltype(dst.data())>::type;
            if (dst.size()) {
                Check(hipMemcpyAsync(dst.data(), from.data(), dst.size() * sizeof(T), hipMemcpyDeviceToDevice, s),
                      "hipMemcpyAsync D2D");
            }
        };
        copy(cars_, src.cars_);
        copy(controls_, src.controls_);
        copy(configs_, src.configs_);
        copy(teams_, src.teams_);
        copy(wheels_, src.wheels_);
        copy(balls_, src.balls_);
        copy(envs_, src.envs_);
        copy(pads_, src.pads_);
        copy(car_count_, src.car_count_);
        copy(internal_, src.internal_);
        if (src.settled_ready_) copy(settled_, src.settled_);
        Check(hipStreamSynchronize(s), "hipStreamSynchronize");
        host_counts_ = src.host_counts_;
        has_cars_ = src.has_cars_;
        RebuildCarOrder(s);
        seed_ = src.seed_;
        kickoff_settled_ = src.kickoff_settled_;
        settled_ready_ = src.settled_ready_;
        mutator_ = src.mutator_;
        kick_dirty_ = true;
        DestroyGraph();
    }

    // ---- Full-state snapshot (checkpoint / resume) -------------------------------------------------------------
    // The reference's full-state surface is Arena::GetFullState / SetFullState and Serialize
    // (src/rocketsim_compat_dev.cpp:369-447, include/rocketgpu/rocketsim_compat.h:293-325): every per-slot and
    // per-env array, retained manifolds included, so that a restored arena continues bit for bit.  Here the
    // snapshot is one flat blob: a header naming the batch it came from, then the public records in the
    // reference's layouts (car states, controls, car configs, teams, ball states, env states, boost pads, car
    // counts), then this build's own per-arena internal state (broadphase and retained ball manifolds) and the
    // settled kickoff ball.  A blob loads only into a batch of the same shape and arena geometry; its digest and
    // every index it holds are checked before anything is written.
    struct SnapshotHeader {
        char magic[8];
        uint32_t version;
        uint32_t env_count;
        uint32_t max_cars;
        uint32_t pad_count;
        uint32_t mesh_count;
        uint32_t tri_count;
        uint32_t node_count;
        uint32_t internal_bytes;
        uint64_t geometry_digest;
        uint64_t body_bytes;
        uint64_t body_digest;
        uint8_t kickoff_settled;
        uint8_t settled_ready;
        uint8_t reserved[6];
    };
    static constexpr char kSnapshotMagic[8] = {'R', 'G', 'S', 'N', 'A', 'P', 'S', 'T'};
    static constexpr uint32_t kSnapshotVersion = 2;

    size_t SnapshotBodyBytes() const {
        const size_t cars = static_cast<size_t>(car_capacity());
        const size_t envs = static_cast<size_t>(cfg_.env_count);
        return cars * (sizeof(CarState) + sizeof(CarControls) + sizeof(CarConfig) + 1) +
               envs * (sizeof(BallState) + sizeof(EnvState) + sizeof(int32_t) + sizeof(core::ArenaInternal)) +
               pads_.size() * sizeof(BoostPadState) + cars * sizeof(WheelRuntimeState) + sizeof(BallState) +
               sizeof(uint64_t);
    }
    size_t SnapshotBytes() const { return sizeof(SnapshotHeader) + SnapshotBodyBytes(); }

    // Digest of the arena geometry and step constants, over their fields (padding bytes excluded).
    uint64_t GeometryDigest() const {
        uint64_t h = rgw::kFnvOffset;
        const auto mix = [&h](const auto& v) { h = rgw::Fnv1a(h, &v, sizeof(v)); };
        for (const core::MeshTriangle& t : mesh_host_.tris) {
            mix(t.a), mix(t.b), mix(t.c), mix(t.normal), mix(t.min), mix(t.max), mix(t.flags);
            mix(t.edge_angle_v0v1), mix(t.edge_angle_v1v2), mix(t.edge_angle_v2v0), mix(t.qmin), mix(t.qmax);
        }
        for (const core::MeshBvhNode& n : mesh_host_.nodes) mix(n.qmin), mix(n.qmax), mix(n.escape_or_tri);
        for (const core::MeshInfo& m : mesh_host_.meshes) {
            mix(m.aabb_min), mix(m.aabb_max), mix(m.q.bvh_min), mix(m.q.bvh_max), mix(m.q.scale);
            mix(m.mesh_index), mix(m.tri_start), mix(m.tri_count), mix(m.node_start), mix(m.node_count);
            mix(m.ray_order);
        }
        const core::StepParams& p = params_;
        mix(p.tick_dt), mix(p.gravity_z), mix(p.ball_radius), mix(p.ball_mass), mix(p.ball_drag);
        mix(p.ball_damping_scale), mix(p.ball_friction), mix(p.ball_restitution), mix(p.ball_max_speed);
        mix(p.ball_max_ang_speed), mix(p.frame_frac), mix(p.enable_scoring), mix(p.goal_point), mix(p.goal_normal);
        mix(p.goal_min), mix(p.goal_max), mix(p.ball_spawn), mix(p.pad_count), mix(p.pad_big_mask);
        return h;
    }

    void SaveState(void* dst, size_t bytes, hipStream_t s) {
        if (bytes < SnapshotBytes()) {
            throw InvalidArgument("snapshot buffer holds " + std::to_string(bytes) + " B; the batch needs " +
                                  std::to_string(SnapshotBytes()));
        }
        if (!dst) throw InvalidArgument("snapshot buffer must not be null");
        unsigned char* body = static_cast<unsigned char*>(dst) + sizeof(SnapshotHeader);
        unsigned char* at = body;
        const auto take = [&](const auto& arr) {
            using T = typename std::remove_pointer<decltype(arr.data())>::type;
            arr.Download(reinterpret_cast<T*>(at), arr.size(), s);
            at += arr.size() * sizeof(T);
        };
        take(cars_);
        take(controls_);
        take(configs_);
        take(teams_);
        take(balls_);
        take(envs_);
        take(pads_);
        take(car_count_);
        take(internal_);
        take(wheels_);
        Check(hipStreamSynchronize(s), "hipStreamSynchronize");
        BallState settled{};
        if (settled_ready_) {
            settled_.Download(&settled, 1, s);
            Check(hipStreamSynchronize(s), "hipStreamSynchronize");
        }
        std::memcpy(at, &settled, sizeof(settled));
        at += sizeof(settled);
        std::memcpy(at, &seed_, sizeof(seed_));
        at += sizeof(seed_);
        SnapshotHeader hdr{};
        std::memcpy(hdr.magic, kSnapshotMagic, sizeof(hdr.magic));
        FillHeader(hdr);
        hdr.body_bytes = static_cast<uint64_t>(at - body);
        hdr.body_digest = rgw::Fnv1a(rgw::kFnvOffset, body, static_cast<size_t>(at - body));
        hdr.kickoff_settled = kickoff_settled_ ? 1 : 0;
        hdr.settled_ready = settled_ready_ ? 1 : 0;
        std::memcpy(dst, &hdr, sizeof(hdr));
    }

    void LoadState(const void* src, size_t bytes, hipStream_t s) {
        if (!src) throw InvalidArgument("snapshot must not be null");
        if (bytes != SnapshotBytes()) {
            throw InvalidArgument("snapshot is " + std::to_string(bytes) + " B; this batch's snapshots are " +
                                  std::to_string(SnapshotBytes()));
        }
        SnapshotHeader hdr;
        std::memcpy(&hdr, src, sizeof(hdr));
        SnapshotHeader want{};
        FillHeader(want);
        if (std::memcmp(hdr.magic, kSnapshotMagic, sizeof(hdr.magic)) != 0 || hdr.version != kSnapshotVersion) {
            throw InvalidArgument("not a batch snapshot of this version");
        }
        if (hdr.env_count != want.env_count || hdr.max_cars != want.max_cars || hdr.pad_count != want.pad_count ||
            hdr.mesh_count != want.mesh_count || hdr.tri_count != want.tri_count || hdr.node_count != want.node_count ||
            hdr.internal_bytes != want.internal_bytes || hdr.geometry_digest != want.geometry_digest ||
            hdr.body_bytes != SnapshotBodyBytes()) {
            throw InvalidArgument("snapshot was taken from a batch of another shape or arena");
        }
        const unsigned char* body = static_cast<const unsigned char*>(src) + sizeof(SnapshotHeader);
        if (rgw::Fnv1a(rgw::kFnvOffset, body, hdr.body_bytes) != hdr.body_digest) {
            throw InvalidArgument("snapshot digest mismatch (the data is damaged)");
        }
        // Section offsets, then the internal records' indices checked before any upload.
        const size_t cars = static_cast<size_t>(car_capacity()), envs = static_cast<size_t>(cfg_.env_count);
        const size_t off_internal = cars * (sizeof(CarState) + sizeof(CarControls) + sizeof(CarConfig) + 1) +
                                    envs * (sizeof(BallState) + sizeof(EnvState) + sizeof(int32_t)) +
                                    pads_.size() * sizeof(BoostPadState);
        std::vector<core::ArenaInternal> internal(envs);
        std::memcpy(static_cast<void*>(internal.data()), body + off_internal, envs * sizeof(core::ArenaInternal));
        for (size_t e = 0; e < envs; ++e) {
            if (!InternalRecordValid(internal[e])) {
                throw InvalidArgument("snapshot holds an invalid internal record for env " + std::to_string(e));
            }
        }
        const size_t off_counts = off_internal - envs * sizeof(int32_t);
        std::vector<int32_t> counts(envs);
        for (size_t e = 0; e < envs; ++e) {
            std::memcpy(&counts[e], body + off_counts + e * sizeof(int32_t), sizeof(int32_t));
            if (counts[e] < 0 || counts[e] > cfg_.max_cars_per_env) {
                throw InvalidArgument("snapshot holds an invalid car count");
            }
        }
        const unsigned char* at = body;
        const auto put = [&](auto& arr) {
            using T = typename std::remove_pointer<decltype(arr.data())>::type;
            arr.Upload(reinterpret_cast<const T*>(at), arr.size(), s);
            at += arr.size() * sizeof(T);
        };
        put(cars_);
        put(controls_);
        put(configs_);
        put(teams_);
        put(balls_);
        put(envs_);
        put(pads_);
        put(car_count_);
        put(internal_);
        put(wheels_);
        kickoff_settled_ = hdr.kickoff_settled != 0;
        mutator_.kickoff_settled = hdr.kickoff_settled;
        if (hdr.settled_ready) {
            if (settled_.size() == 0) settled_.Resize(1);
            settled_.Upload(reinterpret_cast<const BallState*>(at), 1, s);
        }
        at += sizeof(BallState);
        std::memcpy(&seed_, at, sizeof(seed_));
        settled_ready_ = hdr.settled_ready != 0;
        host_counts_ = counts;
        has_cars_ = std::any_of(counts.begin(), counts.end(), [](int32_t c) { return c > 0; });
        RebuildCarOrder(s);
        kick_dirty_ = true;
        DestroyGraph();
        Check(hipStreamSynchronize(s), "hipStreamSynchronize");
    }

    void Step(int32_t ticks, hipStream_t s) {
        if (ticks < 0) throw InvalidArgument("ticks must be non-negative");
        if (ticks == 0) return;
        Launch(nullptr, nullptr, cfg_.env_count, ticks, s);
    }

    // Step(ticks) through the split schedule (heavy arenas by their last contact walk on `gh` lanes, light ones on 8,
    // workgroups of 64 lanes) with the lists in the library's fixed `scratch` array: no device memory is allocated
    // or freed.  Ball-only batches (the car kernel keeps its per-tick state lane-private and has no lists: Step
    // allocates nothing for car batches).  The counts are cleared by a kernel at the start of every call; calls
    // that overlap on one GPU race on the lists.
    void StepSplitNoAlloc(int32_t ticks, hipStream_t s) {
        if (ticks < 0) throw InvalidArgument("ticks must be non-negative");
        if (cfg_.env_count > kNoAllocMaxArenas) {
            throw InvalidArgument("step_split_no_alloc supports at most " + std::to_string(kNoAllocMaxArenas) +
                                  " arenas; this batch has " + std::to_string(cfg_.env_count));
        }
        if (has_cars_) throw InvalidArgument("step_split_no_alloc steps ball-only batches; use step for car arenas");
        if (ticks == 0) return;
        StepArgs a = BaseArgs();
        a.count = cfg_.env_count;
        a.arenas_per_wave = 8;
        const int gh = stage_ == kStageSplit16 ? 16 : 32;
        int32_t* const counts = no_alloc_scratch_;
        int32_t* const lists = no_alloc_scratch_ + kNoAllocListOffset;
        if (a.event_windows) {
            LaunchSplit<true>(a, ticks, lists, counts, gh, 64, s);
        } else {
            LaunchSplit<false>(a, ticks, lists, counts, gh, 64, s);
        }
    }

    // Steps a list of distinct arenas (device ids), each for its own tick count (device array, or `ticks` for all
    // when null).  The serving loop's launch.
    void StepArenas(const int32_t* device_ids, const int32_t* device_ticks, int32_t count, int32_t ticks,
                    hipStream_t s) {
        if (count < 0 || count > cfg_.env_count) throw InvalidArgument("arena count outside 0..env_count");
        if (count > 0 && device_ids == nullptr) throw InvalidArgument("arena id list is null");
        if (device_ticks == nullptr && ticks < 0) throw InvalidArgument("ticks must be non-negative");
        if (count == 0 || (device_ticks == nullptr && ticks == 0)) return;
        Launch(device_ids, device_ticks, count, ticks, s);
    }

    // `launches` step launches of `ticks_per_launch` ticks each, captured once into a hipGraph (on a private stream)
    // and replayed on `s`; the graph is re-captured when the shape or the launch block changes.  The captured kernels
    // hold the batch's own device pointers, which live as long as the batch.
    void StepGraph(int32_t ticks_per_launch, int32_t launches, hipStream_t s) {
        if (ticks_per_launch <= 0 || launches <= 0 || launches > 1024) {
            throw InvalidArgument("graph step needs positive ticks and 1..1024 launches");
        }
        if (!graph_exec_ || graph_ticks_ != ticks_per_launch || graph_launches_ != launches ||
            graph_shape_ != LaunchShape()) {
            DestroyGraph();
            if (!capture_stream_) {
                Check(hipStreamCreateWithFlags(&capture_stream_, hipStreamNonBlocking), "hipStreamCreate");
            }
            Check(hipStreamBeginCapture(capture_stream_, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
            for (int32_t l = 0; l < launches; ++l) Launch(nullptr, nullptr, cfg_.env_count, ticks_per_launch, capture_stream_);
            hipGraph_t graph = nullptr;
            Check(hipStreamEndCapture(capture_stream_, &graph), "hipStreamEndCapture");
            const hipError_t e = hipGraphInstantiate(&graph_exec_, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            Check(e, "hipGraphInstantiate");
            graph_ticks_ = ticks_per_launch;
            graph_launches_ = launches;
            graph_shape_ = LaunchShape();
        }
        Check(hipGraphLaunch(graph_exec_, s), "hipGraphLaunch");
    }

    ~BatchSim() {
        DestroyGraph();
        if (capture_stream_) (void)hipStreamDestroy(capture_stream_);
    }

    // Lanes per workgroup of the step kernel (64, 128 or 256) and state staging (0 in place, 1 private copy);
    // chosen by the launch-shape tuner.  Results are identical for every shape.
    void SetBlock(int block) {
        if (block != 64 && block != 128 && block != 256) throw InvalidArgument("block must be 64, 128 or 256");
        block_ = block;
    }
    void SetStage(int stage) {
        if (stage < kStageInPlace || stage > kStageSplit16) throw InvalidArgument("stage must be 0 to 7");
        stage_ = stage;
    }
    void SetArenasPerWave(int n) {
        if (n != 64 && n != 32 && n != 16 && n != 8 && n != 4) throw InvalidArgument("arenas per wave must be 64, 32, 16, 8 or 4");
        per_wave_ = n;
    }
    int LaunchShape() const { return (block_ * 4 + stage_) * 128 + per_wave_; }

    // Rejects a shape whose LDS slices do not fit one workgroup's LDS; allocates what the shape's launches need.
    void CheckLaunchShape() {
        // A batch with cars at stage 4 gives each arena 2 lanes (32 arenas per wave) or 4 (16, 8 or 4).
        const bool car_pairs = has_cars_ && stage_ == kStageGroup && per_wave_ == 32;
        if (IsGroupStage(stage_) && !car_pairs && per_wave_ != 16 && per_wave_ != 8 && per_wave_ != 4) {
            throw InvalidArgument("stages 4 and 5 need 16, 8 or 4 arenas per wave (4, 8 or 16 lanes per arena)");
        }
        if (IsSplitStage(stage_) && per_wave_ != 8) {
            throw InvalidArgument("stages 6 and 7 need 8 arenas per wave (8 lanes per light arena)");
        }
        const size_t lds = LdsBytesPerLane(stage_) * static_cast<size_t>(block_ / 64) * per_wave_;
        int device = 0, limit = 0;
        Check(hipGetDevice(&device), "hipGetDevice");
        Check(hipDeviceGetAttribute(&limit, hipDeviceAttributeMaxSharedMemoryPerBlock, device), "hipDeviceGetAttribute");
        if (lds > static_cast<size_t>(limit)) {
            throw InvalidArgument("launch shape needs " + std::to_string(lds) + " B of LDS per workgroup; the device allows " +
                                  std::to_string(limit));
        }
        if (lds > 65536) {
            AllowLargeLds<false>(limit);
            AllowLargeLds<true>(limit);
        }
        // The split stages' lists are allocated here, not at launch: a launch may be under graph capture.
        if (IsSplitStage(stage_)) {
            const size_t n = static_cast<size_t>(cfg_.env_count);
            if (split_lists_.size() < n + n) split_lists_.Resize(n + n);
            if (split_counts_.size() < 2) split_counts_.Resize(2);
        }
    }
    // Stages 6 and 7: the windows a contact walk must load for its arena to count as heavy, and the heavy list's
    // capacity as a divisor of env_count.
    void SetSplit(int rounds, int cap_divisor) {
        if (rounds < 1 || rounds > 65535) throw InvalidArgument("split rounds must be 1 to 65535");
        if (cap_divisor < 1 || cap_divisor > 65536) throw InvalidArgument("split capacity divisor must be 1 to 65536");
        split_rounds_ = rounds;
        split_cap_divisor_ = cap_divisor;
        DestroyGraph();
    }
    int block() const { return block_; }
    int stage() const { return stage_; }
    int arenas_per_wave() const { return per_wave_; }

    // Same argument checks and messages as the reference (src/rocketgpu.cpp:2412-2421).
    void StepWithDeviceControls(const void* controls, int32_t ticks, int32_t stride, hipStream_t s) {
        if (ticks <= 0) return;
        if (controls == nullptr) throw InvalidArgument("device_controls_by_tick must not be null");
        if (stride != 0 && stride < car_capacity()) {
            throw InvalidArgument("control_stride must be zero or at least the total car capacity");
        }
        if (has_cars_) {
            LaunchCars(nullptr, nullptr, cfg_.env_count, ticks, static_cast<const CarControls*>(controls), stride, s);
            return;
        }
        Step(ticks, s);
    }

    void DeviceView(rg_device_view* v) const {
        std::memset(v, 0, sizeof(*v));
        v->car_states = cars_.data();
        v->controls = controls_.data();
        v->ball_states = balls_.data();
        v->env_states = envs_.data();
        v->env_car_count = car_count_.data();
        v->boost_pad_states = pads_.data();
        v->teams = teams_.data();
        v->car_configs = configs_.data();
    }

    // ball_path (rocketgpu.h:1069-1073): device out [env][samples][6] and result [env][2].  The arenas must hold
    // no cars, as in the reference; this build's arenas never do.
    void BallPath(int32_t samples, int32_t ticks, float goal_y, float* out, int32_t* result, hipStream_t s) {
        if (samples <= 0 || ticks <= 0) return;
        if (!out || !result) throw InvalidArgument("ball_path output pointers must not be null");
        if (has_cars_) throw InvalidArgument("ball_path requires arenas without cars");
        const StepArgs a = BaseArgs();
        const int grid = (cfg_.env_count + 63) / 64;
        if (a.event_windows) {
            hipLaunchKernelGGL(BallPathKernel<true>, dim3(grid), dim3(64), 0, s, a, samples, ticks, goal_y, out, result);
        } else {
            hipLaunchKernelGGL(BallPathKernel<false>, dim3(grid), dim3(64), 0, s, a, samples, ticks, goal_y, out, result);
        }
        Check(hipGetLastError(), "BallPathKernel launch");
    }

    void EnvDigests(uint64_t* out, hipStream_t s) {
        const int block = 256;
        const int grid = (cfg_.env_count + block - 1) / block;
        hipLaunchKernelGGL(EnvDigestKernel, dim3(grid), dim3(block), 0, s, balls_.data(), cars_.data(),
                           car_count_.data(), cfg_.env_count, cfg_.max_cars_per_env, out);
        Check(hipGetLastError(), "EnvDigestKernel launch");
    }

private:
    void RejectCarCommands() const {
        if (has_cars_) throw InvalidArgument("env commands on arenas with cars are not built yet");
    }

    void FillHeader(SnapshotHeader& h) const {
        std::memcpy(h.magic, kSnapshotMagic, sizeof(h.magic));
        h.version = kSnapshotVersion;
        h.env_count = static_cast<uint32_t>(cfg_.env_count);
        h.max_cars = static_cast<uint32_t>(cfg_.max_cars_per_env);
        h.pad_count = static_cast<uint32_t>(params_.pad_count);
        h.mesh_count = static_cast<uint32_t>(mesh_host_.meshes.size());
        h.tri_count = static_cast<uint32_t>(mesh_host_.tris.size());
        h.node_count = static_cast<uint32_t>(mesh_host_.nodes.size());
        h.internal_bytes = static_cast<uint32_t>(sizeof(core::ArenaInternal));
        h.geometry_digest = GeometryDigest();
    }

    // Every count and index of an internal record within the arrays the step indexes with it.
    bool InternalRecordValid(const core::ArenaInternal& st) const {
        const int meshes = static_cast<int>(mesh_host_.meshes.size());
        const core::SweepState& h = st.bp.sweep;
        if (h.handle_count > core::kSweepMaxHandles || h.static_count > h.handle_count ||
            h.next_fresh > core::kSweepMaxHandles + 1 || h.free_count > core::kMaxBodies ||
            h.obj_count > core::kMaxBodies) {
            return false;
        }
        for (int i = 0; i < core::kMaxBodies; ++i) {
            if (h.free_list[i] > core::kSweepMaxHandles || h.body_handle[i] > core::kSweepMaxHandles ||
                h.obj[i] >= core::kMaxBodies) {
                return false;
            }
        }
        for (int a = 0; a < 3; ++a) {
            for (int i = 0; i < core::kSweepMaxEdges; ++i) {
                if (h.edge_handle[a][i] > core::kSweepMaxHandles) return false;
            }
            for (int k = 0; k <= core::kSweepMaxHandles; ++k) {
                if (h.min_edge[k][a] >= core::kSweepMaxEdges || h.max_edge[k][a] >= core::kSweepMaxEdges) return false;
            }
        }
        for (int k = 0; k <= core::kSweepMaxHandles; ++k) {
            if (h.static_id[k] >= meshes || h.static_id[k] < -core::kMaxBodies) return false;
        }
        if (st.bp.pair_count > core::kMaxManifoldEntries || st.bp.manifold_count > core::kMaxManifoldEntries ||
            st.bp.predictive_count > core::kMaxBodies) {
            return false;
        }
        const auto key_ok = [&](uint32_t key) {
            const int sid = core::KeyStatic(key);
            return sid == -1 || sid == core::kPredictiveStaticId || (sid >= 0 && sid < meshes);
        };
        for (int i = 0; i < st.bp.pair_count; ++i) {
            if (!key_ok(st.bp.pairs[i])) return false;
        }
        for (int i = 0; i < st.bp.manifold_count; ++i) {
            if (!key_ok(st.bp.manifolds[i])) return false;
        }
        for (int i = 0; i < st.bp.predictive_count; ++i) {
            if (!key_ok(st.bp.predictive[i])) return false;
        }
        if (meshes < 16 && (st.ball_mesh_retained >> meshes) != 0) return false;
        for (int m = 0; m < core::kMaxArenaMeshes; ++m) {
            if (st.ball_mesh[m].count > 4) return false;
        }
        for (int c = 0; c < core::kMaxArenaCarSlots; ++c) {
            for (int k = 0; k < core::kCarWorldManifoldsPerCar; ++k) {
                const core::CarWorldManifold& w = st.car_world[c][k];
                if (w.count > 4 || (w.count > 0 && w.mesh >= meshes)) return false;
            }
        }
        for (int k = 0; k < core::kCarPairSlots; ++k) {
            if (st.car_pair[k].count > 4) return false;
        }
        return true;
    }

    void DestroyGraph() {
        if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
        graph_exec_ = nullptr;
    }

    template <typename T>
    static void ZeroedDefaults(std::vector<T>& v) {
        std::memset(static_cast<void*>(v.data()), 0, v.size() * sizeof(T));
        for (T& x : v) new (&x) T;
    }

    // The batch's own arrays and geometry; work-list fields are set by the caller.
    StepArgs BaseArgs() const {
        StepArgs a{};
        a.balls = balls_.data();
        a.envs = envs_.data();
        a.internal = internal_.data();
        a.env_car_count = car_count_.data();
        a.mesh.tris = tris_.data();
        a.mesh.nodes = nodes_.data();
        a.mesh.meshes = infos_.data();
        a.mesh.mesh_count = static_cast<int32_t>(mesh_host_.meshes.size());
        a.mesh.tri_count = static_cast<int32_t>(mesh_host_.tris.size());
        a.params = params_;
        a.env_count = cfg_.env_count;
        a.pads = pads_.data();
        a.controls = controls_.data();
        a.max_cars = cfg_.max_cars_per_env;
        a.settled = kickoff_settled_ && settled_ready_ ? settled_.data() : nullptr;
        a.event_windows = events_.windows;
        a.event_records = events_.records;
        a.event_capacity = events_.capacity_per_env;
        a.event_goal_only = events_.goal_only ? 1 : 0;
        return a;
    }

    void LaunchApplyCommands(const EnvCommand* commands, int32_t count, hipStream_t s) {
        const StepArgs a = BaseArgs();
        const int grid = (count + 63) / 64;
        hipLaunchKernelGGL(ApplyEnvCommandsKernel, dim3(grid), dim3(64), 0, s, a, commands, count);
        Check(hipGetLastError(), "ApplyEnvCommandsKernel launch");
    }

    // The settled kickoff ball (the reference's record_kickoff_settled, src/rocketgpu.cpp:1538-1612): a fresh
    // arena gets a new-match reset with the game's countdown (the ball re-created 7.737 uu above its spawn, under
    // the 2D constraint), then kKickoffCountdownTicks ticks and one more at a time until the countdown has ended.
    // Its ball is what a settled kickoff reset copies.  Computed once per batch, on this GPU, by the step kernel.
    void EnsureSettledBall() {
        if (!kickoff_settled_ || settled_ready_) return;
        std::vector<BallState> ball(1);
        std::vector<EnvState> env(1);
        ZeroedDefaults(ball);
        ZeroedDefaults(env);
        std::vector<core::ArenaInternal> st(1);
        std::memset(static_cast<void*>(st.data()), 0, sizeof(core::ArenaInternal));
        core::ResetToKickoffNoCars(ball[0], env[0], st[0], nullptr, 0, 0, nullptr, params_);
        EnvCommand c{};
        c.reset_to_kickoff = 1;
        c.reset_scoreboard = 1;
        c.kickoff_seed_mode = 3;
        c.kickoff_countdown_ticks = core::kKickoffCountdownGameTimers;
        core::ApplyEnvCommandNoCars(ball[0], env[0], st[0], nullptr, 0, c, nullptr, params_);
        DeviceArray<EnvState> d_env;
        DeviceArray<core::ArenaInternal> d_st;
        d_env.Resize(1);
        d_st.Resize(1);
        settled_.Upload(ball.data(), 1, nullptr);
        d_env.Upload(env.data(), 1, nullptr);
        d_st.Upload(st.data(), 1, nullptr);
        StepArgs a = BaseArgs();
        a.balls = settled_.data();
        a.envs = d_env.data();
        a.internal = d_st.data();
        a.event_windows = nullptr;  // a private arena: no events
        a.event_records = nullptr;
        a.env_count = 1;
        a.count = 1;
        a.arenas_per_wave = 64;
        int32_t ticks = core::kKickoffCountdownTicks;
        for (int k = 0;; ++k) {
            hipLaunchKernelGGL((StepBallArenasKernel<kStageInPlace, false>), dim3(1), dim3(64), 0, nullptr, a, ticks);
            Check(hipGetLastError(), "StepBallArenasKernel launch");
            d_env.Download(env.data(), 1, nullptr);
            Check(hipStreamSynchronize(nullptr), "hipStreamSynchronize");
            if (env[0].countdown_phase == 0) break;
            if (k == 16) throw std::runtime_error("kickoff_settled: the countdown did not end");
            ticks = 1;
        }
        settled_ready_ = true;
    }

    // Car arenas in the batch: one lane per arena with the full car tick (ball arenas in the same launch take the
    // ball tick).  Physics events of car arenas are not written.
    CarStepArgs CarArgs(const CarControls* controls, int64_t stride) const {
        CarStepArgs c = car_args_;
        c.cars = cars_.data();
        c.wheels = wheels_.data();
        c.cfgs = configs_.data();
        c.teams = teams_.data();
        c.controls = controls;
        c.control_stride = stride;
        c.seed = seed_;
        return c;
    }

    // The settled kickoff rows (the reference's ensure_kickoff_settled / record_kickoff_settled,
    // src/rocketgpu.cpp:1439-1612), recomputed after teams or car configs change.  The batch's car slots are grouped
    // by (CarConfig bytes, team) in slot order; each new group is run in a batch of its own on this arena, one arena
    // per group with max(max_cars, 5) cars of that group: a new-match reset with the game's countdown (kickoff seed
    // mode 3: car k at kickoff spot k, k >= 5 at the respawn spots), kKickoffCountdownTicks ticks, then one at a
    // time until every countdown has ended.  Row (group, k) is car k's car and wheel state then; the ball row is the
    // first run's ball.  False when settled kickoffs are off.
    bool EnsureKickoffRows() {
        if (!kickoff_settled_) return false;
        if (!kick_dirty_) return true;
        const size_t cap = static_cast<size_t>(car_capacity());
        const int spawns = std::max(cfg_.max_cars_per_env, 5);
        std::vector<CarConfig> configs(cap);
        std::vector<uint8_t> teams(cap);
        configs_.Download(configs.data(), cap, nullptr);
        teams_.Download(teams.data(), cap, nullptr);
        Check(hipStreamSynchronize(nullptr), "hipStreamSynchronize");
        std::vector<int32_t> group(cap);
        std::vector<size_t> first_new;
        const size_t known = kick_groups_.size();
        for (size_t i = 0; i < cap; ++i) {
            size_t g = 0;
            while (g < kick_groups_.size() &&
                   !(kick_groups_[g].team == teams[i] &&
                     std::memcmp(&kick_groups_[g].config, &configs[i], sizeof(CarConfig)) == 0)) {
                ++g;
            }
            if (g == kick_groups_.size()) {
                kick_groups_.push_back({configs[i], teams[i]});
                first_new.push_back(i);
            }
            group[i] = static_cast<int32_t>(g);
        }
        if (!first_new.empty()) {
            try {
                RecordKickoffRows(first_new.size(), spawns, known);
            } catch (...) {
                kick_groups_.resize(known);
                kick_rows_car_.resize(known * static_cast<size_t>(spawns));
                kick_rows_wheels_.resize(known * static_cast<size_t>(spawns));
                throw;
            }
        }
        const size_t rows = kick_rows_car_.size();
        kick_cars_.Resize(rows);
        kick_wheels_.Resize(rows);
        kick_group_.Resize(cap);
        kick_ball_.Resize(1);
        kick_cars_.Upload(kick_rows_car_.data(), rows, nullptr);
        kick_wheels_.Upload(kick_rows_wheels_.data(), rows, nullptr);
        kick_group_.Upload(group.data(), cap, nullptr);
        kick_ball_.Upload(&kick_rows_ball_, 1, nullptr);
        Check(hipStreamSynchronize(nullptr), "hipStreamSynchronize");
        kick_spawns_ = spawns;
        kick_dirty_ = false;
        return true;
    }

    void RecordKickoffRows(size_t groups, int spawns, size_t known) {
        BatchSim run(*this, static_cast<int>(groups), spawns, false);
        const size_t m = groups * static_cast<size_t>(spawns);
        std::vector<int32_t> counts(groups, spawns);
        std::vector<CarConfig> configs(m);
        std::vector<uint8_t> teams(m);
        for (size_t i = 0; i < m; ++i) {
            configs[i] = kick_groups_[known + i / spawns].config;
            teams[i] = kick_groups_[known + i / spawns].team;
        }
        run.SetEnvCarCount(counts.data(), groups, nullptr);
        run.SetTeams(teams.data(), m, nullptr);
        run.SetCarConfigs(configs.data(), m, nullptr);
        EnvCommand c{};
        c.reset_to_kickoff = 1;
        c.reset_scoreboard = 1;
        c.kickoff_seed_mode = 3;
        c.kickoff_countdown_ticks = core::kKickoffCountdownGameTimers;
        const std::vector<EnvCommand> cmds(groups, c);
        run.ApplyEnvCommands(cmds.data(), groups, nullptr);
        run.Step(core::kKickoffCountdownTicks, nullptr);
        std::vector<EnvState> envs(groups);
        for (int k = 0;; ++k) {
            run.envs_.Download(envs.data(), groups, nullptr);
            Check(hipStreamSynchronize(nullptr), "hipStreamSynchronize");
            if (std::all_of(envs.begin(), envs.end(), [](const EnvState& e) { return e.countdown_phase == 0; })) break;
            if (k == 16) throw std::runtime_error("kickoff_settled: the countdown did not end");
            run.Step(1, nullptr);
        }
        std::vector<CarState> cars(m);
        std::vector<WheelRuntimeState> wheels(m);
        run.cars_.Download(cars.data(), m, nullptr);
        run.wheels_.Download(wheels.data(), m, nullptr);
        BallState ball;
        run.balls_.Download(&ball, 1, nullptr);
        Check(hipStreamSynchronize(nullptr), "hipStreamSynchronize");
        kick_rows_car_.insert(kick_rows_car_.end(), cars.begin(), cars.end());
        kick_rows_wheels_.insert(kick_rows_wheels_.end(), wheels.begin(), wheels.end());
        if (known == 0) kick_rows_ball_ = ball;
    }

    core::KickoffRows KickoffRowsArg(bool on) const {
        core::KickoffRows r{};
        if (on) {
            r.cars = kick_cars_.data();
            r.wheels = kick_wheels_.data();
            r.group = kick_group_.data();
            r.ball = kick_ball_.data();
            r.spawns = kick_spawns_;
        }
        return r;
    }

    void LaunchApplyCommandsCars(const EnvCommand* commands, int32_t count, bool rows, hipStream_t s) {
        const StepArgs a = BaseArgs();
        const int grid = (count + 63) / 64;
        hipLaunchKernelGGL(ApplyEnvCommandsCarsKernel, dim3(grid), dim3(64), 0, s, a, CarArgs(controls_.data(), 0),
                           KickoffRowsArg(rows), commands, count);
        Check(hipGetLastError(), "ApplyEnvCommandsCarsKernel launch");
    }

    // The plain launch's lane order: arenas by car count, most cars first (index order within a count), so that the
    // arenas sharing a wavefront run the same per-car loops.  Each arena's result does not depend on its lane.
    void RebuildCarOrder(hipStream_t s) {
        const int n = cfg_.env_count;
        std::vector<int32_t> order(static_cast<size_t>(n));
        for (int i = 0; i < n; ++i) order[static_cast<size_t>(i)] = i;
        std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) {
            return host_counts_[static_cast<size_t>(x)] > host_counts_[static_cast<size_t>(y)];
        });
        if (car_order_.size() != static_cast<size_t>(n)) car_order_.Resize(static_cast<size_t>(n));
        car_order_.Upload(order.data(), static_cast<size_t>(n), s);
        Check(hipStreamSynchronize(s), "hipStreamSynchronize");
    }

    void LaunchCars(const int32_t* ids, const int32_t* ticks_list, int32_t count, int32_t ticks,
                    const CarControls* controls, int64_t stride, hipStream_t s) {
        StepArgs a = BaseArgs();
        a.arena_ids = ids == nullptr && ticks_list == nullptr && count == cfg_.env_count &&
                              car_order_.size() == static_cast<size_t>(cfg_.env_count)
                          ? car_order_.data()
                          : ids;
        a.arena_ticks = ticks_list;
        a.count = count;
        a.arenas_per_wave = per_wave_;
        const CarStepArgs c = CarArgs(controls, stride);
        const int grid = (count + per_wave_ - 1) / per_wave_;
        if (stage_ == kStageLdsScratch) {
            const size_t lds = sizeof(core::CarArenaTick) * static_cast<size_t>(per_wave_);
            if (lds > car_lds_allowed_) {
                int device = 0, limit = 0;
                Check(hipGetDevice(&device), "hipGetDevice");
                Check(hipDeviceGetAttribute(&limit, hipDeviceAttributeMaxSharedMemoryPerBlock, device),
                      "hipDeviceGetAttribute");
                if (lds > static_cast<size_t>(limit)) {
                    throw InvalidArgument("car arenas with LDS staging need " + std::to_string(lds) +
                                          " B of LDS per workgroup; the device allows " + std::to_string(limit));
                }
                Check(hipFuncSetAttribute(reinterpret_cast<const void*>(&StepCarArenasKernel<true, 1>),
                                          hipFuncAttributeMaxDynamicSharedMemorySize, limit),
                      "hipFuncSetAttribute");
                car_lds_allowed_ = static_cast<size_t>(limit);
            }
            hipLaunchKernelGGL((StepCarArenasKernel<true, 1>), dim3(grid), dim3(64), lds, s, a, c, ticks);
        } else if (stage_ == kStageGroup && per_wave_ <= 16) {
            hipLaunchKernelGGL((StepCarArenasKernel<false, 4>), dim3(grid), dim3(64), 0, s, a, c, ticks);
        } else if (stage_ == kStageGroup && per_wave_ == 32) {
            hipLaunchKernelGGL((StepCarArenasKernel<false, 2>), dim3(grid), dim3(64), 0, s, a, c, ticks);
        } else {
            hipLaunchKernelGGL((StepCarArenasKernel<false, 1>), dim3(grid), dim3(64), 0, s, a, c, ticks);
        }
        Check(hipGetLastError(), "StepCarArenasKernel launch");
    }

    void Launch(const int32_t* ids, const int32_t* ticks_list, int32_t count, int32_t ticks, hipStream_t s) {
        if (has_cars_) {
            LaunchCars(ids, ticks_list, count, ticks, controls_.data(), 0, s);
            return;
        }
        StepArgs a = BaseArgs();
        a.arena_ids = ids;
        a.arena_ticks = ticks_list;
        a.count = count;
        a.arenas_per_wave = per_wave_;
        if (a.event_windows) {
            LaunchStep<true>(a, count, ticks, s);
        } else {
            LaunchStep<false>(a, count, ticks, s);
        }
    }

    // The LDS-staged kernels may use more than 64 KB of dynamic LDS.
    template <bool kEvents>
    static void AllowLargeLds(int limit) {
        const void* kernels[] = {reinterpret_cast<const void*>(&StepBallArenasKernel<kStageLdsScratch, kEvents>),
                                 reinterpret_cast<const void*>(&StepBallArenasKernel<kStageLds, kEvents>),
                                 reinterpret_cast<const void*>(&StepBallArenasGroupKernel<4, true, kEvents>),
                                 reinterpret_cast<const void*>(&StepBallArenasGroupKernel<8, true, kEvents>),
                                 reinterpret_cast<const void*>(&StepBallArenasGroupKernel<16, true, kEvents>),
                                 reinterpret_cast<const void*>(&StepBallArenasSplitKernel<32, kEvents>),
                                 reinterpret_cast<const void*>(&StepBallArenasSplitKernel<16, kEvents>)};
        for (const void* k : kernels) {
            Check(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, limit), "hipFuncSetAttribute");
        }
    }

    // Stages 6 and 7 over the whole batch: clear the list counts, sort the arenas by their last contact walk, and
    // step both lists in one launch (the heavy arenas' workgroups first, so they start first).
    // `lists` holds 2 env_count ids and `counts` 2 words; `gh` lanes per heavy arena (32 or 16), `block` lanes per
    // workgroup (64, 128 or 256).
    template <bool kEvents>
    void LaunchSplit(const StepArgs& a, int32_t ticks, int32_t* lists, int32_t* counts, int gh, int block,
                     hipStream_t s) {
        const int n = cfg_.env_count;
        // The heavy list holds at most env_count arenas (cap_divisor >= 1): the lists fit 2 env_count words.
        const int cap = n / split_cap_divisor_ > 1 ? n / split_cap_divisor_ : 1;
        hipLaunchKernelGGL(ClearSplitCountsKernel, dim3(1), dim3(64), 0, s, counts);
        hipLaunchKernelGGL(PartitionArenasKernel, dim3((n + 255) / 256), dim3(256), 0, s, internal_.data(), n,
                           split_rounds_, cap, lists, counts);
        const int heavy_blocks = (cap * gh + block - 1) / block;
        const int light_blocks = static_cast<int>((static_cast<int64_t>(n) * 8 + block - 1) / block);
        const size_t lds = sizeof(core::BallTickScratch) * static_cast<size_t>(block / 8);
        const dim3 grid(static_cast<unsigned>(heavy_blocks + light_blocks)), dims(block);
        if (gh == 32) {
            hipLaunchKernelGGL((StepBallArenasSplitKernel<32, kEvents>), grid, dims, lds, s, a, ticks, lists, counts,
                               cap, heavy_blocks);
        } else {
            hipLaunchKernelGGL((StepBallArenasSplitKernel<16, kEvents>), grid, dims, lds, s, a, ticks, lists, counts,
                               cap, heavy_blocks);
        }
        Check(hipGetLastError(), "StepBallArenasSplitKernel launch");
    }

    template <bool kEvents>
    void LaunchStep(const StepArgs& a, int32_t count, int32_t ticks, hipStream_t s) {
        if (IsSplitStage(stage_) && a.arena_ids == nullptr && a.arena_ticks == nullptr) {
            LaunchSplit<kEvents>(a, ticks, split_lists_.data(), split_counts_.data(),
                                 stage_ == kStageSplit32 ? 32 : 16, block_, s);
            return;
        }
        if (IsGroupStage(stage_)) {
            // 64 / per_wave lanes per arena (a subset launch of stage 6 or 7 runs as stage 5).
            const int lanes = 64 / per_wave_;
            const int64_t threads = static_cast<int64_t>(count) * lanes;
            const dim3 grid(static_cast<unsigned>((threads + block_ - 1) / block_)), block(block_);
            const size_t lds = LdsBytesPerLane(stage_) * static_cast<size_t>(block_ / lanes);
            const bool in_lds = stage_ == kStageGroupLds || IsSplitStage(stage_);
            switch (lanes) {
            case 4:
                if (in_lds) hipLaunchKernelGGL((StepBallArenasGroupKernel<4, true, kEvents>), grid, block, lds, s, a, ticks);
                else hipLaunchKernelGGL((StepBallArenasGroupKernel<4, false, kEvents>), grid, block, 0, s, a, ticks);
                break;
            case 8:
                if (in_lds) hipLaunchKernelGGL((StepBallArenasGroupKernel<8, true, kEvents>), grid, block, lds, s, a, ticks);
                else hipLaunchKernelGGL((StepBallArenasGroupKernel<8, false, kEvents>), grid, block, 0, s, a, ticks);
                break;
            case 16:
                if (in_lds) {
                    hipLaunchKernelGGL((StepBallArenasGroupKernel<16, true, kEvents>), grid, block, lds, s, a, ticks);
                } else {
                    hipLaunchKernelGGL((StepBallArenasGroupKernel<16, false, kEvents>), grid, block, 0, s, a, ticks);
                }
                break;
            default:
                throw InvalidArgument("stages 4 and 5 need 16, 8 or 4 arenas per wave (4, 8 or 16 lanes per arena)");
            }
            Check(hipGetLastError(), "StepBallArenasGroupKernel launch");
            return;
        }
        const int64_t threads = (static_cast<int64_t>(count) + per_wave_ - 1) / per_wave_ * 64;
        const int grid = static_cast<int>((threads + block_ - 1) / block_);
        const size_t lds = LdsBytesPerLane(stage_) * static_cast<size_t>(block_ / 64) * per_wave_;
        switch (stage_) {
        case kStageInPlace:
            hipLaunchKernelGGL((StepBallArenasKernel<kStageInPlace, kEvents>), dim3(grid), dim3(block_), 0, s, a, ticks);
            break;
        case kStagePrivate:
            hipLaunchKernelGGL((StepBallArenasKernel<kStagePrivate, kEvents>), dim3(grid), dim3(block_), 0, s, a, ticks);
            break;
        case kStageLdsScratch:
            hipLaunchKernelGGL((StepBallArenasKernel<kStageLdsScratch, kEvents>), dim3(grid), dim3(block_), lds, s, a,
                               ticks);
            break;
        default:
            hipLaunchKernelGGL((StepBallArenasKernel<kStageLds, kEvents>), dim3(grid), dim3(block_), lds, s, a, ticks);
            break;
        }
        Check(hipGetLastError(), "StepBallArenasKernel launch");
    }

    static void Expect(size_t n, size_t want, const char* what) {
        if (n != want) throw InvalidArgument(std::string(what) + ": count must equal the array size");
    }

    rg_sim_config cfg_;
    ArenaMeshHost mesh_host_;
    core::StepParams params_{};
    DeviceArray<core::MeshTriangle> tris_;
    DeviceArray<core::MeshBvhNode> nodes_;
    DeviceArray<core::MeshInfo> infos_;
    DeviceArray<BallState> balls_;
    DeviceArray<EnvState> envs_;
    DeviceArray<core::ArenaInternal> internal_;
    DeviceArray<int32_t> car_count_;
    DeviceArray<int32_t> car_order_;  // plain-launch lane order (RebuildCarOrder)
    DeviceArray<CarState> cars_;
    DeviceArray<CarControls> controls_;
    DeviceArray<WheelRuntimeState> wheels_;
    DeviceArray<uint32_t> unsupported_;
    std::vector<int32_t> host_counts_;
    bool has_cars_ = false;
    uint64_t seed_ = 0;
    CarStepArgs car_args_{};
    // Settled kickoff rows: the (config, team) groups seen so far, their rows on the host and the device, and each
    // car slot's group.
    struct KickoffGroup {
        CarConfig config;
        uint8_t team;
    };
    std::vector<KickoffGroup> kick_groups_;
    std::vector<CarState> kick_rows_car_;
    std::vector<WheelRuntimeState> kick_rows_wheels_;
    BallState kick_rows_ball_{};
    DeviceArray<CarState> kick_cars_;
    DeviceArray<WheelRuntimeState> kick_wheels_;
    DeviceArray<int32_t> kick_group_;
    DeviceArray<BallState> kick_ball_;
    int kick_spawns_ = 0;
    bool kick_dirty_ = true;
    DeviceArray<uint8_t> teams_;
    DeviceArray<CarConfig> configs_;
    DeviceArray<BoostPadState> pads_;
    DeviceArray<EnvCommand> commands_;
    DeviceArray<BallState> settled_;
    bool kickoff_settled_ = true;
    rg_mutator_config mutator_ = DefaultMutatorConfig();
    bool settled_ready_ = false;
    EventStream events_{};
    DeviceArray<uint64_t> digests_;
    DeviceArray<int32_t> split_lists_;   // stages 6 and 7: heavy then light arena ids
    DeviceArray<int32_t> split_counts_;
    int32_t* no_alloc_scratch_ = nullptr;
    size_t car_lds_allowed_ = 65536;  // dynamic LDS the LDS-staged car kernel may take without a larger attribute  // device address of `scratch` on this batch's device
    int split_rounds_ = kSplitRoundsDefault;
    int split_cap_divisor_ = kSplitCapDivisorDefault;
    int block_ = 64;
    int stage_ = kStageInPlace;
    int per_wave_ = 64;
    hipStream_t capture_stream_ = nullptr;
    hipGraphExec_t graph_exec_ = nullptr;
    int32_t graph_ticks_ = 0, graph_launches_ = 0, graph_shape_ = 0;
};

}  // namespace rg

struct rg_batch {
    std::unique_ptr<rg::BatchSim> sim;
};

namespace {

template <typename F>
int Guard(F&& f) {
    try {
        f();
        return RG_OK;
    } catch (const rg::InvalidArgument& e) {
        rg::g_last_error = e.what();
        return RG_EINVAL;
    } catch (const std::invalid_argument& e) {
        rg::g_last_error = e.what();
        return RG_EINVAL;
    } catch (const std::out_of_range& e) {
        rg::g_last_error = e.what();
        return RG_ERANGE;
    } catch (const std::exception& e) {
        rg::g_last_error = e.what();
        return RG_ERUNTIME;
    }
}

hipStream_t S(void* s) { return static_cast<hipStream_t>(s); }

rgw::ArenaRulesSpec RulesFromLayout(const rg_arena_layout& l) {
    rgw::ArenaRulesSpec r{};
    for (int t = 0; t < 2; ++t) {
        const float* g = l.goals[t];
        r.goals[t].point = rgw::V3{g[0], g[1], g[2]};
        r.goals[t].normal = rgw::V3{g[3], g[4], g[5]};
        r.goals[t].min = rgw::V3{g[6], g[7], g[8]};
        r.goals[t].max = rgw::V3{g[9], g[10], g[11]};
        const float len = std::sqrt(g[3] * g[3] + g[4] * g[4] + g[5] * g[5]);
        if (!(len > 0.0f) || !std::isfinite(len)) throw rg::InvalidArgument("ArenaGoal normal must be nonzero");
    }
    for (int t = 0; t < 2; ++t) {
        for (int k = 0; k < 5; ++k) {
            const float* q = l.kickoff_spawns[t][k];
            r.kickoff[t][k] = rgw::SpawnSpot{q[0], q[1], q[2], q[3]};
        }
        for (int k = 0; k < 4; ++k) {
            const float* q = l.respawn_spawns[t][k];
            r.respawn[t][k] = rgw::SpawnSpot{q[0], q[1], q[2], q[3]};
        }
    }
    r.ball_spawn = rgw::V3{l.ball_spawn[0], l.ball_spawn[1], l.ball_spawn[2]};
    if (l.pad_count < 0 || l.pad_count > rgw::kMaxPads) {
        throw rg::InvalidArgument("ArenaLayout::boost_pads holds at most kMaxBoostPads pads");
    }
    if (l.pad_count > 0 && (!l.pad_positions || !l.pad_is_big)) throw rg::InvalidArgument("null pad arrays");
    r.pad_count = l.pad_count;
    for (int k = 0; k < l.pad_count; ++k) {
        r.pad_pos[k] = rgw::V3{l.pad_positions[3 * k], l.pad_positions[3 * k + 1], l.pad_positions[3 * k + 2]};
        r.pad_big[k] = l.pad_is_big[k] ? 1 : 0;
    }
    return r;
}

}  // namespace

extern "C" {

const char* rg_last_error(void) { return rg::g_last_error.c_str(); }
void rg_set_last_error(const char* message) { rg::g_last_error = message ? message : ""; }
const char* rg_version(void) { return "rocketgpu_amd 0.2 (gfx950)"; }

int rg_batch_create(const rg_sim_config* config, const rg_arena_layout* layout, rg_batch** out) {
    return Guard([&] {
        if (!config || !layout || !out) throw rg::InvalidArgument("null argument");
        std::vector<std::vector<rg::Vec3>> verts(static_cast<size_t>(layout->mesh_count));
        std::vector<std::vector<int32_t>> idx(static_cast<size_t>(layout->mesh_count));
        size_t vo = 0, io = 0;
        for (int m = 0; m < layout->mesh_count; ++m) {
            for (int v = 0; v < layout->vertex_counts[m]; ++v, ++vo) {
                const float* p = layout->vertices_uu + 3 * vo;
                if (!(std::abs(p[0]) <= 11500.0f && std::abs(p[1]) <= 11500.0f && std::abs(p[2]) <= 11500.0f)) {
                    throw rg::InvalidArgument(
                        "ArenaLayout: a mesh vertex must be finite and within kArenaLayoutMaxCoord (11500 uu)");
                }
                verts[static_cast<size_t>(m)].push_back(rg::Vec3(p[0], p[1], p[2]));
            }
            for (int i = 0; i < layout->index_counts[m]; ++i, ++io) idx[static_cast<size_t>(m)].push_back(layout->indices[io]);
        }
        auto b = std::make_unique<rg_batch>();
        b->sim = std::make_unique<rg::BatchSim>(*config, rg::GroupLayoutMeshes(verts, idx), RulesFromLayout(*layout));
        *out = b.release();
    });
}

int rg_batch_create_procedural(const rg_sim_config* config, double scale, double fillet_radius, rg_batch** out) {
    return Guard([&] {
        if (!config || !out) throw rg::InvalidArgument("null argument");
        rgw::ArenaParams ap;
        ap.scale = scale;
        ap.fillet_radius = fillet_radius;
        const rgw::ArenaGeometry g = rgw::BuildProceduralArena(ap);
        std::vector<std::vector<rg::Vec3>> verts(1);
        for (const rgw::V3& v : g.vertices) verts[0].push_back(rg::Vec3(v.x, v.y, v.z));
        std::vector<std::vector<int32_t>> idx(1, g.indices);
        auto b = std::make_unique<rg_batch>();
        b->sim = std::make_unique<rg::BatchSim>(*config, rg::GroupLayoutMeshes(verts, idx),
                                                rgw::ProceduralArenaRules(ap));
        *out = b.release();
    });
}

void rg_batch_destroy(rg_batch* batch) { delete batch; }

int64_t rg_batch_env_count(const rg_batch* b) { return b ? b->sim->env_count() : -1; }
int64_t rg_batch_car_capacity(const rg_batch* b) { return b ? b->sim->car_capacity() : -1; }

int rg_batch_set_env_car_count(rg_batch* b, const int32_t* counts, size_t count, void* stream) {
    return Guard([&] { b->sim->SetEnvCarCount(counts, count, S(stream)); });
}

int rg_batch_upload(rg_batch* b, int id, const void* src, size_t count, void* stream) {
    return Guard([&] { b->sim->Upload(id, src, count, S(stream)); });
}

int rg_batch_upload_mode(rg_batch* b, int id, const void* src, size_t count, int32_t mode, void* stream) {
    return Guard([&] {
        if (!b || !b->sim) throw rg::InvalidArgument("null batch");
        if (mode != RG_UPLOAD_PUBLIC_STATE && mode != RG_UPLOAD_PRESERVE_INTERNAL) {
            throw rg::InvalidArgument("upload mode must be RG_UPLOAD_PUBLIC_STATE or RG_UPLOAD_PRESERVE_INTERNAL");
        }
        if (count && !src) throw rg::InvalidArgument("host_src must not be null");
        b->sim->Upload(id, src, count, S(stream), mode == RG_UPLOAD_PRESERVE_INTERNAL);
    });
}

int rg_batch_download(rg_batch* b, int id, void* dst, size_t count, void* stream) {
    return Guard([&] { b->sim->Download(id, dst, count, S(stream)); });
}

int rg_batch_download_async(rg_batch* b, int id, void* dst, size_t count, void* stream) {
    return Guard([&] { b->sim->Download(id, dst, count, S(stream), false); });
}

int rg_batch_copy_state_from(rg_batch* b, const rg_batch* src, void* stream) {
    return Guard([&] {
        if (!src) throw rg::InvalidArgument("source batch must not be null");
        b->sim->CopyStateFrom(*src->sim, S(stream));
    });
}

void rg_mutator_config_defaults(rg_mutator_config* out) {
    if (out) *out = rg::DefaultMutatorConfig();
}

int rg_batch_set_mutator_config(rg_batch* b, const rg_mutator_config* config) {
    return Guard([&] {
        if (!config) throw rg::InvalidArgument("config must not be null");
        b->sim->SetMutatorConfig(*config);
    });
}

int rg_batch_get_mutator_config(const rg_batch* b, rg_mutator_config* out) {
    return Guard([&] {
        if (!out) throw rg::InvalidArgument("out must not be null");
        *out = b->sim->MutatorConfig();
    });
}

int rg_batch_step(rg_batch* b, int32_t ticks, void* stream) {
    return Guard([&] { b->sim->Step(ticks, S(stream)); });
}

int rg_batch_step_split_no_alloc(rg_batch* b, int32_t ticks, void* stream) {
    return Guard([&] { b->sim->StepSplitNoAlloc(ticks, S(stream)); });
}

int rg_batch_step_with_device_controls(rg_batch* b, const void* controls, int32_t ticks, int32_t stride,
                                       void* stream) {
    return Guard([&] { b->sim->StepWithDeviceControls(controls, ticks, stride, S(stream)); });
}

int64_t rg_batch_boost_pad_capacity(const rg_batch* b) { return b ? b->sim->boost_pad_capacity() : -1; }

int rg_batch_set_teams(rg_batch* b, const uint8_t* host_teams, size_t count, void* stream) {
    return Guard([&] { b->sim->SetTeams(host_teams, count, S(stream)); });
}

int rg_batch_set_car_configs(rg_batch* b, const void* host_configs, size_t count, void* stream) {
    return Guard([&] { b->sim->SetCarConfigs(static_cast<const rg::CarConfig*>(host_configs), count, S(stream)); });
}

int rg_batch_set_kickoff_settled(rg_batch* b, int32_t on) {
    return Guard([&] { b->sim->SetKickoffSettled(on != 0); });
}

int rg_batch_set_split(rg_batch* b, int32_t rounds, int32_t cap_divisor) {
    return Guard([&] { b->sim->SetSplit(rounds, cap_divisor); });
}

int rg_batch_set_physics_event_stream(rg_batch* b, void* device_records, void* device_windows,
                                      uint32_t capacity_per_env, uint32_t env_count, uint32_t cars_per_env,
                                      int32_t goal_only) {
    return Guard([&] {
        rg::EventStream st;
        st.records = static_cast<rg::PhysicsEvent*>(device_records);
        st.windows = static_cast<rg::PhysicsEventWindow*>(device_windows);
        st.capacity_per_env = capacity_per_env;
        st.env_count = env_count;
        st.cars_per_env = cars_per_env;
        st.goal_only = goal_only != 0;
        b->sim->SetPhysicsEventStream(st);
    });
}

int rg_batch_state_snapshot_size(const rg_batch* b, size_t* bytes) {
    return Guard([&] {
        if (!bytes) throw rg::InvalidArgument("bytes must not be null");
        *bytes = b->sim->SnapshotBytes();
    });
}

int rg_batch_save_state(rg_batch* b, void* host_dst, size_t bytes, void* stream) {
    return Guard([&] { b->sim->SaveState(host_dst, bytes, S(stream)); });
}

int rg_batch_load_state(rg_batch* b, const void* host_src, size_t bytes, void* stream) {
    return Guard([&] { b->sim->LoadState(host_src, bytes, S(stream)); });
}

int rg_batch_reset_to_kickoff(rg_batch* b, uint64_t seed, void* stream) {
    return Guard([&] { b->sim->ResetToKickoff(seed, S(stream)); });
}

int rg_batch_apply_env_commands(rg_batch* b, const void* host_commands, size_t count, void* stream) {
    return Guard([&] {
        b->sim->ApplyEnvCommands(static_cast<const rg::EnvCommand*>(host_commands), count, S(stream));
    });
}

int rg_batch_apply_env_commands_device(rg_batch* b, const void* device_commands, size_t count, void* stream) {
    return Guard([&] {
        b->sim->ApplyEnvCommandsDevice(static_cast<const rg::EnvCommand*>(device_commands), count, S(stream));
    });
}

int rg_batch_step_with_device_commands(rg_batch* b, const void* device_controls_by_tick, int32_t control_stride,
                                       const void* device_env_commands_by_tick, int32_t env_command_stride,
                                       int32_t ticks, void* stream) {
    return Guard([&] {
        b->sim->StepWithDeviceCommands(static_cast<const rg::CarControls*>(device_controls_by_tick), control_stride,
                                       static_cast<const rg::EnvCommand*>(device_env_commands_by_tick),
                                       env_command_stride, ticks, S(stream));
    });
}

int rg_batch_step_arenas(rg_batch* b, const int32_t* device_arena_ids, const int32_t* device_ticks, int32_t count,
                         int32_t ticks, void* stream) {
    return Guard([&] { b->sim->StepArenas(device_arena_ids, device_ticks, count, ticks, S(stream)); });
}

int rg_batch_step_graph(rg_batch* b, int32_t ticks_per_launch, int32_t launches, void* stream) {
    return Guard([&] { b->sim->StepGraph(ticks_per_launch, launches, S(stream)); });
}

int rg_batch_set_launch_shape(rg_batch* b, int32_t block, int32_t stage, int32_t arenas_per_wave) {
    return Guard([&] {
        const int old_block = b->sim->block(), old_stage = b->sim->stage(), old_per_wave = b->sim->arenas_per_wave();
        try {
            b->sim->SetBlock(block);
            b->sim->SetStage(stage);
            b->sim->SetArenasPerWave(arenas_per_wave);
            b->sim->CheckLaunchShape();
        } catch (...) {
            b->sim->SetBlock(old_block);
            b->sim->SetStage(old_stage);
            b->sim->SetArenasPerWave(old_per_wave);
            throw;
        }
    });
}

int rg_batch_synchronize(rg_batch* b, void* stream) {
    (void)b;
    return Guard([&] { rg::Check(hipStreamSynchronize(S(stream)), "hipStreamSynchronize"); });
}

int rg_batch_device_view(rg_batch* b, rg_device_view* out) {
    return Guard([&] { b->sim->DeviceView(out); });
}

int rg_batch_ball_path(rg_batch* b, int32_t samples, int32_t ticks, float goal_y, float* device_out,
                       int32_t* device_result, void* stream) {
    return Guard([&] { b->sim->BallPath(samples, ticks, goal_y, device_out, device_result, S(stream)); });
}

int rg_batch_env_digests(rg_batch* b, uint64_t* device_out, void* stream) {
    return Guard([&] { b->sim->EnvDigests(device_out, S(stream)); });
}

int rg_car_preset_config(int32_t preset, void* out) {
    return Guard([&] {
        if (preset < 0 || preset >= rg::kPresetCount) throw rg::InvalidArgument("car preset must be 0..7");
        if (!out) throw rg::InvalidArgument("out must not be null");
        const rg::CarConfig c = rg::CarPresetConfig(preset);
        std::memcpy(out, &c, sizeof(rg::CarConfig));
    });
}

int rg_workload_request_controls(uint64_t seed, int64_t request, int64_t first_arena, int64_t arena_count,
                                 int32_t cars_per_arena, float p_boost, void* out) {
    return Guard([&] {
        if (request < 0 || first_arena < 0 || arena_count < 0 || cars_per_arena < 0 ||
            (arena_count * cars_per_arena > 0 && !out)) {
            throw rg::InvalidArgument("bad request controls arguments");
        }
        rgw::Controls* o = static_cast<rgw::Controls*>(out);
        for (int64_t a = 0; a < arena_count; ++a) {
            for (int32_t k = 0; k < cars_per_arena; ++k) {
                o[a * cars_per_arena + k] = rgw::RequestControls(seed, static_cast<uint64_t>(request),
                                                                 static_cast<uint64_t>(first_arena + a),
                                                                 static_cast<uint64_t>(k), p_boost);
            }
        }
    });
}

int64_t rg_batch_arena_bytes(const rg_batch* b, int32_t cars) {
    if (!b || cars < 0) return -1;
    const int64_t pads = rg_batch_boost_pad_capacity(b) / rg_batch_env_count(b);
    return static_cast<int64_t>(sizeof(rg::BallState) + sizeof(rg::EnvState) + sizeof(rg::core::ArenaInternal) +
                                sizeof(int32_t)) +
           pads * static_cast<int64_t>(sizeof(rg::BoostPadState)) +
           cars * static_cast<int64_t>(sizeof(rg::CarState) + sizeof(rg::WheelRuntimeState) +
                                       sizeof(rg::CarControls) + sizeof(rg::CarConfig) + 1);
}

uint64_t rg_fold_digests(const uint64_t* d, int64_t n) {
    return n > 0 && d ? rgw::Fnv1a(rgw::kFnvOffset, d, static_cast<size_t>(n) * sizeof(uint64_t)) : rgw::kFnvOffset;
}

int rg_workload_controls(uint64_t seed, int64_t first_arena, int64_t arena_count, int32_t cars_per_arena,
                         int32_t ticks, float p_boost, void* out) {
    return Guard([&] {
        if (arena_count < 0 || cars_per_arena < 0 || ticks < 0 ||
            (arena_count * cars_per_arena * ticks > 0 && !out)) {
            throw rg::InvalidArgument("bad workload request");
        }
        const int64_t cars = arena_count * cars_per_arena;
        rgw::Controls* o = static_cast<rgw::Controls*>(out);
        for (int64_t c = 0; c < cars; ++c) {
            rgw::ControlStream cs;
            cs.seed(seed, static_cast<uint64_t>(first_arena + c / cars_per_arena),
                    static_cast<uint64_t>(c % cars_per_arena), p_boost);
            for (int32_t t = 0; t < ticks; ++t) o[static_cast<int64_t>(t) * cars + c] = cs.next();
        }
    });
}

int rg_workload_ball_states(uint64_t seed, int64_t first_arena, int64_t count, double scale, void* out) {
    return Guard([&] {
        if (count < 0 || (count > 0 && !out)) throw rg::InvalidArgument("bad workload request");
        rg::BallState* b = static_cast<rg::BallState*>(out);
        for (int64_t i = 0; i < count; ++i) {
            const rgw::BallLaunch l = rgw::SeededBallLaunch(seed, static_cast<uint64_t>(first_arena + i), scale);
            std::memset(static_cast<void*>(&b[i]), 0, sizeof(rg::BallState));
            rg::BallState* s = new (&b[i]) rg::BallState;
            s->pos = rg::Vec3(l.pos[0], l.pos[1], l.pos[2]);
            s->vel = rg::Vec3(l.vel[0], l.vel[1], l.vel[2]);
            s->ang_vel = rg::Vec3(l.ang_vel[0], l.ang_vel[1], l.ang_vel[2]);
        }
    });
}

int64_t rg_state_layouts_json(char* buf, size_t capacity) {
    // {"name": {"itemsize": n, "fields": [[field, typestr, offset, [shape...]], ...]}, ...}; scalars have shape [],
    // as the reference's state_layouts() gives them (python/rocketgpu_bindings.cpp:543-560).
    std::string j = "{";
    bool first_field = true;
    const auto begin = [&](const char* name, size_t itemsize, bool first) {
        j += std::string(first ? "" : "]}, ") + "\"" + name + "\": {\"itemsize\": " + std::to_string(itemsize) +
             ", \"fields\": [";
        first_field = true;
    };
    const auto field = [&](const char* name, const char* type, size_t off, int n) {
        char tmp[192];
        std::snprintf(tmp, sizeof(tmp), "%s[\"%s\", \"%s\", %zu, [%s]]", first_field ? "" : ", ", name, type, off,
                      n == 0 ? "" : std::to_string(n).c_str());
        j += tmp;
        first_field = false;
    };
    // A 2-D field (an array of vectors): shape [n, m].
    const auto field2 = [&](const char* name, const char* type, size_t off, int n, int m) {
        char tmp[192];
        std::snprintf(tmp, sizeof(tmp), "%s[\"%s\", \"%s\", %zu, [%d, %d]]", first_field ? "" : ", ", name, type, off,
                      n, m);
        j += tmp;
        first_field = false;
    };
#define RG_F(T, f, type, n) field(#f, type, offsetof(T, f), n)
#define RG_SUB(T, m, S, f, type) field(#m "." #f, type, offsetof(T, m) + offsetof(T::S, f), 0)
    begin("ball_states", sizeof(rg::BallState), true);
    RG_SUB(rg::BallState, hs_info, HeatseekerInfo, y_target_dir, "<f4");
    RG_SUB(rg::BallState, hs_info, HeatseekerInfo, cur_target_speed, "<f4");
    RG_SUB(rg::BallState, hs_info, HeatseekerInfo, time_since_hit, "<f4");
    RG_SUB(rg::BallState, ds_info, DropshotInfo, charge_level, "<i4");
    RG_SUB(rg::BallState, ds_info, DropshotInfo, accumulated_hit_force, "<f4");
    RG_SUB(rg::BallState, ds_info, DropshotInfo, y_target_dir, "<f4");
    RG_SUB(rg::BallState, ds_info, DropshotInfo, has_damaged, "|u1");
    RG_SUB(rg::BallState, ds_info, DropshotInfo, last_damage_tick, "<u8");
    RG_F(rg::BallState, pos, "<f4", 3);
    RG_F(rg::BallState, rot, "<f4", 4);
    RG_F(rg::BallState, vel, "<f4", 3);
    RG_F(rg::BallState, ang_vel, "<f4", 3);
    RG_F(rg::BallState, vel_impulse_cache, "<f4", 3);
    RG_F(rg::BallState, tick_count_when_extra_impulse_applied, "<u8", 0);
    RG_F(rg::BallState, tick_count_since_update, "<u8", 0);
    RG_F(rg::BallState, pos_bt, "<f4", 3);
    RG_F(rg::BallState, vel_bt, "<f4", 3);
    RG_F(rg::BallState, basis_forward, "<f4", 3);
    RG_F(rg::BallState, basis_right, "<f4", 3);
    RG_F(rg::BallState, basis_up, "<f4", 3);
    RG_F(rg::BallState, basis_authoritative, "<u4", 0);
    RG_F(rg::BallState, vel_impulse_cache_bt, "<f4", 3);
    RG_F(rg::BallState, sleeping, "<u4", 0);
    RG_F(rg::BallState, deactivation_time, "<f4", 0);
    RG_F(rg::BallState, constrained_2d, "<u2", 0);
    begin("env_states", sizeof(rg::EnvState), false);
    RG_F(rg::EnvState, tick, "<u4", 0);
    RG_F(rg::EnvState, blue_score, "<u4", 0);
    RG_F(rg::EnvState, orange_score, "<u4", 0);
    RG_F(rg::EnvState, last_scored_team, "<i4", 0);
    RG_F(rg::EnvState, goal_scored_this_tick, "|u1", 0);
    RG_F(rg::EnvState, car_visit_order_mode, "|u1", 0);
    RG_F(rg::EnvState, car_visit_present_mask, "|u1", 0);
    RG_F(rg::EnvState, kickoff_resets, "|u1", 0);
    RG_F(rg::EnvState, car_visit_order_packed, "<u4", 0);
    RG_F(rg::EnvState, countdown_phase, "|u1", 0);
    RG_F(rg::EnvState, countdown_fires, "|u1", 0);
    RG_F(rg::EnvState, frame_ticks, "<u2", 0);
    RG_F(rg::EnvState, countdown_count, "<f4", 0);
    RG_F(rg::EnvState, frame_frac_prev, "<f4", 0);
    RG_F(rg::EnvState, kickoff_match_order, "<u2", 0);
    RG_F(rg::EnvState, kickoff_match_index, "|u1", 0);
    RG_F(rg::EnvState, goal_latched, "|u1", 0);
    RG_F(rg::EnvState, respawn_match_order, "|u1", 0);
    begin("car_states", sizeof(rg::CarState), false);
field("pos", "<f4", offsetof(rg::CarState, pos), 3);
    field("rot", "<f4", offsetof(rg::CarState, rot), 4);
    field("vel", "<f4", offsetof(rg::CarState, vel), 3);
    field("ang_vel", "<f4", offsetof(rg::CarState, ang_vel), 3);
    field("flip_rel_torque", "<f4", offsetof(rg::CarState, flip_rel_torque), 3);
    field("jump_time", "<f4", offsetof(rg::CarState, jump_time), 0);
    field("flip_time", "<f4", offsetof(rg::CarState, flip_time), 0);
    field("air_time", "<f4", offsetof(rg::CarState, air_time), 0);
    field("air_time_since_jump", "<f4", offsetof(rg::CarState, air_time_since_jump), 0);
    field("boost", "<f4", offsetof(rg::CarState, boost), 0);
    field("time_since_boosted", "<f4", offsetof(rg::CarState, time_since_boosted), 0);
    field("boosting_time", "<f4", offsetof(rg::CarState, boosting_time), 0);
    field("supersonic_time", "<f4", offsetof(rg::CarState, supersonic_time), 0);
    field("handbrake_val", "<f4", offsetof(rg::CarState, handbrake_val), 0);
    field("auto_flip_timer", "<f4", offsetof(rg::CarState, auto_flip_timer), 0);
    field("boost_amount", "<f4", offsetof(rg::CarState, boost_amount), 0);
    field("demo_respawn_timer", "<f4", offsetof(rg::CarState, demo_respawn_timer), 0);
    field("tick_count_since_update", "<u8", offsetof(rg::CarState, tick_count_since_update), 0);
    field("ball_contact_last_tick", "|u1", offsetof(rg::CarState, ball_contact_last_tick), 0);
    field("ball_contact_this_tick", "|u1", offsetof(rg::CarState, ball_contact_this_tick), 0);
    field("is_on_ground", "|u1", offsetof(rg::CarState, is_on_ground), 0);
    field("wheels_with_contact", "|u1", offsetof(rg::CarState, wheels_with_contact), 4);
    field("has_jumped", "|u1", offsetof(rg::CarState, has_jumped), 0);
    field("has_double_jumped", "|u1", offsetof(rg::CarState, has_double_jumped), 0);
    field("has_flipped", "|u1", offsetof(rg::CarState, has_flipped), 0);
    field("is_flipping", "|u1", offsetof(rg::CarState, is_flipping), 0);
    field("is_jumping", "|u1", offsetof(rg::CarState, is_jumping), 0);
    field("is_boosting", "|u1", offsetof(rg::CarState, is_boosting), 0);
    field("is_supersonic", "|u1", offsetof(rg::CarState, is_supersonic), 0);
    field("is_auto_flipping", "|u1", offsetof(rg::CarState, is_auto_flipping), 0);
    field("is_demoed", "|u1", offsetof(rg::CarState, is_demoed), 0);
    field("just_respawned", "|u1", offsetof(rg::CarState, just_respawned), 0);
    field("demo_no_contact_response", "|u1", offsetof(rg::CarState, demo_no_contact_response), 0);
    field("respawn_slot_override_plus1", "|u1", offsetof(rg::CarState, respawn_slot_override_plus1), 0);
    field("rlbotv5_boost_hold", "|u1", offsetof(rg::CarState, rlbotv5_boost_hold), 0);
    field("world_contact.has_contact", "|u1", offsetof(rg::CarState, world_contact) + offsetof(decltype(rg::CarState::world_contact), has_contact), 0);
    field("world_contact.wheel_ball_mask", "|u1", offsetof(rg::CarState, world_contact) + offsetof(decltype(rg::CarState::world_contact), wheel_ball_mask), 0);
    field("world_contact.reserved1", "|u1", offsetof(rg::CarState, world_contact) + offsetof(decltype(rg::CarState::world_contact), reserved1), 0);
    field("world_contact.reserved2", "|u1", offsetof(rg::CarState, world_contact) + offsetof(decltype(rg::CarState::world_contact), reserved2), 0);
    field("world_contact.contact_normal", "<f4", offsetof(rg::CarState, world_contact) + offsetof(decltype(rg::CarState::world_contact), contact_normal), 3);
    field("car_contact.other_car_id", "<u4", offsetof(rg::CarState, car_contact) + offsetof(decltype(rg::CarState::car_contact), other_car_id), 0);
    field("car_contact.cooldown_timer", "<f4", offsetof(rg::CarState, car_contact) + offsetof(decltype(rg::CarState::car_contact), cooldown_timer), 0);
    field("vel_impulse_cache", "<f4", offsetof(rg::CarState, vel_impulse_cache), 3);
    field("ball_hit_info.is_valid", "|u1", offsetof(rg::CarState, ball_hit_info) + offsetof(rg::BallHitInfo, is_valid), 0);
    field("ball_hit_info.extra_pending", "|u1", offsetof(rg::CarState, ball_hit_info) + offsetof(rg::BallHitInfo, extra_pending), 0);
    field("ball_hit_info.extra_tick_is_previous", "|u1", offsetof(rg::CarState, ball_hit_info) + offsetof(rg::BallHitInfo, extra_tick_is_previous), 0);
    field("ball_hit_info.tick_count_when_hit", "<u8", offsetof(rg::CarState, ball_hit_info) + offsetof(rg::BallHitInfo, tick_count_when_hit), 0);
    field("ball_hit_info.tick_count_when_extra_impulse_applied", "<u8", offsetof(rg::CarState, ball_hit_info) + offsetof(rg::BallHitInfo, tick_count_when_extra_impulse_applied), 0);
    field("ball_hit_info.hit_car_id", "<u4", offsetof(rg::CarState, ball_hit_info) + offsetof(rg::BallHitInfo, hit_car_id), 0);
    field("ball_hit_info.relative_pos_on_ball", "<f4", offsetof(rg::CarState, ball_hit_info) + offsetof(rg::BallHitInfo, relative_pos_on_ball), 3);
    field("ball_hit_info.ball_pos", "<f4", offsetof(rg::CarState, ball_hit_info) + offsetof(rg::BallHitInfo, ball_pos), 3);
    field("ball_hit_info.extra_hit_vel", "<f4", offsetof(rg::CarState, ball_hit_info) + offsetof(rg::BallHitInfo, extra_hit_vel), 3);
    field("last_controls.throttle", "<f4", offsetof(rg::CarState, last_controls) + offsetof(rg::CarControls, throttle), 0);
    field("last_controls.steer", "<f4", offsetof(rg::CarState, last_controls) + offsetof(rg::CarControls, steer), 0);
    field("last_controls.pitch", "<f4", offsetof(rg::CarState, last_controls) + offsetof(rg::CarControls, pitch), 0);
    field("last_controls.yaw", "<f4", offsetof(rg::CarState, last_controls) + offsetof(rg::CarControls, yaw), 0);
    field("last_controls.roll", "<f4", offsetof(rg::CarState, last_controls) + offsetof(rg::CarControls, roll), 0);
    field("last_controls.jump", "|u1", offsetof(rg::CarState, last_controls) + offsetof(rg::CarControls, jump), 0);
    field("last_controls.boost", "|u1", offsetof(rg::CarState, last_controls) + offsetof(rg::CarControls, boost), 0);
    field("last_controls.handbrake", "|u1", offsetof(rg::CarState, last_controls) + offsetof(rg::CarControls, handbrake), 0);
    field("pos_bt", "<f4", offsetof(rg::CarState, pos_bt), 3);
    field("vel_bt", "<f4", offsetof(rg::CarState, vel_bt), 3);
    field("tick_start_pos_bt", "<f4", offsetof(rg::CarState, tick_start_pos_bt), 3);
    field("pad_frame_pos_bt", "<f4", offsetof(rg::CarState, pad_frame_pos_bt), 3);
    field("vel_touched_solve_stage", "|u1", offsetof(rg::CarState, vel_touched_solve_stage), 0);
    field("basis_authoritative", "|u1", offsetof(rg::CarState, basis_authoritative), 0);
    field("ext_force_impulse_bt_valid", "|u1", offsetof(rg::CarState, ext_force_impulse_bt_valid), 0);
    field("pad_frame_pos_valid", "|u1", offsetof(rg::CarState, pad_frame_pos_valid), 0);
    field("basis_forward", "<f4", offsetof(rg::CarState, basis_forward), 3);
    field("basis_right", "<f4", offsetof(rg::CarState, basis_right), 3);
    field("basis_up", "<f4", offsetof(rg::CarState, basis_up), 3);
    field("ext_force_impulse_bt", "<f4", offsetof(rg::CarState, ext_force_impulse_bt), 3);
    field("ext_torque_impulse", "<f4", offsetof(rg::CarState, ext_torque_impulse), 3);
    field("solve_delta_linear_bt", "<f4", offsetof(rg::CarState, solve_delta_linear_bt), 3);
    field("vel_impulse_cache_bt", "<f4", offsetof(rg::CarState, vel_impulse_cache_bt), 3);
    field("solve_delta_angular", "<f4", offsetof(rg::CarState, solve_delta_angular), 3);
    field("bump_event_bumps", "|u1", offsetof(rg::CarState, bump_event_bumps), 7);
    field("bump_event_demos", "|u1", offsetof(rg::CarState, bump_event_demos), 7);
    field("boost_pad_touching", "|u1", offsetof(rg::CarState, boost_pad_touching), 0);
    field("bump_event_reserved1", "|u1", offsetof(rg::CarState, bump_event_reserved1), 0);
    field("previous_ball_contact_start", "<u8", offsetof(rg::CarState, previous_ball_contact_start), 0);
    field("last_ball_contact_start", "<u8", offsetof(rg::CarState, last_ball_contact_start), 0);
    field("last_world_surface_contact", "<u8", offsetof(rg::CarState, last_world_surface_contact), 0);
    field("first_world_after_previous_ball", "<u8", offsetof(rg::CarState, first_world_after_previous_ball), 0);
    field("first_world_after_last_ball", "<u8", offsetof(rg::CarState, first_world_after_last_ball), 0);
    field("last_flip_or_jump_start", "<u8", offsetof(rg::CarState, last_flip_or_jump_start), 0);
    field("last_directional_flip_start", "<u8", offsetof(rg::CarState, last_directional_flip_start), 0);
    field("last_ball_started_with_world", "|u1", offsetof(rg::CarState, last_ball_started_with_world), 0);
    field("flip_or_jump_last_tick", "|u1", offsetof(rg::CarState, flip_or_jump_last_tick), 0);
    field("internal_jump_time", "<f4", offsetof(rg::CarState, internal_jump_time), 0);
    field("internal_flip_time", "<f4", offsetof(rg::CarState, internal_flip_time), 0);
    field("flip_timer_live", "|u1", offsetof(rg::CarState, flip_timer_live), 0);
    field("air_state", "|u1", offsetof(rg::CarState, air_state), 0);
    field("driving_disabled", "<u2", offsetof(rg::CarState, driving_disabled), 0);
    field("dodge_timeout", "<f4", offsetof(rg::CarState, dodge_timeout), 0);
    field("dodge_elapsed", "<f4", offsetof(rg::CarState, dodge_elapsed), 0);
    field("dodge_dir_x", "<f4", offsetof(rg::CarState, dodge_dir_x), 0);
    field("dodge_dir_y", "<f4", offsetof(rg::CarState, dodge_dir_y), 0);
    field("latest_touch_tick", "<u8", offsetof(rg::CarState, latest_touch_tick), 0);
    field("latest_touch_location", "<f4", offsetof(rg::CarState, latest_touch_location), 3);
    field("latest_touch_normal", "<f4", offsetof(rg::CarState, latest_touch_normal), 3);
    field("internal_vel", "<f4", offsetof(rg::CarState, internal_vel), 3);
    field("internal_ang_vel", "<f4", offsetof(rg::CarState, internal_ang_vel), 3);
    field("reported_vel", "<f4", offsetof(rg::CarState, reported_vel), 3);
    field("reported_ang_vel", "<f4", offsetof(rg::CarState, reported_ang_vel), 3);
    field("internal_has_jumped", "|u1", offsetof(rg::CarState, internal_has_jumped), 0);
    field("internal_has_double_jumped", "|u1", offsetof(rg::CarState, internal_has_double_jumped), 0);
    field("internal_has_flipped", "|u1", offsetof(rg::CarState, internal_has_flipped), 0);
    field("internal_is_supersonic", "|u1", offsetof(rg::CarState, internal_is_supersonic), 0);
    field("reported_flags", "|u1", offsetof(rg::CarState, reported_flags), 0);
    field("game_has_jumped", "|u1", offsetof(rg::CarState, game_has_jumped), 0);
    field("double_jump_ticks", "|u1", offsetof(rg::CarState, double_jump_ticks), 0);
    field("jump_hold_broken", "|u1", offsetof(rg::CarState, jump_hold_broken), 0);
    field("supersonic_set_v3", "|u1", offsetof(rg::CarState, supersonic_set_v3), 0);
    field("rlbotv5_dodged", "|u1", offsetof(rg::CarState, rlbotv5_dodged), 0);
    field2("set_wheel_normals", "<f4", offsetof(rg::CarState, set_wheel_normals), 4, 3);
    field("set_wheel_normals_mask", "|u1", offsetof(rg::CarState, set_wheel_normals_mask), 0);
    field("cal_pre_flags", "|u1", offsetof(rg::CarState, cal_pre_flags), 0);
    field("cal_pre_forward_speed_uu", "<f4", offsetof(rg::CarState, cal_pre_forward_speed_uu), 0);
    field("cal_pre_pitchlock_restart_time", "<f4", offsetof(rg::CarState, cal_pre_pitchlock_restart_time), 0);
    field("cal_tick_start_ang_vel", "<f4", offsetof(rg::CarState, cal_tick_start_ang_vel), 3);
    field("tick_start_basis_forward", "<f4", offsetof(rg::CarState, tick_start_basis_forward), 3);
    field("tick_start_basis_right", "<f4", offsetof(rg::CarState, tick_start_basis_right), 3);
    field("tick_start_basis_up", "<f4", offsetof(rg::CarState, tick_start_basis_up), 3);
    field2("restored_inv_inertia_bt", "<f4", offsetof(rg::CarState, restored_inv_inertia_bt), 3, 3);
    field("restored_inv_inertia_valid", "|u1", offsetof(rg::CarState, restored_inv_inertia_valid), 0);
    field2("pending_bump_bt", "<f4", offsetof(rg::CarState, pending_bump_bt), 7, 3);
    field("pending_bump_hitter", "<u4", offsetof(rg::CarState, pending_bump_hitter), 7);
    field("pending_bump_count", "|u1", offsetof(rg::CarState, pending_bump_count), 0);
    field("auto_flip_torque_scale", "<f4", offsetof(rg::CarState, auto_flip_torque_scale), 0);
    begin("controls", sizeof(rg::CarControls), false);
    RG_F(rg::CarControls, throttle, "<f4", 0);
    RG_F(rg::CarControls, steer, "<f4", 0);
    RG_F(rg::CarControls, pitch, "<f4", 0);
    RG_F(rg::CarControls, yaw, "<f4", 0);
    RG_F(rg::CarControls, roll, "<f4", 0);
    RG_F(rg::CarControls, jump, "|u1", 0);
    RG_F(rg::CarControls, boost, "|u1", 0);
    RG_F(rg::CarControls, handbrake, "|u1", 0);
    begin("car_configs", sizeof(rg::CarConfig), false);
    RG_F(rg::CarConfig, hitbox_size, "<f4", 3);
    RG_F(rg::CarConfig, hitbox_published, "<f4", 3);
    RG_F(rg::CarConfig, hitbox_offset, "<f4", 3);
    RG_F(rg::CarConfig, front_wheels_offset, "<f4", 3);
    RG_F(rg::CarConfig, back_wheels_offset, "<f4", 3);
    RG_F(rg::CarConfig, front_wheel_radius, "<f4", 0);
    RG_F(rg::CarConfig, back_wheel_radius, "<f4", 0);
    RG_F(rg::CarConfig, front_suspension_rest, "<f4", 0);
    RG_F(rg::CarConfig, back_suspension_rest, "<f4", 0);
    RG_F(rg::CarConfig, dodge_deadzone, "<f4", 0);
    RG_F(rg::CarConfig, three_wheels, "|u1", 0);
    RG_F(rg::CarConfig, body, "<u2", 0);
    RG_F(rg::CarConfig, steer_angle_curve_deg, "<f4", 6);
    RG_F(rg::CarConfig, powerslide_steer_angle_curve_deg, "<f4", 2);
    RG_F(rg::CarConfig, spawn_z, "<f4", 0);
    begin("teams", 1, false);
    field("team", "|u1", 0, 0);
    begin("env_car_count", sizeof(int32_t), false);
    field("count", "<i4", 0, 0);
    begin("boost_pad_states", sizeof(rg::BoostPadState), false);
    RG_F(rg::BoostPadState, cooldown, "<f4", 0);
    RG_F(rg::BoostPadState, is_active, "|u1", 0);
    RG_F(rg::BoostPadState, is_big, "|u1", 0);
    RG_F(rg::BoostPadState, pickup_frac_q, "<u2", 0);
    RG_F(rg::BoostPadState, prev_locked_car_id, "<u4", 0);
    RG_F(rg::BoostPadState, cur_locked_car_id, "<u4", 0);
    // The MutatorConfig record of rg_batch_set_mutator_config (a host record, not a device array).
    begin("mutator_config", sizeof(rg_mutator_config), false);
#define RG_MUT_LAYOUT(type, name, def, target) \
    field(#name, std::is_same<type, float>::value ? "<f4" : "|u1", offsetof(rg_mutator_config, name), 0);
    RG_MUTATOR_FIELDS(RG_MUT_LAYOUT)
#undef RG_MUT_LAYOUT
#undef RG_SUB
#undef RG_F
    j += "]}}";
    if (buf && capacity > 0) {
        const size_t n = std::min(capacity - 1, j.size());
        std::memcpy(buf, j.data(), n);
        buf[n] = 0;
    }
    return static_cast<int64_t>(j.size());
}

}  // extern "C"
