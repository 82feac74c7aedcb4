// rocketgpu_amd::BatchSim: the reference's C++ BatchSim interface (include/rocketgpu/rocketgpu.h:988-1089) over the
// C-ABI of rg_capi.h, header-only.  Method names, argument order, default arguments and exception types are the
// reference's; streams are hipStream_t (the reference's cudaStream_t).  Records are the byte-compatible ones of
// rg_types.h.  Errors from the library are thrown as the reference throws them: std::invalid_argument (RG_EINVAL),
// std::out_of_range (RG_ERANGE), std::runtime_error (RG_ERUNTIME).
//
//   rocketgpu_amd::SimConfig cfg;                      // rg_sim_config with the reference's defaults
//   cfg.env_count = 4096;
//   rocketgpu_amd::BatchSim sim(cfg, 1.0, 256.0);     // procedural arena PA(1, 256), or (cfg, layout)
//   sim.set_env_car_count(counts.data(), counts.size());
//   sim.reset_to_kickoff(seed);
//   sim.step_with_device_controls(d_controls, 8, sim.car_capacity(), stream);
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>

#include "rocketgpu_amd/rg_capi.h"
#include "rocketgpu_amd/rg_types.h"

namespace rocketgpu_amd {

using rg::BallState;
using rg::BoostPadState;
using rg::CarConfig;
using rg::CarControls;
using rg::CarState;
using rg::EnvCommand;
using rg::EnvState;
using rg::Team;
using rg::WheelRuntimeState;

// SimConfig (rocketgpu.h:842-856): the fields this build reads, with the reference's defaults.
struct SimConfig : rg_sim_config {
    SimConfig() : rg_sim_config{} {
        env_count = 1;
        max_cars_per_env = 6;
        tick_dt = 1.0f / 120.0f;
        enable_boost_pads = 1;
        enable_car_car_collisions = 1;
        enable_demos = 1;
        enable_scoring = 1;
        respawn_rng_seed = 0;
    }
};

// MutatorConfig (rocketgpu.h:708-840) with the reference's defaults.
struct MutatorConfig : rg_mutator_config {
    MutatorConfig() : rg_mutator_config{} { rg_mutator_config_defaults(this); }
};

enum class StateUploadMode { PublicState, PreserveInternal };

namespace detail {
inline void Check(int rc) {
    if (rc == RG_OK) return;
    const std::string msg = rg_last_error();
    if (rc == RG_EINVAL) throw std::invalid_argument(msg);
    if (rc == RG_ERANGE) throw std::out_of_range(msg);
    throw std::runtime_error(msg);
}
}  // namespace detail

class BatchSim {
public:
    // BatchSim(const SimConfig&, const ArenaLayout&)  (rocketgpu.h:992).
    BatchSim(const SimConfig& config, const rg_arena_layout& layout) {
        detail::Check(rg_batch_create(&config, &layout, &b_));
    }
    // The procedural arena PA(scale, fillet_radius) as the layout.
    BatchSim(const SimConfig& config, double scale, double fillet_radius) {
        detail::Check(rg_batch_create_procedural(&config, scale, fillet_radius, &b_));
    }
    ~BatchSim() {
        if (b_) rg_batch_destroy(b_);
    }
    BatchSim(BatchSim&& other) noexcept : b_(other.b_) { other.b_ = nullptr; }
    BatchSim& operator=(BatchSim&& other) noexcept {
        if (this != &other) {
            if (b_) rg_batch_destroy(b_);
            b_ = other.b_;
            other.b_ = nullptr;
        }
        return *this;
    }
    BatchSim(const BatchSim&) = delete;
    BatchSim& operator=(const BatchSim&) = delete;

    void set_mutator_config(const MutatorConfig& mutator) { detail::Check(rg_batch_set_mutator_config(b_, &mutator)); }
    MutatorConfig mutator_config() const {
        MutatorConfig m;
        detail::Check(rg_batch_get_mutator_config(b_, &m));
        return m;
    }

    size_t env_count() const noexcept { return static_cast<size_t>(rg_batch_env_count(b_)); }
    size_t car_capacity() const noexcept { return static_cast<size_t>(rg_batch_car_capacity(b_)); }
    size_t boost_pad_capacity() const noexcept { return static_cast<size_t>(rg_batch_boost_pad_capacity(b_)); }

    void set_env_car_count(const int32_t* host_counts, size_t count, hipStream_t stream = nullptr) {
        detail::Check(rg_batch_set_env_car_count(b_, host_counts, count, stream));
    }
    void set_teams(const Team* host_teams, size_t count, hipStream_t stream = nullptr) {
        detail::Check(rg_batch_set_teams(b_, reinterpret_cast<const uint8_t*>(host_teams), count, stream));
    }
    void set_car_configs(const CarConfig* host_configs, size_t count, hipStream_t stream = nullptr) {
        detail::Check(rg_batch_set_car_configs(b_, host_configs, count, stream));
    }

    void upload_car_states(const CarState* host_states, size_t count, hipStream_t stream = nullptr,
                           StateUploadMode mode = StateUploadMode::PublicState) {
        Upload(RG_ARRAY_CAR_STATES, host_states, count, stream, mode);
    }
    void upload_ball_states(const BallState* host_states, size_t count, hipStream_t stream = nullptr,
                            StateUploadMode mode = StateUploadMode::PublicState) {
        Upload(RG_ARRAY_BALL_STATES, host_states, count, stream, mode);
    }
    void upload_controls(const CarControls* host_controls, size_t count, hipStream_t stream = nullptr) {
        Upload(RG_ARRAY_CONTROLS, host_controls, count, stream, StateUploadMode::PublicState);
    }
    void upload_env_states(const EnvState* host_states, size_t count, hipStream_t stream = nullptr) {
        Upload(RG_ARRAY_ENV_STATES, host_states, count, stream, StateUploadMode::PublicState);
    }
    void upload_boost_pad_states(const BoostPadState* host_states, size_t count, hipStream_t stream = nullptr) {
        Upload(RG_ARRAY_BOOST_PAD_STATES, host_states, count, stream, StateUploadMode::PublicState);
    }
    void upload_wheel_runtime_states(const WheelRuntimeState* host_states, size_t count, hipStream_t stream = nullptr) {
        Upload(RG_ARRAY_WHEEL_STATES, host_states, count, stream, StateUploadMode::PreserveInternal);
    }

    void download_car_states(CarState* host_states, size_t count, hipStream_t stream = nullptr) const {
        detail::Check(rg_batch_download(b_, RG_ARRAY_CAR_STATES, host_states, count, stream));
    }
    void download_wheel_runtime_states(WheelRuntimeState* host_states, size_t count,
                                       hipStream_t stream = nullptr) const {
        detail::Check(rg_batch_download(b_, RG_ARRAY_WHEEL_STATES, host_states, count, stream));
    }
    void download_ball_states(BallState* host_states, size_t count, hipStream_t stream = nullptr) const {
        detail::Check(rg_batch_download(b_, RG_ARRAY_BALL_STATES, host_states, count, stream));
    }
    void download_env_states(EnvState* host_states, size_t count, hipStream_t stream = nullptr) const {
        detail::Check(rg_batch_download(b_, RG_ARRAY_ENV_STATES, host_states, count, stream));
    }
    void download_boost_pad_states(BoostPadState* host_states, size_t count, hipStream_t stream = nullptr) const {
        detail::Check(rg_batch_download(b_, RG_ARRAY_BOOST_PAD_STATES, host_states, count, stream));
    }
    void download_car_states_async(CarState* host_states, size_t count, hipStream_t stream = nullptr) const {
        detail::Check(rg_batch_download_async(b_, RG_ARRAY_CAR_STATES, host_states, count, stream));
    }
    void download_ball_states_async(BallState* host_states, size_t count, hipStream_t stream = nullptr) const {
        detail::Check(rg_batch_download_async(b_, RG_ARRAY_BALL_STATES, host_states, count, stream));
    }
    void download_env_states_async(EnvState* host_states, size_t count, hipStream_t stream = nullptr) const {
        detail::Check(rg_batch_download_async(b_, RG_ARRAY_ENV_STATES, host_states, count, stream));
    }
    void download_boost_pad_states_async(BoostPadState* host_states, size_t count,
                                         hipStream_t stream = nullptr) const {
        detail::Check(rg_batch_download_async(b_, RG_ARRAY_BOOST_PAD_STATES, host_states, count, stream));
    }

    void copy_state_from(const BatchSim& source, hipStream_t stream = nullptr) {
        detail::Check(rg_batch_copy_state_from(b_, source.b_, stream));
    }

    void reset_to_kickoff(uint64_t seed = 0, hipStream_t stream = nullptr) {
        detail::Check(rg_batch_reset_to_kickoff(b_, seed, stream));
    }
    void apply_env_commands(const EnvCommand* host_commands, size_t count, hipStream_t stream = nullptr) {
        detail::Check(rg_batch_apply_env_commands(b_, host_commands, count, stream));
    }
    void apply_env_commands_device(const EnvCommand* device_commands, size_t count, hipStream_t stream = nullptr) {
        detail::Check(rg_batch_apply_env_commands_device(b_, device_commands, count, stream));
    }
    void step(int32_t ticks = 1, hipStream_t stream = nullptr) { detail::Check(rg_batch_step(b_, ticks, stream)); }
    void ball_path(int32_t samples, int32_t ticks, float goal_y, float* out, int32_t* result,
                   hipStream_t stream = nullptr) {
        detail::Check(rg_batch_ball_path(b_, samples, ticks, goal_y, out, result, stream));
    }
    void step_with_device_controls(const CarControls* device_controls_by_tick, int32_t ticks, int32_t control_stride,
                                   hipStream_t stream = nullptr) {
        detail::Check(rg_batch_step_with_device_controls(b_, device_controls_by_tick, ticks, control_stride, stream));
    }
    void step_with_device_commands(const CarControls* device_controls_by_tick, int32_t control_stride,
                                   const EnvCommand* device_env_commands_by_tick, int32_t env_command_stride,
                                   int32_t ticks, hipStream_t stream = nullptr) {
        detail::Check(rg_batch_step_with_device_commands(b_, device_controls_by_tick, control_stride,
                                                         device_env_commands_by_tick, env_command_stride, ticks,
                                                         stream));
    }
    void synchronize(hipStream_t stream = nullptr) const { detail::Check(rg_batch_synchronize(b_, stream)); }

    // The C handle, for the entry points with no reference counterpart (launch shape, snapshots, serving).
    rg_batch* handle() const noexcept { return b_; }

private:
    void Upload(int id, const void* src, size_t count, hipStream_t stream, StateUploadMode mode) {
        detail::Check(rg_batch_upload_mode(b_, id, src, count,
                                           mode == StateUploadMode::PreserveInternal ? RG_UPLOAD_PRESERVE_INTERNAL
                                                                                     : RG_UPLOAD_PUBLIC_STATE,
                                           stream));
    }

    rg_batch* b_ = nullptr;
};

}  // namespace rocketgpu_amd
