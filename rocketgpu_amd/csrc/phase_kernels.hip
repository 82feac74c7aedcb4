// Single phases of the car half of the tick as gfx950 kernels, one lane per car, over byte-compatible records in
// device memory.  Test-only library (rocketgpu_amd/_build/librg_phases.so): the phases are checked here, against the
// reference's own phase functions, before the whole car tick that will call them exists.  BatchSim does not use it.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "batch_common.hpp"

using namespace rg;

namespace {

// Car slot k of env e (lane e * max_cars + k): its pre-tick (StepEnvPreTickCar); unsupported[i] = 1 marks a car
// it left alone.
__global__ void __launch_bounds__(64) CarPreTickKernel(CarState* cars, const CarControls* controls,
                                                       const CarConfig* configs, WheelRuntimeState* wheels,
                                                       const EnvState* envs, const uint8_t* teams, int32_t n_env,
                                                       int32_t max_cars, uint64_t seed, core::CarParams m,
                                                       core::CarSpawnRules rules, uint8_t* unsupported) {
    const int i = static_cast<int>(blockIdx.x) * static_cast<int>(blockDim.x) + static_cast<int>(threadIdx.x);
    if (i >= n_env * max_cars) return;
    const int e = i / max_cars;
    CarState car = cars[i];
    const CarConfig cfg = configs[i];
    if (!core::CarPreTickSupported(car, cfg, m)) {
        unsupported[i] = 1;
        return;
    }
    unsupported[i] = 0;
    WheelRuntimeState ws;
    if (wheels) ws = wheels[i];
    const core::CarRespawnContext rc{&rules, seed, e, envs[e].tick, static_cast<uint32_t>(i + 1), teams[i]};
    core::CarPreTick(car, controls[i], cfg, m, wheels ? &ws : nullptr, rc);
    cars[i] = car;
    if (wheels) wheels[i] = ws;
}

// Car k's post-tick (cooldowns, last controls, bumps and velocity caches, tick counter, reported fields).
__global__ void __launch_bounds__(64) CarPostTickKernel(CarState* cars, const CarControls* controls,
                                                        const WheelRuntimeState* wheels, int32_t n, core::CarParams m,
                                                        int32_t report) {
    const int k = static_cast<int>(blockIdx.x) * static_cast<int>(blockDim.x) + static_cast<int>(threadIdx.x);
    if (k >= n) return;
    CarState car = cars[k];
    WheelRuntimeState ws;
    if (wheels) ws = wheels[k];
    core::CarPostTick(car, controls[k], m, wheels ? &ws : nullptr, report != 0);
    cars[k] = car;
}

// Car k's force phase (0: jump hold, 1: boost, 2: air control, 3: flip torque); queues may be null.
__global__ void __launch_bounds__(64) CarForcePhaseKernel(CarState* cars, const CarControls* controls,
                                                          const CarConfig* configs, core::CarForceQueue* queues,
                                                          int32_t n, int32_t phase, int32_t flags, core::CarParams m) {
    const int k = static_cast<int>(blockIdx.x) * static_cast<int>(blockDim.x) + static_cast<int>(threadIdx.x);
    if (k >= n) return;
    CarState car = cars[k];
    core::CarForceQueue q;
    if (queues) q = queues[k];
    core::CarForceQueue* qp = queues ? &q : nullptr;
    if (phase == 0) {
        core::CarUpdateJumpState(car, controls[k], m, qp);
    } else if (phase == 1) {
        core::CarApplyBoost(car, controls[k], m, qp);
    } else {
        const CarConfig cfg = configs[k];
        const CarControls ctl = controls[k];
        const core::CarFrame frame = core::MakeCarFrame(car, cfg, m.car_mass);
        if (phase == 2) {
            core::CarApplyAirControl(car, ctl, cfg, m, frame, (flags & 1) != 0, (flags & 2) != 0, qp);
        } else {
            core::CarApplyFlipTorque(car, ctl, cfg, m, frame, (flags & 1) != 0 ? ctl.pitch : 2.0f, qp);
        }
    }
    cars[k] = car;
    if (queues) queues[k] = q;
}

// The step core's math functions on the device: fn 0 GameSinf, 1 GameCosf, 2 UcrtAtan2f(x, y), 3 UcrtPowf(x, y),
// 4 UcrtAcosf, 5 SafeAtan2(x, y).
__device__ inline float MathFn(int fn, float x, float y) {
    switch (fn) {
    case 0: return core::GameSinf(x);
    case 1: return core::GameCosf(x);
    case 2: return core::UcrtAtan2f(x, y);
    case 3: return core::UcrtPowf(x, y);
    case 4: return core::UcrtAcosf(x);
    default: return core::SafeAtan2(x, y);
    }
}

// splitmix64's output function: the mixing of the math checksums.
__device__ inline uint64_t Mix64(uint64_t z) {
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// Sum over inputs i in [first, first + count) (the float with bits i) of Mix64(i << 32 | result bits), NaN results
// counted as 0x7fc00000: an order-free checksum of a one-argument function over a range of bit patterns.
__global__ void __launch_bounds__(256) MathChecksumKernel(int fn, uint64_t first, uint64_t count, uint64_t* sum) {
    uint64_t acc = 0;
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t k = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; k < count; k += stride) {
        const uint64_t i = first + k;
        const float r = MathFn(fn, core::BitsFloat(static_cast<uint32_t>(i)), 0.0f);
        const uint32_t b = r != r ? 0x7fc00000u : core::FloatBits(r);
        acc += Mix64((i << 32) | b);
    }
    atomicAdd(reinterpret_cast<unsigned long long*>(sum), static_cast<unsigned long long>(acc));
}

__global__ void __launch_bounds__(256) MathEvalKernel(int fn, const float* x, const float* y, float* out, int64_t n) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) out[i] = MathFn(fn, x[i], y[i]);
}

}  // namespace

extern "C" {

// The checksum of MathChecksumKernel for fn 0, 1 or 4 over bit patterns [first, first + count), into *host_sum.
// Returns a hipError_t.
int rgp_math_checksum(int32_t fn, uint64_t first, uint64_t count, uint64_t* host_sum) {
    if ((fn != 0 && fn != 1 && fn != 4) || host_sum == nullptr || first + count > (1ull << 32)) {
        return static_cast<int>(hipErrorInvalidValue);
    }
    uint64_t* d = nullptr;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&d), sizeof(uint64_t));
    if (e != hipSuccess) return static_cast<int>(e);
    e = hipMemset(d, 0, sizeof(uint64_t));
    if (e == hipSuccess) {
        hipLaunchKernelGGL(MathChecksumKernel, dim3(65536), dim3(256), 0, nullptr, fn, first, count, d);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(host_sum, d, sizeof(uint64_t), hipMemcpyDeviceToHost);
    (void)hipFree(d);
    return static_cast<int>(e);
}

// out[i] = fn(x[i], y[i]) for i < n (device pointers).  Returns a hipError_t.
int rgp_math_eval(int32_t fn, const float* x, const float* y, float* out, int64_t n, void* stream) {
    if (fn < 0 || fn > 5) return static_cast<int>(hipErrorInvalidValue);
    if (n <= 0) return 0;
    hipLaunchKernelGGL(MathEvalKernel, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0,
                       static_cast<hipStream_t>(stream), fn, x, y, out, n);
    return static_cast<int>(hipGetLastError());
}

// One force phase (0: jump hold, 1: boost, 2: air control, 3: flip torque) on n cars in place (device pointers;
// configs: one CarConfig per car, read by phases 2 and 3; queues: one force queue per car, or null); flags: for air
// control bit 0 updates air control and bit 1 applies the air throttle, for the flip torque bit 0 passes the
// control's pitch as the raw pitch; params: a host list of the force-phase parameters.  Returns a
// hipError_t, or hipErrorInvalidValue for an unknown phase, missing configs or a null parameter list.
int rgp_car_force_phase(void* cars, const void* controls, const void* configs, void* queues, int32_t n,
                        int32_t phase, int32_t flags, const float* params, void* stream) {
    if (phase < 0 || phase > 3 || (phase >= 2 && configs == nullptr) || params == nullptr) {
        return static_cast<int>(hipErrorInvalidValue);
    }
    if (n <= 0) return 0;
    const core::CarParams m = ForcePhaseParams(params);
    hipLaunchKernelGGL(CarForcePhaseKernel, dim3((n + 63) / 64), dim3(64), 0, static_cast<hipStream_t>(stream),
                       static_cast<CarState*>(cars), static_cast<const CarControls*>(controls),
                       static_cast<const CarConfig*>(configs), static_cast<core::CarForceQueue*>(queues), n, phase,
                       flags, m);
    return static_cast<int>(hipGetLastError());
}

// n cars in place (device pointers; wheels may be null, and are only read).  Returns a hipError_t.
int rgp_car_posttick(void* cars, const void* controls, const void* wheels, int32_t n, int32_t report, void* stream) {
    if (n <= 0) return 0;
    const core::CarParams m{};
    hipLaunchKernelGGL(CarPostTickKernel, dim3((n + 63) / 64), dim3(64), 0, static_cast<hipStream_t>(stream),
                       static_cast<CarState*>(cars), static_cast<const CarControls*>(controls),
                       static_cast<const WheelRuntimeState*>(wheels), n, m, report);
    return static_cast<int>(hipGetLastError());
}

// n_env x max_cars car slots in place on PA(scale, fillet) (device pointers; wheels may be null), the default car
// parameters; unsupported: one device byte per slot.  Returns a hipError_t.
int rgp_car_pretick(void* cars, const void* controls, const void* configs, void* wheels, const void* envs,
                    const void* teams, int32_t n_env, int32_t max_cars, double scale, double fillet, uint64_t seed,
                    void* unsupported, void* stream) {
    const int32_t n = n_env * max_cars;
    if (n <= 0) return 0;
    const core::CarParams m{};
    rgw::ArenaParams ap;
    ap.scale = scale;
    ap.fillet_radius = fillet;
    const core::CarSpawnRules rules = SpawnRules(rgw::ProceduralArenaRules(ap));
    hipLaunchKernelGGL(CarPreTickKernel, dim3((n + 63) / 64), dim3(64), 0, static_cast<hipStream_t>(stream),
                       static_cast<CarState*>(cars), static_cast<const CarControls*>(controls),
                       static_cast<const CarConfig*>(configs), static_cast<WheelRuntimeState*>(wheels),
                       static_cast<const EnvState*>(envs), static_cast<const uint8_t*>(teams), n_env, max_cars, seed,
                       m, rules, static_cast<uint8_t*>(unsupported));
    return static_cast<int>(hipGetLastError());
}

}  // extern "C"
