// Single phases of the car half of the tick as gfx950 kernels, one lane per car, over byte-compatible records in
// device memory.  Test-only library (rocketgpu_amd/_build/librg_phases.so): the phases are checked here, against the
// reference's own phase functions, before the whole car tick that will call them exists.  BatchSim does not use it.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "core/rg_car_phases.hpp"

using namespace rg;

namespace {

// Car k's pre-tick (StepEnvPreTickCar without respawns); unsupported[k] = 1 marks a car it left alone.
__global__ void __launch_bounds__(64) CarPreTickKernel(CarState* cars, const CarControls* controls,
                                                       const CarConfig* configs, WheelRuntimeState* wheels, int32_t n,
                                                       core::CarParams m, uint8_t* unsupported) {
    const int k = static_cast<int>(blockIdx.x) * static_cast<int>(blockDim.x) + static_cast<int>(threadIdx.x);
    if (k >= n) return;
    CarState car = cars[k];
    const CarConfig cfg = configs[k];
    if (!core::CarPreTickSupported(car, cfg, m)) {
        unsupported[k] = 1;
        return;
    }
    unsupported[k] = 0;
    WheelRuntimeState ws;
    if (wheels) ws = wheels[k];
    core::CarPreTick(car, controls[k], cfg, m, wheels ? &ws : nullptr);
    cars[k] = car;
    if (wheels) wheels[k] = ws;
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

}  // namespace

extern "C" {

// n cars in place (device pointers; wheels may be null, and are only read).  Returns a hipError_t.
int rgp_car_posttick(void* cars, const void* controls, const void* wheels, int32_t n, int32_t report, void* stream) {
    if (n <= 0) return 0;
    const core::CarParams m{};
    hipLaunchKernelGGL(CarPostTickKernel, dim3((n + 63) / 64), dim3(64), 0, static_cast<hipStream_t>(stream),
                       static_cast<CarState*>(cars), static_cast<const CarControls*>(controls),
                       static_cast<const WheelRuntimeState*>(wheels), n, m, report);
    return static_cast<int>(hipGetLastError());
}

// n cars in place (device pointers; wheels may be null), the default car parameters; unsupported: n device bytes.
// Returns a hipError_t.
int rgp_car_pretick(void* cars, const void* controls, const void* configs, void* wheels, int32_t n,
                    void* unsupported, void* stream) {
    if (n <= 0) return 0;
    const core::CarParams m{};
    hipLaunchKernelGGL(CarPreTickKernel, dim3((n + 63) / 64), dim3(64), 0, static_cast<hipStream_t>(stream),
                       static_cast<CarState*>(cars), static_cast<const CarControls*>(controls),
                       static_cast<const CarConfig*>(configs), static_cast<WheelRuntimeState*>(wheels), n, m,
                       static_cast<uint8_t*>(unsupported));
    return static_cast<int>(hipGetLastError());
}

}  // extern "C"
