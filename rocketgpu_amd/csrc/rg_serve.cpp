// Serving loop: continuous batching of step requests over arenas resident in HBM.
//
// Simulated actor clients each own a disjoint range of the batch's arenas (ball-only or with cars) and send
// requests "step my arenas k ticks" with Poisson arrivals; a request may bring fresh held controls for the client's
// cars, and every reset_every-th request of a client first resets its arenas to a kickoff.  The loop admits
// requests whose client is idle into the next launch while the launch's budgets allow (arenas per launch, and HBM
// bytes of the arena records the launch touches), keeps up to max_in_flight launches on separate streams (their
// arena sets are disjoint, so launches overlap on the device), and retires a request when the event after its
// launch completes.  Latency = completion - arrival (queueing included).  An arena's result depends only on its own
// requests (ticks, controls, resets), so a request's result is the same alone and in any mix; tests check that
// byte for byte, and replay served arenas on the reference build.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "rocketgpu_amd/rg_capi.h"
#include "rocketgpu_amd/rg_types.h"
#include "rocketgpu_amd/workload.h"

extern "C" void rg_set_last_error(const char* message);  // rg_batch.hip

namespace {

void HipCheck(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

void CapiCheck(int rc) {
    if (rc != RG_OK) throw std::runtime_error(rg_last_error());
}

struct Request {
    int32_t client;
    int32_t ticks;
    double arrival_s;
};

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    int32_t* host_ids = nullptr;
    int32_t* host_ticks = nullptr;
    int32_t* dev_ids = nullptr;
    int32_t* dev_ticks = nullptr;
    rgw::Controls* host_controls = nullptr;  // fresh controls of the launch's arenas, in arena order
    rg::EnvCommand* dev_commands = nullptr;  // env_count commands, zero except the arenas being reset
    rg::EnvCommand* host_reset = nullptr;    // one reset command per arena of the largest client
    std::vector<std::pair<int32_t, int32_t>> reset_ranges;
    std::vector<int32_t> requests;  // indices into the request table
    bool busy = false;
};

double Exponential(uint64_t& state, double rate) {
    const double u = rgw::UnitFloat(rgw::SplitMix64(state));
    return -std::log(1.0 - u) / rate;
}

void CheckConfig(const rg_serve_config* cfg) {
    if (!cfg) throw std::invalid_argument("null argument");
    if (cfg->clients <= 0 || cfg->min_arenas <= 0 || cfg->max_arenas < cfg->min_arenas) {
        throw std::invalid_argument("clients and arenas per client must be positive");
    }
    if (cfg->horizon_count <= 0 || cfg->horizon_count > 4) throw std::invalid_argument("1..4 horizons");
    for (int h = 0; h < cfg->horizon_count; ++h) {
        if (cfg->horizons[h] <= 0) throw std::invalid_argument("horizons must be positive");
    }
    if (!(cfg->request_rate_hz > 0.0) || !(cfg->duration_s > 0.0)) {
        throw std::invalid_argument("rate and duration must be positive");
    }
    if (cfg->max_in_flight <= 0 || cfg->max_in_flight > 8) throw std::invalid_argument("1..8 launches in flight");
    if (cfg->reset_every < 0 || cfg->launch_budget_bytes < 0) throw std::invalid_argument("negative budget or reset");
}

// Client c's arena range: counts uniform in [min_arenas, max_arenas], ranges back to back from arena 0.
void ClientArenas(const rg_serve_config* cfg, std::vector<int32_t>& first, std::vector<int32_t>& count) {
    uint64_t rng = cfg->seed ^ 0x5e7a11c0ffeeull;
    first.assign(cfg->clients, 0);
    count.assign(cfg->clients, 0);
    int64_t next = 0;
    for (int c = 0; c < cfg->clients; ++c) {
        const int span = cfg->max_arenas - cfg->min_arenas + 1;
        count[c] = cfg->min_arenas + static_cast<int32_t>(rgw::SplitMix64(rng) % static_cast<uint64_t>(span));
        first[c] = static_cast<int32_t>(next);
        next += count[c];
    }
}

// Client c's requests: Poisson arrivals at rate / clients over the run, each with a horizon drawn from the list.
template <typename F>
void ClientArrivals(const rg_serve_config* cfg, int32_t c, F&& add) {
    const double per_client = cfg->request_rate_hz / cfg->clients;
    uint64_t s = rgw::StreamSeed(cfg->seed, static_cast<uint64_t>(c), 0, 0x5e);
    double t = Exponential(s, per_client);
    while (t < cfg->duration_s) {
        const int h = static_cast<int>(rgw::SplitMix64(s) % static_cast<uint64_t>(cfg->horizon_count));
        add(t, cfg->horizons[h]);
        t += Exponential(s, per_client);
    }
}

}  // namespace

extern "C" int rg_serve_run(rg_batch* batch, const rg_serve_config* cfg, rg_serve_report* report,
                            int32_t* host_ticks_per_arena) {
    std::vector<Slot> slots;
    int rc = RG_OK;
    try {
        if (!batch || !report) throw std::invalid_argument("null argument");
        CheckConfig(cfg);
        const int64_t env_count = rg_batch_env_count(batch);
        const int32_t max_cars = env_count > 0 ? static_cast<int32_t>(rg_batch_car_capacity(batch) / env_count) : 0;

        std::vector<int32_t> first, count;
        ClientArenas(cfg, first, count);
        const int64_t next = static_cast<int64_t>(first.back()) + count.back();
        if (next > env_count) throw std::out_of_range("clients own more arenas than the batch holds");
        const int32_t budget = cfg->max_arenas_per_launch;
        if (budget < cfg->max_arenas) throw std::invalid_argument("max_arenas_per_launch below one client's arenas");

        // HBM bytes a request touches: the client's arena records at their car counts.
        std::vector<int32_t> cars(static_cast<size_t>(env_count));
        CapiCheck(rg_batch_download(batch, RG_ARRAY_ENV_CAR_COUNT, cars.data(), cars.size(), nullptr));
        std::vector<int64_t> client_bytes(cfg->clients, 0);
        int64_t resident_bytes = 0;
        for (int c = 0; c < cfg->clients; ++c) {
            for (int32_t k = 0; k < count[c]; ++k) client_bytes[c] += rg_batch_arena_bytes(batch, cars[first[c] + k]);
            resident_bytes += client_bytes[c];
            if (cfg->launch_budget_bytes > 0 && client_bytes[c] > cfg->launch_budget_bytes) {
                throw std::invalid_argument("launch_budget_bytes below one client's arena records");
            }
        }

        // Arrivals: one Poisson stream per client, merged by time; each client's requests numbered in order.
        std::vector<Request> reqs;
        for (int c = 0; c < cfg->clients; ++c) {
            ClientArrivals(cfg, c, [&](double t, int32_t ticks) { reqs.push_back(Request{c, ticks, t}); });
        }
        std::stable_sort(reqs.begin(), reqs.end(),
                         [](const Request& a, const Request& b) { return a.arrival_s < b.arrival_s; });
        std::vector<int64_t> served(cfg->clients, 0);  // requests of each client admitted so far

        // Device addresses of the controls (fresh controls are written there) and the reset commands.
        rg_device_view view{};
        CapiCheck(rg_batch_device_view(batch, &view));
        const bool resets = cfg->reset_every > 0;
        const bool controls = cfg->fresh_controls != 0 && max_cars > 0;
        if (resets) {
            // Record the settled kickoff rows now, outside the timed loop: an empty command list applies nothing.
            CapiCheck(rg_batch_apply_env_commands_device(batch, view.env_car_count, 0, nullptr));
        }

        slots.resize(static_cast<size_t>(cfg->max_in_flight));
        for (Slot& sl : slots) {
            HipCheck(hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking), "hipStreamCreate");
            HipCheck(hipEventCreateWithFlags(&sl.done, hipEventDisableTiming), "hipEventCreate");
            HipCheck(hipHostMalloc(reinterpret_cast<void**>(&sl.host_ids), budget * sizeof(int32_t)), "hipHostMalloc");
            HipCheck(hipHostMalloc(reinterpret_cast<void**>(&sl.host_ticks), budget * sizeof(int32_t)),
                     "hipHostMalloc");
            HipCheck(hipMalloc(reinterpret_cast<void**>(&sl.dev_ids), budget * sizeof(int32_t)), "hipMalloc");
            HipCheck(hipMalloc(reinterpret_cast<void**>(&sl.dev_ticks), budget * sizeof(int32_t)), "hipMalloc");
            if (controls) {
                HipCheck(hipHostMalloc(reinterpret_cast<void**>(&sl.host_controls),
                                       static_cast<size_t>(budget) * max_cars * sizeof(rgw::Controls)),
                         "hipHostMalloc");
            }
            if (resets) {
                HipCheck(hipMalloc(reinterpret_cast<void**>(&sl.dev_commands), env_count * sizeof(rg::EnvCommand)),
                         "hipMalloc");
                HipCheck(hipMemset(sl.dev_commands, 0, env_count * sizeof(rg::EnvCommand)), "hipMemset");
                HipCheck(hipHostMalloc(reinterpret_cast<void**>(&sl.host_reset),
                                       static_cast<size_t>(cfg->max_arenas) * sizeof(rg::EnvCommand)),
                         "hipHostMalloc");
                for (int32_t k = 0; k < cfg->max_arenas; ++k) {
                    rg::EnvCommand cmd{};
                    cmd.reset_to_kickoff = 1;
                    cmd.reset_scoreboard = 1;
                    cmd.kickoff_seed_mode = 3;
                    sl.host_reset[k] = cmd;
                }
            }
        }
        HipCheck(hipDeviceSynchronize(), "hipDeviceSynchronize");

        std::vector<std::deque<int32_t>> pending(static_cast<size_t>(cfg->clients));
        std::vector<uint8_t> busy(static_cast<size_t>(cfg->clients), 0);
        std::deque<int32_t> ready;  // clients with a pending request and no request in flight, oldest first
        std::vector<double> latency;
        latency.reserve(reqs.size());
        if (host_ticks_per_arena) std::memset(host_ticks_per_arena, 0, static_cast<size_t>(env_count) * 4);
        int64_t arena_ticks = 0, launches = 0;
        size_t arrived = 0, done = 0;

        using Clock = std::chrono::steady_clock;
        const auto t0 = Clock::now();
        const auto now_s = [&] { return std::chrono::duration<double>(Clock::now() - t0).count(); };

        while (done < reqs.size()) {
            bool progressed = false;
            const double now = now_s();
            for (; arrived < reqs.size() && reqs[arrived].arrival_s <= now; ++arrived) {
                const int32_t c = reqs[arrived].client;
                if (pending[c].empty() && !busy[c]) ready.push_back(c);
                pending[c].push_back(static_cast<int32_t>(arrived));
                progressed = true;
            }
            for (Slot& sl : slots) {
                if (!sl.busy) continue;
                const hipError_t q = hipEventQuery(sl.done);
                if (q == hipErrorNotReady) continue;
                HipCheck(q, "hipEventQuery");
                const double t_done = now_s();
                for (const int32_t r : sl.requests) {
                    latency.push_back(t_done - reqs[r].arrival_s);
                    const int32_t c = reqs[r].client;
                    busy[c] = 0;
                    if (!pending[c].empty()) ready.push_back(c);
                    ++done;
                }
                sl.requests.clear();
                sl.busy = false;
                progressed = true;
            }
            for (Slot& sl : slots) {
                if (sl.busy || ready.empty()) continue;
                int32_t n = 0;
                int64_t bytes = 0;
                std::deque<int32_t> deferred;
                sl.reset_ranges.clear();
                while (!ready.empty()) {
                    const int32_t c = ready.front();
                    ready.pop_front();
                    if (n + count[c] > budget ||
                        (cfg->launch_budget_bytes > 0 && bytes + client_bytes[c] > cfg->launch_budget_bytes)) {
                        deferred.push_back(c);
                        continue;
                    }
                    const int32_t r = pending[c].front();
                    pending[c].pop_front();
                    busy[c] = 1;
                    const int64_t number = served[c]++;
                    for (int32_t k = 0; k < count[c]; ++k) {
                        sl.host_ids[n + k] = first[c] + k;
                        sl.host_ticks[n + k] = reqs[r].ticks;
                        if (host_ticks_per_arena) host_ticks_per_arena[first[c] + k] += reqs[r].ticks;
                    }
                    if (controls) {
                        CapiCheck(rg_workload_request_controls(cfg->seed, number, first[c], count[c], max_cars,
                                                               cfg->p_boost, sl.host_controls +
                                                                   static_cast<size_t>(n) * max_cars));
                    }
                    if (resets && (number + 1) % cfg->reset_every == 0) sl.reset_ranges.push_back({n, c});
                    n += count[c];
                    bytes += client_bytes[c];
                    arena_ticks += static_cast<int64_t>(count[c]) * reqs[r].ticks;
                    sl.requests.push_back(r);
                }
                while (!deferred.empty()) {
                    ready.push_front(deferred.back());
                    deferred.pop_back();
                }
                if (n == 0) continue;
                HipCheck(hipMemcpyAsync(sl.dev_ids, sl.host_ids, n * sizeof(int32_t), hipMemcpyHostToDevice,
                                        sl.stream), "hipMemcpyAsync");
                HipCheck(hipMemcpyAsync(sl.dev_ticks, sl.host_ticks, n * sizeof(int32_t), hipMemcpyHostToDevice,
                                        sl.stream), "hipMemcpyAsync");
                // Resets first (the settled kickoff of the client's arenas), then the request's controls.
                if (!sl.reset_ranges.empty()) {
                    for (const auto& rc : sl.reset_ranges) {
                        const int32_t c = rc.second;
                        HipCheck(hipMemcpyAsync(sl.dev_commands + first[c], sl.host_reset,
                                                count[c] * sizeof(rg::EnvCommand), hipMemcpyHostToDevice, sl.stream),
                                 "hipMemcpyAsync");
                    }
                    CapiCheck(rg_batch_apply_env_commands_device(batch, sl.dev_commands, static_cast<size_t>(env_count),
                                                                 sl.stream));
                    for (const auto& rc : sl.reset_ranges) {
                        const int32_t c = rc.second;
                        HipCheck(hipMemsetAsync(sl.dev_commands + first[c], 0, count[c] * sizeof(rg::EnvCommand),
                                                sl.stream),
                                 "hipMemsetAsync");
                    }
                }
                if (controls) {
                    // The launch's arenas are whole client ranges: copy each client's rows to its arena slots.
                    int32_t at = 0;
                    for (const int32_t r : sl.requests) {
                        const int32_t c = reqs[r].client;
                        HipCheck(hipMemcpyAsync(static_cast<rgw::Controls*>(view.controls) +
                                                    static_cast<size_t>(first[c]) * max_cars,
                                                sl.host_controls + static_cast<size_t>(at) * max_cars,
                                                static_cast<size_t>(count[c]) * max_cars * sizeof(rgw::Controls),
                                                hipMemcpyHostToDevice, sl.stream),
                                 "hipMemcpyAsync");
                        at += count[c];
                    }
                }
                CapiCheck(rg_batch_step_arenas(batch, sl.dev_ids, sl.dev_ticks, n, 0, sl.stream));
                HipCheck(hipEventRecord(sl.done, sl.stream), "hipEventRecord");
                sl.busy = true;
                ++launches;
                progressed = true;
            }
            if (!progressed) {
                if (arrived < reqs.size() && ready.empty()) {
                    bool any_busy = false;
                    for (const Slot& sl : slots) any_busy = any_busy || sl.busy;
                    const double wait = reqs[arrived].arrival_s - now_s();
                    if (!any_busy && wait > 2e-4) {
                        std::this_thread::sleep_for(std::chrono::duration<double>(wait - 1e-4));
                        continue;
                    }
                }
                std::this_thread::yield();
            }
        }
        const double seconds = now_s();
        std::sort(latency.begin(), latency.end());
        const auto pct = [&](double p) {
            if (latency.empty()) return 0.0;
            const size_t i = std::min(latency.size() - 1, static_cast<size_t>(std::ceil(p * latency.size())) - 1);
            return latency[i] * 1e3;
        };
        double sum = 0.0;
        for (const double l : latency) sum += l;
        report->requests = static_cast<int64_t>(reqs.size());
        report->arena_ticks = arena_ticks;
        report->launches = launches;
        report->seconds = seconds;
        report->arena_ticks_per_s = seconds > 0.0 ? static_cast<double>(arena_ticks) / seconds : 0.0;
        report->p50_ms = pct(0.50);
        report->p99_ms = pct(0.99);
        report->mean_ms = latency.empty() ? 0.0 : sum / latency.size() * 1e3;
        report->max_ms = latency.empty() ? 0.0 : latency.back() * 1e3;
        report->arenas_resident = next;
        report->resident_bytes = resident_bytes;
    } catch (const std::invalid_argument& e) {
        rg_set_last_error(e.what());
        rc = RG_EINVAL;
    } catch (const std::out_of_range& e) {
        rg_set_last_error(e.what());
        rc = RG_ERANGE;
    } catch (const std::exception& e) {
        rg_set_last_error(e.what());
        rc = RG_ERUNTIME;
    }
    for (Slot& sl : slots) {
        if (sl.stream) (void)hipStreamSynchronize(sl.stream);
        if (sl.dev_ids) (void)hipFree(sl.dev_ids);
        if (sl.dev_ticks) (void)hipFree(sl.dev_ticks);
        if (sl.host_ids) (void)hipHostFree(sl.host_ids);
        if (sl.host_ticks) (void)hipHostFree(sl.host_ticks);
        if (sl.host_controls) (void)hipHostFree(sl.host_controls);
        if (sl.host_reset) (void)hipHostFree(sl.host_reset);
        if (sl.dev_commands) (void)hipFree(sl.dev_commands);
        if (sl.done) (void)hipEventDestroy(sl.done);
        if (sl.stream) (void)hipStreamDestroy(sl.stream);
    }
    return rc;
}

extern "C" int rg_serve_client_arenas(const rg_serve_config* cfg, int32_t client, int32_t* first, int32_t* count) {
    try {
        CheckConfig(cfg);
        if (client < 0 || client >= cfg->clients || !first || !count) throw std::invalid_argument("bad client");
        std::vector<int32_t> f, c;
        ClientArenas(cfg, f, c);
        *first = f[client];
        *count = c[client];
        return RG_OK;
    } catch (const std::exception& e) {
        rg_set_last_error(e.what());
        return RG_EINVAL;
    }
}

extern "C" int64_t rg_serve_client_requests(const rg_serve_config* cfg, int32_t client, int32_t* out_ticks,
                                            int64_t capacity) {
    try {
        CheckConfig(cfg);
        if (client < 0 || client >= cfg->clients) throw std::invalid_argument("bad client");
        int64_t n = 0;
        ClientArrivals(cfg, client, [&](double, int32_t ticks) {
            if (out_ticks && n < capacity) out_ticks[n] = ticks;
            ++n;
        });
        return n;
    } catch (const std::exception& e) {
        rg_set_last_error(e.what());
        return -1;
    }
}
