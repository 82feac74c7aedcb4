// Host-compiled twin of the step core, for tests only.  It runs the exact same core sources as the gfx950 kernels,
// compiled for the host, so that the CPU test suite can compare the core against the oracle (the reference CPU
// build) without a GPU, and so a GPU mismatch can be localized to the compiler or the device path.  It is not part
// of the product: BatchSim never calls it and the Python package never loads it.
#include <cmath>
#include <cstring>
#include <memory>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "batch_common.hpp"
#include "core/rg_arena_step.hpp"
#include "core/rg_car_phases.hpp"
#include "core/rg_env_commands.hpp"

using namespace rg;

namespace {

struct Twin {
    ArenaMeshHost mesh;
    core::ArenaMeshView view{};
    core::StepParams params{};
    std::vector<BallState> balls;
    std::vector<EnvState> envs;
    std::vector<core::ArenaInternal> internal;
    std::vector<BoostPadState> pads;
    bool kickoff_settled = true;
    bool settled_ready = false;
    BallState settled{};
    // Physics event stream in host memory (null windows: off).
    PhysicsEvent* event_records = nullptr;
    PhysicsEventWindow* event_windows = nullptr;
    uint32_t event_capacity = 0;
    bool event_goal_only = false;
};

core::EventSink EnvEvents(const Twin& t, size_t e) {
    core::EventSink s;
    if (t.event_windows != nullptr) {
        s.window = t.event_windows + e;
        s.records = t.event_records + e * t.event_capacity;
        s.capacity = t.event_capacity;
        s.goal_only = t.event_goal_only;
    }
    return s;
}

BoostPadState* EnvPads(Twin& t, size_t e) {
    return t.pads.empty() ? nullptr : t.pads.data() + e * static_cast<size_t>(t.params.pad_count);
}

// The settled kickoff ball, by the same procedure as BatchSim::EnsureSettledBall, on the host core.
const BallState* Settled(Twin& t) {
    if (!t.kickoff_settled) return nullptr;
    if (!t.settled_ready) {
        BallState ball;
        EnvState env;
        std::memset(static_cast<void*>(&ball), 0, sizeof(ball));
        std::memset(static_cast<void*>(&env), 0, sizeof(env));
        new (&ball) BallState;
        new (&env) EnvState;
        auto st = std::make_unique<core::ArenaInternal>();
        std::memset(static_cast<void*>(st.get()), 0, sizeof(core::ArenaInternal));
        core::ResetToKickoffNoCars(ball, env, *st, nullptr, 0, 0, nullptr, t.params);
        EnvCommand c{};
        c.reset_to_kickoff = 1;
        c.reset_scoreboard = 1;
        c.kickoff_seed_mode = 3;
        c.kickoff_countdown_ticks = core::kKickoffCountdownGameTimers;
        core::ApplyEnvCommandNoCars(ball, env, *st, nullptr, 0, c, nullptr, t.params);
        auto scratch = std::make_unique<core::BallTickScratch>();
        int ticks = core::kKickoffCountdownTicks;
        for (int k = 0;; ++k) {
            for (int i = 0; i < ticks; ++i) core::StepBallArena(ball, env, *st, t.view, t.params, *scratch);
            if (env.countdown_phase == 0) break;
            if (k == 16) throw std::runtime_error("kickoff_settled: the countdown did not end");
            ticks = 1;
        }
        t.settled = ball;
        t.settled_ready = true;
    }
    return &t.settled;
}

// A tick with the event code compiled in (an off sink writes nothing).
void StepWithEvents(BallState& ball, EnvState& env, core::ArenaInternal& st, const Twin& t,
                    core::BallTickScratch& scratch, const core::EventSink& events) {
    core::StepBallArena<core::SerialWalk, core::SerialWalk, true>(ball, env, st, t.view, t.params, scratch, events);
}

thread_local std::string g_err;

}  // namespace

extern "C" {

const char* rgt_last_error() { return g_err.c_str(); }

void* rgt_create(int env_count, double scale, double fillet_radius) {
    try {
        auto t = std::make_unique<Twin>();
        rgw::ArenaParams ap;
        ap.scale = scale;
        ap.fillet_radius = fillet_radius;
        t->mesh = BuildProceduralArenaMesh(ap);
        t->view.tris = t->mesh.tris.data();
        t->view.nodes = t->mesh.nodes.data();
        t->view.meshes = t->mesh.meshes.data();
        t->view.mesh_count = static_cast<int32_t>(t->mesh.meshes.size());
        t->view.tri_count = static_cast<int32_t>(t->mesh.tris.size());
        t->params = DefaultStepParams(rgw::ProceduralArenaRules(ap));
        t->balls.assign(static_cast<size_t>(env_count), BallState{});
        t->envs.assign(static_cast<size_t>(env_count), EnvState{});
        t->internal.assign(static_cast<size_t>(env_count), core::ArenaInternal{});
        std::memset(t->internal.data(), 0, t->internal.size() * sizeof(core::ArenaInternal));
        t->pads.assign(static_cast<size_t>(env_count) * static_cast<size_t>(t->params.pad_count), BoostPadState{});
        for (size_t e = 0; e < static_cast<size_t>(env_count); ++e) core::ResetBoostPads(EnvPads(*t, e), t->params);
        return t.release();
    } catch (const std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}

void rgt_destroy(void* p) { delete static_cast<Twin*>(p); }

// Same table format as oracle/ref_shim.cpp rgref_mesh_tables.
int rgt_mesh_tables(void* p, long* counts, float* tris, int32_t* nodes, float* meshes, int32_t* mesh_ints) {
    const Twin& t = *static_cast<Twin*>(p);
    counts[0] = static_cast<long>(t.mesh.tris.size());
    counts[1] = static_cast<long>(t.mesh.nodes.size());
    counts[2] = static_cast<long>(t.mesh.meshes.size());
    if (tris) {
        for (size_t i = 0; i < t.mesh.tris.size(); ++i) {
            const core::MeshTriangle& m = t.mesh.tris[i];
            float* o = tris + 16 * i;
            const Vec3 v[4] = {m.a, m.b, m.c, m.normal};
            for (int k = 0; k < 4; ++k) {
                o[3 * k] = v[k].x;
                o[3 * k + 1] = v[k].y;
                o[3 * k + 2] = v[k].z;
            }
            o[12] = m.edge_angle_v0v1;
            o[13] = m.edge_angle_v1v2;
            o[14] = m.edge_angle_v2v0;
            o[15] = static_cast<float>(m.flags & 0xffu);
        }
    }
    if (nodes) {
        for (size_t i = 0; i < t.mesh.nodes.size(); ++i) {
            const core::MeshBvhNode& n = t.mesh.nodes[i];
            int32_t* o = nodes + 4 * i;
            for (int a = 0; a < 3; ++a) o[a] = n.qmin[a] | (n.qmax[a] << 16);
            o[3] = n.escape_or_tri;
        }
    }
    if (meshes) {
        for (size_t i = 0; i < t.mesh.meshes.size(); ++i) {
            const core::MeshInfo& m = t.mesh.meshes[i];
            float* o = meshes + 15 * i;
            const Vec3 v[5] = {m.aabb_min, m.aabb_max, m.q.bvh_min, m.q.bvh_max, m.q.scale};
            for (int k = 0; k < 5; ++k) {
                o[3 * k] = v[k].x;
                o[3 * k + 1] = v[k].y;
                o[3 * k + 2] = v[k].z;
            }
            int32_t* oi = mesh_ints + 4 * i;
            oi[0] = m.tri_start;
            oi[1] = m.tri_count;
            oi[2] = m.node_start;
            oi[3] = m.node_count;
        }
    }
    return 0;
}

// which: 1 BallState, 2 EnvState (public upload: balls are canonicalized like BatchSim::upload_ball_states).
int rgt_upload(void* p, int which, const void* src, long n) {
    Twin& t = *static_cast<Twin*>(p);
    if (which == 1) {
        std::memcpy(t.balls.data(), src, static_cast<size_t>(n) * sizeof(BallState));
        for (long i = 0; i < n; ++i) CanonicalizeBall(t.balls[static_cast<size_t>(i)]);
        return 0;
    }
    if (which == 2) {
        std::memcpy(t.envs.data(), src, static_cast<size_t>(n) * sizeof(EnvState));
        return 0;
    }
    if (which == 3) {
        std::memcpy(t.pads.data(), src, static_cast<size_t>(n) * sizeof(BoostPadState));
        return 0;
    }
    return -1;
}

int rgt_download(void* p, int which, void* dst, long n) {
    Twin& t = *static_cast<Twin*>(p);
    if (which == 1) {
        std::memcpy(dst, t.balls.data(), static_cast<size_t>(n) * sizeof(BallState));
        return 0;
    }
    if (which == 2) {
        std::memcpy(dst, t.envs.data(), static_cast<size_t>(n) * sizeof(EnvState));
        return 0;
    }
    if (which == 3) {
        std::memcpy(dst, t.pads.data(), static_cast<size_t>(n) * sizeof(BoostPadState));
        return 0;
    }
    return -1;
}

void rgt_set_kickoff_settled(void* p, int on) { static_cast<Twin*>(p)->kickoff_settled = on != 0; }

int rgt_settled_ball(void* p, void* out) {
    try {
        const BallState* b = Settled(*static_cast<Twin*>(p));
        if (!b) return -1;
        std::memcpy(out, b, sizeof(BallState));
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

int rgt_reset_to_kickoff(void* p, uint64_t seed) {
    Twin& t = *static_cast<Twin*>(p);
    const BallState* settled = Settled(t);
    for (size_t e = 0; e < t.balls.size(); ++e) {
        core::ResetToKickoffNoCars(t.balls[e], t.envs[e], t.internal[e], EnvPads(t, e), static_cast<int>(e), seed,
                                   settled, t.params);
    }
    return 0;
}

int rgt_apply_env_commands(void* p, const void* commands, long n) {
    Twin& t = *static_cast<Twin*>(p);
    const EnvCommand* c = static_cast<const EnvCommand*>(commands);
    const BallState* settled = Settled(t);
    for (long e = 0; e < n; ++e) {
        const size_t i = static_cast<size_t>(e);
        core::ApplyEnvCommandNoCars(t.balls[i], t.envs[i], t.internal[i], EnvPads(t, i), static_cast<int>(e), c[e],
                                    settled, t.params);
    }
    return 0;
}

// step_with_device_commands on the host: commands[tick * stride + env] before each env's tick.
int rgt_step_with_commands(void* p, const void* commands, int stride, int ticks) {
    Twin& t = *static_cast<Twin*>(p);
    const EnvCommand* c = static_cast<const EnvCommand*>(commands);
    const BallState* settled = Settled(t);
    auto scratch = std::make_unique<core::BallTickScratch>();
    for (int k = 0; k < ticks; ++k) {
        for (size_t e = 0; e < t.balls.size(); ++e) {
            core::ApplyTickCommandNoCars(t.balls[e], t.envs[e], t.internal[e], EnvPads(t, e), static_cast<int>(e),
                                         c[static_cast<size_t>(k) * static_cast<size_t>(stride) + e], settled,
                                         t.params);
            StepWithEvents(t.balls[e], t.envs[e], t.internal[e], t, *scratch, EnvEvents(t, e));
        }
    }
    return 0;
}

// The car pre-tick over a batch of n_env envs x max_cars slots on PA(scale, fillet) (car slot k of env e at
// e * max_cars + k; wheels may be null; envs give the tick, teams one byte per slot), with the respawn seed.  Returns
// the number of cars it does not step (three-wheel bodies); those are left as they were.
long rgt_car_pretick(int n_env, int max_cars, double scale, double fillet, uint64_t seed, void* cars,
                     const void* controls, const void* configs, void* wheels, const void* envs, const uint8_t* teams) {
    const core::CarParams m{};
    rgw::ArenaParams ap;
    ap.scale = scale;
    ap.fillet_radius = fillet;
    const core::CarSpawnRules rules = SpawnRules(rgw::ProceduralArenaRules(ap));
    CarState* c = static_cast<CarState*>(cars);
    const CarControls* ctl = static_cast<const CarControls*>(controls);
    const CarConfig* cc = static_cast<const CarConfig*>(configs);
    WheelRuntimeState* w = static_cast<WheelRuntimeState*>(wheels);
    const EnvState* ev = static_cast<const EnvState*>(envs);
    long skipped = 0;
    for (int e = 0; e < n_env; ++e) {
        for (int k = 0; k < max_cars; ++k) {
            const size_t i = static_cast<size_t>(e) * max_cars + k;
            if (!core::CarPreTickSupported(c[i], cc[i], m)) {
                ++skipped;
                continue;
            }
            const core::CarRespawnContext rc{&rules, seed, e, ev[e].tick, static_cast<uint32_t>(i + 1), teams[i]};
            core::CarPreTick(c[i], ctl[i], cc[i], m, w ? &w[i] : nullptr, rc);
        }
    }
    return skipped;
}

// The car post-tick on n independent cars (wheels may be null), reporting fields when report is nonzero.
void rgt_car_posttick(long n, void* cars, const void* controls, const void* wheels, int report) {
    const core::CarParams m{};
    CarState* c = static_cast<CarState*>(cars);
    const CarControls* ctl = static_cast<const CarControls*>(controls);
    const WheelRuntimeState* w = static_cast<const WheelRuntimeState*>(wheels);
    for (long k = 0; k < n; ++k) core::CarPostTick(c[k], ctl[k], m, w ? &w[k] : nullptr, report != 0);
}

// One force phase (0: jump hold, 1: boost, 2: air control, 3: flip torque) on n independent cars, with the
// parameters of a force-phase list; configs: one CarConfig per car (read by phases 2 and 3, may be null otherwise);
// flags: for air control bit 0 updates air control and bit 1 applies the air throttle, for the flip torque bit 0
// passes the control's pitch as the raw pitch (else none); queues: one force queue per car, or null for the form
// that changes the velocity at once.  Returns 0, or 1 for an unknown phase or missing configs.
int rgt_car_force_phase(long n, void* cars, const void* controls, const void* configs, void* queues, int phase,
                        int flags, const float* params) {
    if (phase < 0 || phase > 3 || (phase >= 2 && configs == nullptr)) return 1;
    const core::CarParams m = ForcePhaseParams(params);
    CarState* c = static_cast<CarState*>(cars);
    const CarControls* ctl = static_cast<const CarControls*>(controls);
    const CarConfig* cfg = static_cast<const CarConfig*>(configs);
    core::CarForceQueue* q = static_cast<core::CarForceQueue*>(queues);
    for (long k = 0; k < n; ++k) {
        core::CarForceQueue* qk = q ? &q[k] : nullptr;
        if (phase == 0) {
            core::CarUpdateJumpState(c[k], ctl[k], m, qk);
        } else if (phase == 1) {
            core::CarApplyBoost(c[k], ctl[k], m, qk);
        } else {
            const core::CarFrame frame = core::MakeCarFrame(c[k], cfg[k], m.car_mass);
            if (phase == 2) {
                core::CarApplyAirControl(c[k], ctl[k], cfg[k], m, frame, (flags & 1) != 0, (flags & 2) != 0, qk);
            } else {
                core::CarApplyFlipTorque(c[k], ctl[k], cfg[k], m, frame, (flags & 1) != 0 ? ctl[k].pitch : 2.0f, qk);
            }
        }
    }
    return 0;
}

// The game's axis quantization over n inputs.
void rgt_quantize_axis(const float* x, float* out, long n) {
    for (long i = 0; i < n; ++i) out[i] = core::QuantizeAxisInput(x[i]);
}

// The default car parameters in the oracle's rgref_car_params order.
int rgt_car_params(float* out) {
    const core::CarParams m{};
    const float v[] = {m.tick_dt, m.frame_frac, m.car_mass, m.car_max_speed, m.car_max_ang_speed,
                       m.supersonic_start_speed, m.supersonic_maintain_min_speed, m.supersonic_maintain_max_time,
                       m.jump_immediate_speed, m.jump_min_time, m.jump_reset_time_pad, m.jump_max_time,
                       m.doublejump_max_delay, m.flip_torque_time, m.flip_initial_vel_scale,
                       m.flip_forward_impulse_max_speed_scale, m.flip_side_impulse_max_speed_scale,
                       m.flip_backward_impulse_max_speed_scale, m.flip_backward_impulse_scale_x, m.boost_spawn_amount,
                       static_cast<float>(m.unlimited_flips), static_cast<float>(m.unlimited_double_jumps),
                       m.jump_accel, m.boost_max, m.boost_used_per_second, m.boost_min_time, m.boost_accel_ground,
                       m.recharge_boost_per_second, m.recharge_boost_delay,
                       static_cast<float>(m.recharge_boost_enabled), m.car_air_throttle_accel,
                       m.flip_pitchlock_extra_time, m.car_air_control_pitch, m.car_air_control_yaw,
                       m.car_air_control_roll, m.car_air_damp_pitch, m.car_air_damp_yaw, m.car_air_damp_roll,
                       m.flip_torque_x, m.flip_torque_y};
    std::memcpy(out, v, sizeof(v));
    return static_cast<int>(sizeof(v) / sizeof(float));
}

// The physics event stream in host buffers (capacity records per env, one window per env); null windows: off.
void rgt_set_event_stream(void* p, void* records, void* windows, uint32_t capacity, int goal_only) {
    Twin& t = *static_cast<Twin*>(p);
    t.event_records = static_cast<PhysicsEvent*>(records);
    t.event_windows = static_cast<PhysicsEventWindow*>(windows);
    t.event_capacity = capacity;
    t.event_goal_only = goal_only != 0;
}

int rgt_step(void* p, int ticks) {
    Twin& t = *static_cast<Twin*>(p);
    auto scratch = std::make_unique<core::BallTickScratch>();
    for (int k = 0; k < ticks; ++k) {
        for (size_t e = 0; e < t.balls.size(); ++e) {
            StepWithEvents(t.balls[e], t.envs[e], t.internal[e], t, *scratch, EnvEvents(t, e));
        }
    }
    return 0;
}

// Per-env body digests of the ball arenas (no cars), then their fold into `batch` (may be null).
int rgt_env_digests(void* p, uint64_t* out, uint64_t* batch) {
    const Twin& t = *static_cast<Twin*>(p);
    for (size_t e = 0; e < t.balls.size(); ++e) {
        const BallState& b = t.balls[e];
        const float rot[4] = {b.rot.w, b.rot.x, b.rot.y, b.rot.z};
        out[e] = rgw::BodyDigest(rgw::kFnvOffset, &b.pos.x, &b.vel.x, &b.ang_vel.x, rot);
    }
    if (batch) *batch = rgw::Fnv1a(rgw::kFnvOffset, out, t.balls.size() * sizeof(uint64_t));
    return 0;
}

// ball_path on the host: the same loop as the device kernel (BallPathEnv, env_step.hpp:7217-7259).
int rgt_ball_path(void* p, int samples, int ticks, float goal_y, float* out, int32_t* result) {
    Twin& t = *static_cast<Twin*>(p);
    auto scratch = std::make_unique<core::BallTickScratch>();
    for (size_t e = 0; e < t.balls.size(); ++e) {
        BallState& ball = t.balls[e];
        bool scored = std::fabs(ball.pos.y) >= goal_y;
        int goal_tick = 0;
        float* o = out + e * static_cast<size_t>(samples) * 6;
        for (int s = 0; s < samples; ++s) {
            for (int k = 0; k < ticks && !scored; ++k) {
                StepWithEvents(ball, t.envs[e], t.internal[e], t, *scratch, EnvEvents(t, e));
                if (std::fabs(ball.pos.y) >= goal_y) {
                    scored = true;
                    goal_tick = s * ticks + k + 1;
                }
            }
            const float v[6] = {ball.pos.x, ball.pos.y, ball.pos.z, ball.vel.x, ball.vel.y, ball.vel.z};
            std::memcpy(o + 6 * s, v, sizeof(v));
        }
        result[2 * e] = scored ? (ball.pos.y > 0.0f ? 1 : 2) : 0;
        result[2 * e + 1] = goal_tick;
    }
    return 0;
}

// Invariant of the mesh trees the stackless walk relies on: a parent's quantized box contains every node of its
// subtree.  Returns the number of violations.
long rgt_bvh_containment_violations(void* p) {
    const Twin& t = *static_cast<Twin*>(p);
    long bad = 0;
    for (const core::MeshInfo& info : t.mesh.meshes) {
        const core::MeshBvhNode* nodes = t.mesh.nodes.data() + info.node_start;
        for (int i = 0; i < info.node_count; ++i) {
            if (nodes[i].escape_or_tri >= 0) continue;
            const int end = i - nodes[i].escape_or_tri;
            for (int j = i + 1; j < end && j < info.node_count; ++j) {
                for (int a = 0; a < 3; ++a) {
                    if (nodes[j].qmin[a] < nodes[i].qmin[a] || nodes[j].qmax[a] > nodes[i].qmax[a]) ++bad;
                }
            }
        }
    }
    return bad;
}

int rgt_math(int fn, const float* x, const float* y, float* out, long n) {
    for (long i = 0; i < n; ++i) {
        switch (fn) {
        case 0: out[i] = core::GameSinf(x[i]); break;
        case 1: out[i] = core::GameCosf(x[i]); break;
        case 2: out[i] = core::UcrtAtan2f(x[i], y[i]); break;
        case 3: out[i] = core::UcrtPowf(x[i], y[i]); break;
        case 4: out[i] = core::UcrtAcosf(x[i]); break;
        case 5: out[i] = core::SafeAtan2(x[i], y[i]); break;
        default: return -1;
        }
    }
    return 0;
}


// ---- Car arenas (test-only twin of the car step) ----------------------------------------------------------------

namespace {
struct CarTwin {
    ArenaMeshHost mesh;
    core::ArenaMeshView view{};
    core::StepParams params{};
    core::CarParams m{};
    core::CarSpawnRules rules{};
    std::vector<Vec3> pad_pos;
    int env_count = 0, max_cars = 0;
    uint64_t seed = 0;
    std::vector<BallState> balls;
    std::vector<EnvState> envs;
    std::vector<core::ArenaInternal> internal;
    std::vector<BoostPadState> pads;
    std::vector<CarState> cars;
    std::vector<WheelRuntimeState> wheels;
    std::vector<CarConfig> cfgs;
    std::vector<CarControls> controls;
    std::vector<uint8_t> teams;
    std::vector<int32_t> counts;
};
}  // namespace

// A batch of env_count car arenas with max_cars slots each on PA(scale, fillet); `seed` is the respawn seed.  Every
// record starts as its default; the caller uploads the states (raw bytes, no canonicalization).
void* rgt_cars_create(int env_count, int max_cars, double scale, double fillet_radius, uint64_t seed) {
    try {
        auto t = std::make_unique<CarTwin>();
        rgw::ArenaParams ap;
        ap.scale = scale;
        ap.fillet_radius = fillet_radius;
        t->mesh = BuildProceduralArenaMesh(ap);
        t->view.tris = t->mesh.tris.data();
        t->view.nodes = t->mesh.nodes.data();
        t->view.meshes = t->mesh.meshes.data();
        t->view.mesh_count = static_cast<int32_t>(t->mesh.meshes.size());
        t->view.tri_count = static_cast<int32_t>(t->mesh.tris.size());
        const rgw::ArenaRulesSpec rules = rgw::ProceduralArenaRules(ap);
        t->params = DefaultStepParams(rules);
        t->rules = SpawnRules(rules);
        for (int k = 0; k < rules.pad_count; ++k) {
            t->pad_pos.push_back(Vec3(rules.pad_pos[k].x, rules.pad_pos[k].y, rules.pad_pos[k].z));
        }
        t->m.frame_frac = t->params.frame_frac;
        t->env_count = env_count;
        t->max_cars = max_cars;
        t->seed = seed;
        const size_t n = static_cast<size_t>(env_count), cap = n * static_cast<size_t>(max_cars);
        t->balls.assign(n, BallState{});
        t->envs.assign(n, EnvState{});
        t->internal.assign(n, core::ArenaInternal{});
        std::memset(static_cast<void*>(t->internal.data()), 0, n * sizeof(core::ArenaInternal));
        t->pads.assign(n * static_cast<size_t>(t->params.pad_count), BoostPadState{});
        t->cars.assign(cap, CarState{});
        t->wheels.assign(cap, WheelRuntimeState{});
        t->cfgs.assign(cap, CarConfig{});
        t->controls.assign(cap, CarControls{});
        t->teams.assign(cap, 0);
        t->counts.assign(n, 0);
        return t.release();
    } catch (const std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}

void rgt_cars_destroy(void* p) { delete static_cast<CarTwin*>(p); }

// which: 0 CarState, 1 BallState, 2 EnvState, 3 BoostPadState, 4 CarControls, 5 CarConfig, 6 env car counts (i32),
// 7 teams (u8), 8 WheelRuntimeState; 9 the unsupported words (u32 per env; download only).
static void* CarTwinArray(CarTwin& t, int which, size_t& item) {
    switch (which) {
    case 0: item = sizeof(CarState); return t.cars.data();
    case 1: item = sizeof(BallState); return t.balls.data();
    case 2: item = sizeof(EnvState); return t.envs.data();
    case 3: item = sizeof(BoostPadState); return t.pads.data();
    case 4: item = sizeof(CarControls); return t.controls.data();
    case 5: item = sizeof(CarConfig); return t.cfgs.data();
    case 6: item = sizeof(int32_t); return t.counts.data();
    case 7: item = 1; return t.teams.data();
    case 8: item = sizeof(WheelRuntimeState); return t.wheels.data();
    default: item = 0; return nullptr;
    }
}

int rgt_cars_upload(void* p, int which, const void* src, long n) {
    CarTwin& t = *static_cast<CarTwin*>(p);
    size_t item = 0;
    void* dst = CarTwinArray(t, which, item);
    if (dst == nullptr) return -1;
    std::memcpy(dst, src, static_cast<size_t>(n) * item);
    return 0;
}

int rgt_cars_download(void* p, int which, void* dst, long n) {
    CarTwin& t = *static_cast<CarTwin*>(p);
    if (which == 9) {
        uint32_t* o = static_cast<uint32_t*>(dst);
        for (long e = 0; e < n; ++e) o[e] = t.internal[static_cast<size_t>(e)].unsupported;
        return 0;
    }
    size_t item = 0;
    void* src = CarTwinArray(t, which, item);
    if (src == nullptr) return -1;
    std::memcpy(dst, src, static_cast<size_t>(n) * item);
    return 0;
}

// `ticks` ticks of every arena with the uploaded controls (BatchSim::step(ticks)).
int rgt_cars_step(void* p, int ticks) {
    CarTwin& t = *static_cast<CarTwin*>(p);
    auto scratch = std::make_unique<core::CarArenaTick>();
    const core::PadLayout pads{t.pad_pos.data(), t.params.pad_count};
    for (int k = 0; k < ticks; ++k) {
        for (int e = 0; e < t.env_count; ++e) {
            const size_t base = static_cast<size_t>(e) * t.max_cars;
            core::CarArenaRefs a;
            a.ball = &t.balls[e];
            a.env = &t.envs[e];
            a.st = &t.internal[e];
            a.pads = t.pads.data() + static_cast<size_t>(e) * t.params.pad_count;
            a.cars = t.cars.data() + base;
            a.wheels = t.wheels.data() + base;
            a.cfgs = t.cfgs.data() + base;
            a.teams = t.teams.data() + base;
            a.controls = t.controls.data() + base;
            a.count = t.counts[e];
            a.env_idx = e;
            a.car_base = static_cast<int>(base);
            core::CarRespawnContext rc{&t.rules, t.seed, e, 0u, 0u, 0};
            core::StepCarArena(a, t.view, t.params, t.m, pads, rc, k + 1 == ticks, *scratch);
        }
    }
    return 0;
}
}  // extern "C"
