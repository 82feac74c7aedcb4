"""Writes the golden fixtures of the parity tests by running the oracle (the reference CPU build, oracle/_ref).

Runs only where the reference sources were built (this container): `python tests/golden/make_fixtures.py`.
Each fixture is an .npz of data only:

  init_ball [N, 288] u8, init_env [N, 44] u8   the uploaded states (public-state upload)
  batch_digest [T] u64                         per tick: FNV-1a over the per-env body digests (env order)
  checkpoint_every, env_digest [K, N] u64      per-env body digests every checkpoint_every ticks
  final_ball [N, 288] u8, final_env [N, 44] u8 every byte after the last tick
  meta (json): workload, sizes, seed

Body digest: FNV-1a 64 over the raw bits of pos, vel, ang_vel, rot of the ball, then each car in slot order
(rocketgpu_amd/workload.h BodyDigest).
"""
from __future__ import annotations

import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)

from oracle import refsim as R  # noqa: E402

FIXTURES = {
    # name: (arenas, ticks, checkpoint_every, seed, scale, fillet)
    "ball_arena_96_t1000": (96, 1000, 100, 7, 1.0, 256.0),
    "ball_arena_1024_t3000": (1024, 3000, 500, 11, 1.0, 256.0),
    "ball_arena_small_scale_64_t1500": (64, 1500, 500, 13, 0.8, 192.0),
    "ball_arena_256_t10000": (256, 10000, 2000, 17, 1.0, 256.0),
}


def make(name, n, ticks, every, seed, scale, fillet):
    s = R.RefSim(n, 2, scale, fillet)
    s.set_env_car_count(np.zeros(n, np.int32))
    init_env = np.tile(R.default_state(R.ENV), (n, 1))
    init_ball = R.workload_ball_states(seed, 0, n, scale)
    s.upload(R.ENV, init_env)
    s.upload(R.BALL, init_ball)
    t0 = time.time()
    batch, env = s.run(np.zeros(n, np.int32), 0, ticks, every)
    dt = time.time() - t0
    final_ball = s.download(R.BALL, n)
    final_env = s.download(R.ENV, n)
    meta = {"workload": "ball arenas (seeded ball launches, no cars)", "arenas": n, "ticks": ticks, "seed": seed,
            "scale": scale, "fillet_radius": fillet, "generator": "oracle/_ref reference CPU build",
            "seconds": round(dt, 2)}
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, init_ball=init_ball, init_env=init_env, batch_digest=batch,
                        checkpoint_every=np.int64(every), env_digest=env, final_ball=final_ball,
                        final_env=final_env, meta=np.frombuffer(json.dumps(meta).encode(), np.uint8))
    print(f"{name}: {n} arenas x {ticks} ticks in {dt:.1f}s -> {os.path.getsize(path)} bytes")


BALL_PATH_FIXTURES = {
    # name: (arenas, samples, ticks per sample, goal_y, seed, scale, fillet)
    "ball_path_256_s40_t6": (256, 40, 6, 5124.2373046875, 21, 1.0, 256.0),
    "ball_path_small_scale_64_s25_t9": (64, 25, 9, 4000.0, 23, 0.8, 192.0),
}


def make_ball_path(name, n, samples, ticks, goal_y, seed, scale, fillet):
    """BatchSim::ball_path on seeded ball launches: out [n, samples, 6] f32, result [n, 2] i32, final states."""
    s = R.RefSim(n, 2, scale, fillet)
    s.set_env_car_count(np.zeros(n, np.int32))
    init_env = np.tile(R.default_state(R.ENV), (n, 1))
    init_ball = R.workload_ball_states(seed, 0, n, scale)
    s.upload(R.ENV, init_env)
    s.upload(R.BALL, init_ball)
    out, res = s.ball_path(samples, ticks, goal_y)
    meta = {"workload": "ball_path over seeded ball launches", "arenas": n, "samples": samples, "ticks": ticks,
            "goal_y": goal_y, "seed": seed, "scale": scale, "fillet_radius": fillet,
            "generator": "oracle/_ref reference CPU build"}
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, init_ball=init_ball, init_env=init_env, out=out, result=res,
                        final_ball=s.download(R.BALL, n), final_env=s.download(R.ENV, n),
                        meta=np.frombuffer(json.dumps(meta).encode(), np.uint8))
    print(f"{name}: {int((res[:, 0] > 0).sum())} of {n} reach a goal -> {os.path.getsize(path)} bytes")


# EnvCommand (rocketgpu.h:683-707), 16 bytes.
COMMAND = np.dtype([("reset_to_kickoff", "u1"), ("reset_scoreboard", "u1"), ("reset_tick_count", "u1"),
                    ("kickoff_seed_mode", "u1"), ("new_round_bodies", "u1"), ("keep_boost_pads", "u1"),
                    ("kickoff_countdown_ticks", "<u2"), ("kickoff_seed", "<u8")])

# One command of each kind the env-command path distinguishes.
COMMAND_VARIANTS = [
    {}, {"reset_to_kickoff": 1}, {"reset_to_kickoff": 1, "reset_scoreboard": 1},
    {"reset_to_kickoff": 1, "kickoff_seed": 12345}, {"reset_to_kickoff": 1, "kickoff_seed_mode": 1, "kickoff_seed": 99},
    {"reset_to_kickoff": 1, "kickoff_seed_mode": 2, "kickoff_seed": 5}, {"reset_to_kickoff": 1, "kickoff_seed_mode": 3},
    {"reset_to_kickoff": 1, "keep_boost_pads": 1}, {"reset_to_kickoff": 1, "kickoff_countdown_ticks": 0xFFFF},
    {"reset_to_kickoff": 1, "kickoff_countdown_ticks": 30, "reset_tick_count": 1}, {"new_round_bodies": 1},
    {"reset_scoreboard": 1, "reset_tick_count": 1},
]

ENV_COMMAND_FIXTURES = {
    # name: (arenas, seed, scale, fillet, stream ticks)
    "env_commands_48": (48, 77, 1.0, 256.0, 240),
}

# The phases of an env-command fixture, in order; states are recorded after each.
ENV_COMMAND_PHASES = ["step 300", "apply_env_commands", "step 700", "apply_env_commands_device", "step 100",
                      "step_with_device_commands", "reset_to_kickoff 777", "step 50"]


def env_command_inputs(n, seed, stream_ticks, pads0):
    """Seeded inputs of an env-command fixture: pads with random cooldowns, one command per env (cycling the
    variants), a sparse per-tick command stream and a per-tick controls stream (2 car slots per env)."""
    rng = np.random.default_rng(seed)
    pads = pads0.copy()
    pads[:, 4] = rng.integers(0, 2, pads.shape[0])
    pads[:, 0:4] = np.frombuffer(rng.uniform(0, 5, pads.shape[0]).astype(np.float32).tobytes(), np.uint8).reshape(-1, 4)
    cmds = np.zeros(n, COMMAND)
    for i in range(n):
        for k, v in COMMAND_VARIANTS[i % len(COMMAND_VARIANTS)].items():
            cmds[i][k] = v
    stream = np.zeros(stream_ticks * n, COMMAND)
    for t in range(stream_ticks):
        for e in rng.choice(n, 2, replace=False):
            for k, v in COMMAND_VARIANTS[rng.integers(len(COMMAND_VARIANTS))].items():
                stream[t * n + e][k] = v
            stream[t * n + e]["kickoff_seed"] = int(rng.integers(0, 1 << 40)) if rng.random() < 0.5 else 0
    # Only the last tick's controls survive a command-stream step (they become the batch's controls).
    controls = np.zeros((stream_ticks * n * 2, 24), np.uint8)
    controls[-n * 2:] = rng.integers(0, 256, (n * 2, 24), dtype=np.uint8)
    return pads, cmds.view(np.uint8).reshape(n, 16), stream.view(np.uint8).reshape(-1, 16), controls


def env_command_run(n, seed, scale, fillet, stream_ticks):
    """The env-command path on ball arenas in the oracle: apply_env_commands (host and device forms),
    step_with_device_commands and reset_to_kickoff between stepping phases; every byte of the balls, envs and pads
    after each phase, and the states a fresh batch starts with."""
    s = R.RefSim(n, 2, scale, fillet)
    s.set_env_car_count(np.zeros(n, np.int32))
    fresh = {"ball": s.download(R.BALL, n), "env": s.download(R.ENV, n), "pads": s.download(R.PAD, n * 34)}
    init_env = np.tile(R.default_state(R.ENV), (n, 1))
    init_ball = R.workload_ball_states(seed, 0, n, scale)
    init_pads, cmds, stream, controls = env_command_inputs(n, seed, stream_ticks, fresh["pads"])
    s.upload(R.ENV, init_env)
    s.upload(R.BALL, init_ball)
    s.upload(R.PAD, init_pads)
    balls, envs, pads = [], [], []
    for phase in ENV_COMMAND_PHASES:
        if phase.startswith("step ") and phase[5:].isdigit():
            s.step(int(phase[5:]))
        elif phase == "apply_env_commands":
            s.apply_env_commands(cmds)
        elif phase == "apply_env_commands_device":
            s.apply_env_commands_device(cmds)
        elif phase == "step_with_device_commands":
            s.step_with_commands(controls, 2 * n, stream, n, stream_ticks)
        else:
            s.reset_to_kickoff(int(phase.split()[-1]))
        balls.append(s.download(R.BALL, n))
        envs.append(s.download(R.ENV, n))
        pads.append(s.download(R.PAD, n * 34))
    return dict(init_ball=init_ball, init_env=init_env, init_pads=init_pads, commands=cmds, command_stream=stream,
                control_stream=controls, fresh_ball=fresh["ball"], fresh_env=fresh["env"], fresh_pads=fresh["pads"],
                ball=np.stack(balls), env=np.stack(envs), pads=np.stack(pads))


def make_env_commands(name, n, seed, scale, fillet, stream_ticks):
    run = env_command_run(n, seed, scale, fillet, stream_ticks)
    meta = {"workload": "env commands on ball arenas (seeded ball launches, no cars)", "arenas": n, "seed": seed,
            "scale": scale, "fillet_radius": fillet, "stream_ticks": stream_ticks, "phases": ENV_COMMAND_PHASES,
            "generator": "oracle/_ref reference CPU build"}
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, **run, meta=np.frombuffer(json.dumps(meta).encode(), np.uint8))
    print(f"{name}: {len(ENV_COMMAND_PHASES)} phases -> {os.path.getsize(path)} bytes")


CAR_PRETICK_FIXTURES = {
    # name: (arenas, cars per arena, ticks, seed, batches kept)
    "car_pretick_3v3": (8, 6, 240, 21, 8),
}


def demo_some_cars(cars, seed, lay):
    """Demoed cars for the respawn branch: a seeded ~20 % of the slots, each with a respawn timer that ends this
    tick, next tick or later, and a respawn override of none, a respawn spot, a kickoff spot or an offset spot."""
    rng = np.random.default_rng(seed)
    c = cars.copy()
    pick = rng.random(len(c)) < 0.2
    for i in np.flatnonzero(pick):
        c[i, lay["is_demoed"]] = 1
        timer = np.float32(rng.choice([0.004, 0.012, 0.5]))
        c[i, lay["demo_respawn_timer"]:lay["demo_respawn_timer"] + 4] = np.frombuffer(timer.tobytes(), np.uint8)
        c[i, lay["respawn_slot_override_plus1"]] = rng.choice([0, 0, 1, 3, 4, 5, 7, 9, 128, 140, 159])
    return c


def make_car_pretick(name, arenas, cars_per_env, ticks, seed, keep):
    """Batches of pre-tick inputs from 3v3 car arenas stepped by the reference, with a seeded fifth of the cars
    demoed (respawns on and after the tick), and the reference's pre-tick of each batch.  The `keep` ticks kept are
    those whose pre-tick starts the most jumps, double jumps and dodges."""
    import rocketgpu_amd as rg
    lay = {f[0]: f[2] for f in rg.state_layouts()["car_states"]["fields"]}
    run = R.car_arena_run(arenas, cars_per_env, ticks, seed)
    score = []
    for t in range(ticks):
        c, _ = R.car_pretick(arenas, cars_per_env, seed, run["cars"][t], run["controls"][t], run["configs"],
                             run["wheels"][t], run["envs"][t], run["teams"])
        score.append(int(((c[:, lay["cal_pre_flags"]] & 7) != 0).sum()))
    snaps = sorted(int(t) for t in np.argsort(-np.array(score), kind="stable")[:keep])
    cars = np.stack([demo_some_cars(run["cars"][t], seed + t, lay) for t in snaps])
    wheels, controls, envs = run["wheels"][snaps], run["controls"][snaps], run["envs"][snaps]
    out_cars, out_wheels = [], []
    for k in range(len(snaps)):
        c, w = R.car_pretick(arenas, cars_per_env, seed, cars[k], controls[k], run["configs"], wheels[k], envs[k],
                             run["teams"])
        out_cars.append(c)
        out_wheels.append(w)
    meta = {"workload": "car pre-tick batches from 3v3 car arenas with random held controls, a fifth of the cars "
                        "demoed", "arenas": arenas, "cars_per_env": cars_per_env, "ticks": ticks, "seed": seed,
            "snapshots": snaps, "generator": "oracle/_ref reference CPU build (StepEnvPreTickCar)"}
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, cars=cars, wheels=wheels, controls=controls, envs=envs, configs=run["configs"],
                        teams=run["teams"], out_cars=np.stack(out_cars), out_wheels=np.stack(out_wheels),
                        meta=np.frombuffer(json.dumps(meta).encode(), np.uint8))
    print(f"{name}: {len(snaps)} batches of {arenas * cars_per_env} cars -> {os.path.getsize(path)} bytes")


CAR_POSTTICK_FIXTURES = {
    # name: (arenas, cars per arena, ticks, seed, rows kept)
    "car_posttick_1v1_3v3": ((8, 4), (2, 6), 200, 31, 256),
}


def perturb_post_inputs(cars, seed):
    """Post-tick inputs that exercise every branch: on a seeded third of the rows pending bumps (1-3 velocity
    deltas in bt), on another third a velocity impulse cache (uu, or bt and uu), and a few demoed cars."""
    import rocketgpu_amd as rg
    lay = {f[0]: f[2] for f in rg.state_layouts()["car_states"]["fields"]}
    rng = np.random.default_rng(seed)
    c = cars.copy()
    n = len(c)
    kind = rng.integers(0, 3, n)
    for i in np.flatnonzero(kind == 1):
        k = int(rng.integers(1, 4))
        c[i, lay["pending_bump_count"]] = k
        bumps = rng.uniform(-3, 3, (7, 3)).astype(np.float32)
        c[i, lay["pending_bump_bt"]:lay["pending_bump_bt"] + 84] = bumps.view(np.uint8).reshape(-1)
    for i in np.flatnonzero(kind == 2):
        uu = rng.uniform(-200, 200, 3).astype(np.float32)
        c[i, lay["vel_impulse_cache"]:lay["vel_impulse_cache"] + 12] = uu.view(np.uint8)
        if rng.random() < 0.5:
            bt = (uu * np.float32(0.02)).astype(np.float32)
            c[i, lay["vel_impulse_cache_bt"]:lay["vel_impulse_cache_bt"] + 12] = bt.view(np.uint8)
    demo = rng.random(n) < 0.05
    c[demo, lay["is_demoed"]] = 1
    return c


def make_car_posttick(name, arenas, cars_per_env, ticks, seed, keep):
    """Post-tick inputs (the pre-tick's outputs of car-arena states, perturbed) and the reference's post-tick of
    each, with and without reported fields."""
    cars, wheels, controls = [], [], []
    for i, (n, c) in enumerate(zip(arenas, cars_per_env)):
        run = R.car_arena_run(n, c, ticks, seed + i)
        for t in range(ticks):
            mc, mw = R.car_pretick(n, c, seed, run["cars"][t], run["controls"][t], run["configs"], run["wheels"][t],
                                   run["envs"][t], run["teams"])
            cars.append(mc)
            wheels.append(mw)
            controls.append(run["controls"][t])
    mid, mid_wheels, controls = np.concatenate(cars), np.concatenate(wheels), np.concatenate(controls)
    rng = np.random.default_rng(seed)
    pick = np.sort(rng.choice(len(mid), keep, replace=False))
    inputs = perturb_post_inputs(mid[pick], seed)
    out_report = R.car_posttick(inputs, controls[pick], mid_wheels[pick], True)
    out_advance = R.car_posttick(inputs, controls[pick], mid_wheels[pick], False)
    meta = {"workload": "car post-tick inputs: pre-ticked car-arena states with seeded pending bumps, velocity "
                        "caches and demoed cars", "arenas": list(arenas), "cars_per_env": list(cars_per_env),
            "ticks": ticks, "seed": seed, "rows": int(keep),
            "generator": "oracle/_ref reference CPU build (StepEnvPostCar)"}
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, cars=inputs, wheels=mid_wheels[pick], controls=controls[pick], out_report=out_report,
                        out_advance=out_advance, meta=np.frombuffer(json.dumps(meta).encode(), np.uint8))
    print(f"{name}: {keep} rows -> {os.path.getsize(path)} bytes")


CAR_FORCE_FIXTURES = {
    # name: (arenas, cars per arena, ticks, seed, rows kept)
    "car_force_phases": (6, 4, 160, 41, 192),
}

# (phase, queued, parameter set, flags) of each output: phase 0 the jump hold, 1 boost, 2 air control (flags bit 0
# updates air control, bit 1 applies the air throttle), 3 the flip torque (flags bit 0: the control's pitch is the
# raw pitch); parameter set 0 the defaults, 1 the defaults with boost recharge on.
FORCE_CASES = ((0, 1, 0, 3), (0, 0, 0, 3), (1, 1, 0, 3), (1, 0, 0, 3), (1, 1, 1, 3), (1, 0, 1, 3),
               (2, 1, 0, 3), (2, 0, 0, 3), (2, 1, 0, 1), (2, 1, 0, 2), (3, 1, 0, 1), (3, 0, 0, 1), (3, 1, 0, 0))


def force_param_sets():
    return np.stack([R.force_params(), R.force_params(recharge_boost_enabled=1.0)])


def perturb_force_inputs(cars, controls, seed):
    """Force-phase inputs that reach every branch: boost amounts at and past the limits (0, tiny, spawn, full,
    over full, negative) with the kept fraction agreeing or not, boost timers on both sides of the minimum time and
    the recharge delay, jump timers on both sides of the reset pad, minimum and maximum times, random jump / boost
    buttons and flags, a few demoed cars, and force queues already holding 0-5 forces.  For air control: flip
    flags and flip timers on both sides of the pitch lock, pitch-lock restart times, tick-start angular velocities,
    authoritative bases, restored inverse-inertia rows and stick axes at zero.  Returns (cars, controls, queues)."""
    import rocketgpu_amd as rg
    lay = {f[0]: f[2] for f in rg.state_layouts()["car_states"]["fields"]}
    rng = np.random.default_rng(seed)
    c = cars.copy()
    k = controls.copy()
    n = len(c)

    def put(i, name, v):
        c[i, lay[name]:lay[name] + 4] = np.frombuffer(np.float32(v).tobytes(), np.uint8)

    for i in range(n):
        boost = np.float32(rng.choice([0.0, 0.2, 33.3, rng.uniform(0, 100), 100.0, 100.5, 120.0, -3.0]))
        put(i, "boost", boost)
        put(i, "boost_amount", np.float32(boost / np.float32(100)) if rng.random() < 0.7 else rng.uniform(0, 1))
        put(i, "boosting_time", rng.choice([0.0, 0.05, 0.0999, 0.1, 0.1001, 0.3]))
        put(i, "time_since_boosted", rng.choice([0.0, 0.1, 0.2416, 0.2417, 0.25, 0.3, 1.0]))
        put(i, "internal_jump_time", rng.choice([0.0, 0.01, 0.0249, 0.03, 0.0499, 0.05, 0.06, 0.19, 0.2, 0.25, 0.5]))
        for name, p in (("is_boosting", 0.5), ("is_on_ground", 0.5), ("is_jumping", 0.4), ("has_jumped", 0.5),
                        ("jump_hold_broken", 0.2), ("is_demoed", 0.08), ("is_flipping", 0.3), ("has_flipped", 0.5),
                        ("basis_authoritative", 0.3)):
            c[i, lay[name]] = 1 if rng.random() < p else 0
        put(i, "internal_flip_time", rng.choice([0.0, 0.02, 0.041, 0.3, 0.5, 0.64, 0.65, 0.94, 0.95, 0.96, 1.5]))
        rel = np.zeros(3, np.float32)
        rel[:2] = ((0, 0), (1, 0), (-1, 0), (0, 1), (0, -1), (0.6, -0.8), (-0.3, 0.95))[rng.integers(7)]
        c[i, lay["flip_rel_torque"]:lay["flip_rel_torque"] + 12] = rel.view(np.uint8)
        put(i, "cal_pre_pitchlock_restart_time", rng.choice([-1.0, 0.0, 0.1, 0.29, 0.3, 0.5]))
        w = rng.uniform(-5, 5, 3).astype(np.float32)
        c[i, lay["cal_tick_start_ang_vel"]:lay["cal_tick_start_ang_vel"] + 12] = w.view(np.uint8)
        if rng.random() < 0.1:
            c[i, lay["restored_inv_inertia_valid"]] = 1
            rows = rng.uniform(0, 0.3, 9).astype(np.float32)
            c[i, lay["restored_inv_inertia_bt"]:lay["restored_inv_inertia_bt"] + 36] = rows.view(np.uint8)
    kc = k.view(R.CONTROLS_DTYPE).reshape(-1)
    kc["jump"] = rng.random(n) < 0.5
    kc["boost"] = rng.random(n) < 0.5
    still = rng.random(n) < 0.15
    for axis in ("pitch", "yaw", "roll"):
        kc[axis][still] = 0.0
    q = np.zeros((n, R.FORCE_QUEUE_BYTES), np.uint8)
    qf = q.view(np.float32)
    qi = q.view(np.int32)
    qf[:, 0:12] = rng.uniform(-40, 40, (n, 12))
    qi[:, 12] = rng.random(n) < 0.3
    qi[:, 13] = rng.integers(0, 6, n)
    qi[:, 14] = rng.integers(0, 6, n)
    qf[:, 15:39] = rng.uniform(-40, 40, (n, 24))
    return c, k, q


def make_car_force_phases(name, arenas, cars_per_env, ticks, seed, keep):
    """Force-phase inputs (the pre-tick's outputs of car-arena states, perturbed, with each car's config) and the
    reference's jump hold, boost and air control on each, with and without a force queue, under the default
    parameters and with boost recharge on."""
    run = R.car_arena_run(arenas, cars_per_env, ticks, seed)
    cars, controls = [], []
    for t in range(ticks):
        mc, _ = R.car_pretick(arenas, cars_per_env, seed, run["cars"][t], run["controls"][t], run["configs"],
                              run["wheels"][t], run["envs"][t], run["teams"])
        cars.append(mc)
        controls.append(run["controls"][t])
    cars, controls = np.concatenate(cars), np.concatenate(controls)
    configs = np.tile(run["configs"], (ticks, 1))
    rng = np.random.default_rng(seed)
    pick = np.sort(rng.choice(len(cars), keep, replace=False))
    c, k, q = perturb_force_inputs(cars[pick], controls[pick], seed)
    cfg = configs[pick]
    params = force_param_sets()
    out_cars, out_queues = [], []
    for phase, queued, ps, flags in FORCE_CASES:
        oc, oq = R.car_force_phase(phase, c, k, q if queued else None, params[ps], cfg, flags)
        out_cars.append(oc)
        out_queues.append(oq if queued else np.zeros_like(q))
    meta = {"workload": "car force-phase inputs: pre-ticked car-arena states with seeded boost and jump timers, "
                        "flags, buttons and queued forces", "arenas": arenas, "cars_per_env": cars_per_env,
            "ticks": ticks, "seed": seed, "rows": int(keep),
            "generator": "oracle/_ref reference CPU build (UpdateJumpState, ApplyBoost, ApplyAirControl)"}
    path = os.path.join(HERE, name + ".npz")
    # Outputs as the XOR with the inputs (mostly zero bytes, so the fixture stays small).
    np.savez_compressed(path, cars=c, controls=k, configs=cfg, queues=q, params=params,
                        cases=np.array(FORCE_CASES, np.int32),
                        out_cars_xor=np.stack(out_cars) ^ c[None], out_queues=np.stack(out_queues),
                        meta=np.frombuffer(json.dumps(meta).encode(), np.uint8))
    print(f"{name}: {keep} rows x {len(FORCE_CASES)} cases -> {os.path.getsize(path)} bytes")


def main(names):
    for name in names or (list(FIXTURES) + list(BALL_PATH_FIXTURES) + list(ENV_COMMAND_FIXTURES) +
                          list(CAR_PRETICK_FIXTURES) + list(CAR_POSTTICK_FIXTURES) + list(CAR_FORCE_FIXTURES)):
        if name in FIXTURES:
            make(name, *FIXTURES[name])
        elif name in ENV_COMMAND_FIXTURES:
            make_env_commands(name, *ENV_COMMAND_FIXTURES[name])
        elif name in CAR_PRETICK_FIXTURES:
            make_car_pretick(name, *CAR_PRETICK_FIXTURES[name])
        elif name in CAR_POSTTICK_FIXTURES:
            make_car_posttick(name, *CAR_POSTTICK_FIXTURES[name])
        elif name in CAR_FORCE_FIXTURES:
            make_car_force_phases(name, *CAR_FORCE_FIXTURES[name])
        else:
            make_ball_path(name, *BALL_PATH_FIXTURES[name])


if __name__ == "__main__":
    main(sys.argv[1:])
