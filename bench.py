"""Benchmark of the batched arena step on one MI355X: arena-ticks/s, one JSON line.

A step is one call of step_with_device_controls(controls[T, car_capacity], T ticks) over the whole batch of arenas,
with the state resident in HBM.  Timing: HIP events recorded on the stream the step kernel is launched on, around
each step; the timed region (K steps) is bracketed by device synchronization.

Headline workload (--workload mixed, the default): 16,384 arenas on the procedural arena PA(1, 256) (20,332
triangles), 55% 1v1 / 27% 2v2 / 18% 3v3 (drawn per arena from the seed), every car's body drawn uniformly from the
six stock hitboxes, teams split evenly, a new-match kickoff (reset_to_kickoff with settled rows) and seeded held
controls ([64 ticks, car_capacity] on the device, step k reading rows 8k mod 64 onwards), T = 8 ticks per step.
Arenas that meet a body contact this build does not compute yet (car-mesh, ball-car, car-car) carry a nonzero
unsupported word from then on; the line reports their share at the end of the run.

The roofline object prices the step kernel at the reference's algorithmic bytes per arena-tick
B(n) = 2 (12588 + 2456 n + 130 n (n - 1)) + 144 n + 4 for n cars (44,195 B at the 55/27/18 mix) against the 8 TB/s
HBM peak; `traffic` is the measured HBM bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, scaled by a plain
copy's counters) when a traffic file for this workload exists, and `limiter` says what the counters show bounds
the kernel.  cpu_baseline times the reference CPU build (oracle/_ref) on this host's cores over the same arenas and
tick window: a warm-up window, then the median of three consecutive windows.

--workload ball keeps the ball-only workload (16,384 arenas, no cars) with its launch-shape tuner, step forms and
serving loop.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--arenas N] [--ticks T] [--workload mixed|ball]
                    [--no-forms] [--no-serve] [--cpu-ticks C] [--calibrate-copy-mib M]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

HBM_PEAK_GBS = 8000.0


def kernel_name(shape):
    """The step kernel a launch shape runs (rocketgpu_amd/csrc/rg_batch.hip)."""
    block, stage, per_wave = shape
    if stage in (6, 7):
        return f"StepBallArenasSplitKernel<{32 if stage == 6 else 16}, false>"
    if stage in (4, 5):
        return f"StepBallArenasGroupKernel<{64 // per_wave}, {'true' if stage == 5 else 'false'}, false>"
    return f"StepBallArenasKernel<{stage}, false>"


# Algorithmic bytes per arena-tick at n cars: B(n) = 2 * (12588 + 2456 n + 130 n (n - 1)) + 144 n + 4.
def bytes_per_arena_tick(n_cars: float) -> float:
    return 2.0 * (12588 + 2456 * n_cars + 130 * n_cars * (n_cars - 1)) + 144 * n_cars + 4


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=5)
    p.add_argument("--arenas", type=int, default=16384)
    p.add_argument("--ticks", type=int, default=8, help="ticks per step (T of step_with_device_controls)")
    p.add_argument("--max-cars", type=int, default=6)
    p.add_argument("--seed", type=int, default=2024)
    p.add_argument("--block", type=int, default=0, help="lanes per workgroup (0: the tuner chooses the shape)")
    p.add_argument("--stage", type=int, default=0,
                   help="state staging with --block: 0 in place, 1 private copy, 2 LDS scratch, 3 LDS state+scratch, "
                        "4 in place with 64/per-wave lanes per arena sharing its tree walks, 5 as 4 with the scratch in LDS, "
                        "6/7 as 5 with 8 lanes per arena and 32/16 for the arenas with the longest contact walks")
    p.add_argument("--per-wave", type=int, default=64, help="arenas per 64-lane wavefront with --block")
    p.add_argument("--split-sweep", action="store_true",
                   help="time launch stages 6 and 7 over a grid of split thresholds and capacities (block 64) "
                        "instead of the benchmark, one JSON line")
    p.add_argument("--no-forms", action="store_true")
    p.add_argument("--no-serve", action="store_true")
    p.add_argument("--serve-arenas", type=int, default=262144)
    p.add_argument("--cpu-ticks", type=int, default=-1,
                   help="ticks of the reference CPU sample (-1: the GPU's timed window; 0: skip)")
    p.add_argument("--traffic", default="")
    p.add_argument("--workload", choices=("mixed", "ball"), default="mixed")
    p.add_argument("--calibrate-copy-mib", type=int, default=0,
                   help="first run one plain device-to-device copy of this many MiB (a known-size reference for "
                        "the FETCH_SIZE / WRITE_SIZE counters)")
    return p.parse_args()


def timed(torch, stream, fn, reps):
    """Average ms per call of fn() from HIP events recorded on `stream` around each call."""
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(reps)]
    for a, b in ev:
        a.record(stream)
        fn()
        b.record(stream)
    stream.synchronize()
    return sum(a.elapsed_time(b) for a, b in ev) / reps


def make_batch(rga, np, n, max_cars, seed):
    sim = rga.BatchSim(n, max_cars)
    sim.set_env_car_count(np.zeros(n, np.int32))
    init_env = sim.download_state_array("env_states")
    init_ball = rga.workload_ball_states(seed, 0, n, 1.0)
    sim.upload("ball_states", init_ball)
    return sim, init_ball, init_env


def host_cores():
    """Threads for the reference CPU build: the cores this process may run on, within the host's CPU share when the
    environment states one (OMP_NUM_THREADS)."""
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    share = os.environ.get("OMP_NUM_THREADS", "")
    if share.isdigit() and int(share) > 0:
        cores = min(cores, int(share))
    return cores


def cpu_baseline(args, init_ball, init_env, start_tick, ticks):
    """The reference CPU build on this host: the same arenas over the same tick window as the GPU's timed steps."""
    if ticks <= 0:
        return None
    cores = host_cores()
    os.environ["ROCKETGPU_CPU_THREADS"] = str(cores)
    import numpy as np
    from oracle import refsim as R

    if not R.available():
        return {"value": None, "unit": "arena-ticks/s", "cores": cores, "kind": "reference",
                "sample": "reference CPU build (oracle/_ref) not present on this machine"}
    n = init_ball.shape[0]
    ref = R.RefSim(n, 2, 1.0, 256.0)
    ref.set_env_car_count(np.zeros(n, np.int32))
    ref.upload(R.ENV, init_env)
    ref.upload(R.BALL, init_ball)
    if start_tick > 0:
        ref.step(start_tick)
    t0 = time.perf_counter()
    ref.step(ticks)
    dt = time.perf_counter() - t0
    return {"value": n * ticks / dt, "unit": "arena-ticks/s", "cores": cores, "kind": "reference",
            "sample": f"{n} ball arenas, ticks {start_tick}..{start_tick + ticks} (the GPU's timed window) in "
                      f"{dt:.2f} s, ROCKETGPU_CPU_THREADS={cores}"}


def split_sweep(args, rga, np, torch, stream, n, T):
    """Stages 6 and 7 at block 64 over heavy-arena thresholds (contact-walk windows) and list capacities: a fresh
    batch per point, the warm-up window, then args.steps launches of T ticks timed with HIP events."""
    def point(stage, rounds, cap_divisor):
        sim, _, _ = make_batch(rga, np, n, args.max_cars, args.seed)
        sim.set_launch_shape(64, stage, 8)
        sim.set_split(rounds, cap_divisor)
        for _ in range(args.warmup):
            sim.step(T, stream=stream)
        ms = timed(torch, stream, lambda: sim.step(T, stream=stream), args.steps)
        sim.close()
        return ms
    out = {"stage5_ms": None, "points": []}
    sim, _, _ = make_batch(rga, np, n, args.max_cars, args.seed)
    sim.set_launch_shape(64, 5, 8)
    for _ in range(args.warmup):
        sim.step(T, stream=stream)
    out["stage5_ms"] = timed(torch, stream, lambda: sim.step(T, stream=stream), args.steps)
    sim.close()
    for stage in (6, 7):
        for rounds in (16, 20, 24, 28, 32, 40):
            for cap_divisor in (32, 16, 8):
                out["points"].append({"stage": stage, "rounds": rounds, "cap_divisor": cap_divisor,
                                      "ms": point(stage, rounds, cap_divisor)})
                print(json.dumps(out["points"][-1]), file=sys.stderr, flush=True)
    print(json.dumps({"split_sweep": out, "arenas": n, "ticks_per_launch": T, "steps": args.steps,
                      "warmup": args.warmup}))


MIX = (0.55, 0.27, 0.18)  # 1v1 / 2v2 / 3v3
CTRL_TICKS = 64           # rows of the seeded control tensor on the device


def mixed_workload(np, n, seed):
    """Per arena car count (2, 4 or 6 at the 55/27/18 mix), per car slot body (0-5: Octane, Dominus, Plank,
    Breakout, Hybrid, Merc, uniform) and team (first half of an arena's cars Blue, the rest Orange)."""
    rng = np.random.default_rng(seed)
    u = rng.random(n)
    counts = np.where(u < MIX[0], 2, np.where(u < MIX[0] + MIX[1], 4, 6)).astype(np.int32)
    presets = rng.integers(0, 6, n * 6).astype(np.uint8)
    slot = np.arange(n * 6) % 6
    teams = (slot >= np.repeat(counts, 6) // 2).astype(np.uint8)
    return counts, presets, teams


def make_car_batch(rga, np, n, seed, workload):
    counts, presets, teams = workload
    table = np.stack([rga.car_preset_config(i) for i in range(6)])
    sim = rga.BatchSim(n, 6)
    sim.set_env_car_count(counts)
    sim.set_teams(teams)
    sim.set_car_configs(table[presets])
    sim.reset_to_kickoff(seed)
    return sim


def cpu_baseline_cars(args, workload, controls, first_step, steps, T):
    """The reference CPU build on this host over the same arenas, kickoff and controls as the GPU: steps up to
    `first_step` untimed, then three consecutive windows of ceil(steps / 3) steps of T ticks; the median rate."""
    if steps <= 0:
        return None
    cores = host_cores()
    os.environ["ROCKETGPU_CPU_THREADS"] = str(cores)
    import numpy as np
    from oracle import refsim as R

    if not R.available():
        return {"value": None, "unit": "arena-ticks/s", "cores": cores, "kind": "reference",
                "sample": "reference CPU build (oracle/_ref) not present on this machine"}
    counts, presets, teams = workload
    n = counts.size
    ref = R.RefSim(n, 6, respawn_seed=args.seed)
    ref.set_env_car_count(counts)
    ref.set_teams(teams)
    ref.set_car_presets(presets)
    ref.reset_to_kickoff(args.seed)
    rows = controls.reshape(CTRL_TICKS, -1)

    def run(k0, k1):
        for k in range(k0, k1):
            r = (k * T) % CTRL_TICKS
            ref.step_with_controls(rows[r:r + T], T, n * 6)

    run(0, first_step)
    per = max(1, -(-steps // 3))
    rates = []
    for w in range(3):
        k0 = first_step + w * per
        t0 = time.perf_counter()
        run(k0, k0 + per)
        rates.append(n * per * T / (time.perf_counter() - t0))
    ref.close()
    return {"value": float(np.median(rates)), "unit": "arena-ticks/s", "cores": cores, "kind": "reference",
            "host_cpus": os.cpu_count(),
            "sample": f"{n} mixed arenas (same kickoff, bodies and controls as the GPU run), warm-up to tick "
                      f"{first_step * T}, then 3 windows of {per * T} ticks; median of "
                      f"{', '.join(f'{r / 1e6:.3f}' for r in rates)} M arena-ticks/s; ROCKETGPU_CPU_THREADS={cores}; "
                      "the reference computes every body contact, this build flags the arenas that meet one"}


def bench_cars(args, rga, np, torch, stream):
    n, T = args.arenas, args.ticks
    if CTRL_TICKS % T:
        raise SystemExit(f"--ticks must divide {CTRL_TICKS}")
    workload = mixed_workload(np, n, args.seed)
    counts = workload[0]
    controls = rga.workload_controls(args.seed, 0, n, 6, CTRL_TICKS)
    ctrl = torch.from_numpy(controls.reshape(CTRL_TICKS, -1)).cuda()
    row_bytes = ctrl.shape[1]

    def ctrl_ptr(k):
        return ctrl.data_ptr() + ((k * T) % CTRL_TICKS) * row_bytes

    # Arenas per wavefront of the car kernel (fewer arenas per wave, more waves per CU): the
    # tuner times each on a fresh batch over the warm-up window plus five steps and keeps the fastest.
    # Staging 0 keeps each arena's per-tick working data in lane-private memory, 2 in LDS (one slot per lane); 4 gives
    # each arena 2 lanes (32 arenas per wave) or 4 (16 or 8), which split its wheel rays' tree walks.
    sweep = {}
    shape = (args.stage, args.per_wave) if args.block else None
    if shape is None:
        key = lambda c: f"stage{c[0]}_perwave{c[1]}"
        for cand in ((0, 64), (0, 32), (0, 16), (4, 32), (4, 16), (4, 8), (2, 32), (2, 16), (2, 8), (2, 4)):
            tsim = make_car_batch(rga, np, n, args.seed, workload)
            tsim.set_launch_shape(64, *cand)
            try:
                for k in range(args.warmup):
                    tsim.step_with_device_controls(ctrl_ptr(k), T, tsim.car_capacity, stream=stream)
                sweep[key(cand)] = timed(torch, stream, lambda: tsim.step_with_device_controls(
                    ctrl_ptr(args.warmup), T, tsim.car_capacity, stream=stream), 5)
            except rga.RocketGpuError as err:  # the LDS slots do not fit this device
                sweep[key(cand)] = str(err)
                tsim.close()
                continue
            tsim.close()
            if shape is None or sweep[key(cand)] < sweep[key(shape)]:
                shape = cand
    stage, per_wave = shape

    sim = make_car_batch(rga, np, n, args.seed, workload)
    sim.set_launch_shape(64, stage, per_wave)
    cap = sim.car_capacity
    torch.cuda.synchronize()
    for k in range(args.warmup):
        sim.step_with_device_controls(ctrl_ptr(k), T, cap, stream=stream)
    stream.synchronize()
    torch.cuda.synchronize()
    events = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
              for _ in range(args.steps)]
    t0 = time.perf_counter()
    for k, (a, b) in enumerate(events):
        a.record(stream)
        sim.step_with_device_controls(ctrl_ptr(args.warmup + k), T, cap, stream=stream)
        b.record(stream)
    stream.synchronize()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_ms = [a.elapsed_time(b) for a, b in events]
    avg_kernel_s = sum(kernel_ms) / len(kernel_ms) / 1e3
    envs = sim.download_state_array("env_states")
    ticks_done = envs[:, 0:4].copy().view(np.uint32)[:, 0]
    assert np.all(ticks_done == (args.warmup + args.steps) * T), "tick counters do not match the steps run"
    unsupported = sim.download_state_array("unsupported").view(np.uint32).ravel()
    sim.close()

    arena_ticks = n * T * args.steps
    value = arena_ticks / wall
    b_arena = float(np.mean([bytes_per_arena_tick(c) for c in counts]))
    # Lanes per arena of the car kernel: 2 or 4 at stage 4 (32 or at most 16 arenas per wave), else 1.
    lanes = (2 if per_wave == 32 else 4) if stage == 4 else 1
    achieved = n * T * b_arena / avg_kernel_s / 1e9
    traffic = None
    traffic_file = args.traffic or os.path.join(REPO, "profiles", "step_car_arenas_16384_traffic.json")
    if os.path.exists(traffic_file):
        with open(traffic_file) as fh:
            tr = json.load(fh)
        shape_of = tr.get("measured_with", {})
        if (tr.get("arenas") == n and tr.get("ticks") == T and tr.get("seed") == args.seed
                and shape_of.get("stage") == stage and shape_of.get("arenas_per_wave") == per_wave):
            traffic = tr.get("hbm_bytes_per_launch")

    forms = None
    if not args.no_forms:
        # Each form on a fresh batch from the same kickoff over the same tick window, with the batch's own control
        # row (zero after the kickoff) instead of the control tensor: T single-tick launches, one T-tick launch,
        # and hipGraph replays of both.
        def form(run):
            fsim = make_car_batch(rga, np, n, args.seed, workload)
            fsim.set_launch_shape(64, stage, per_wave)
            for _ in range(args.warmup):
                run(fsim)
            ms = timed(torch, stream, lambda: run(fsim), args.steps)
            fsim.close()
            return ms

        forms = {"T": T, "window_ticks": [args.warmup * T, (args.warmup + args.steps) * T], "ms_per_T_ticks": {
            "plain_T_launches_of_1_tick": form(lambda s: [s.step(1, stream=stream) for _ in range(T)]),
            "one_launch_of_T_ticks": form(lambda s: s.step(T, stream=stream)),
            "hipgraph_T_launches_of_1_tick": form(lambda s: s.step_graph(1, T, stream=stream)),
            "hipgraph_one_launch_of_T_ticks": form(lambda s: s.step_graph(T, 1, stream=stream))}}

    cpu_steps = args.steps if args.cpu_ticks < 0 else -(-args.cpu_ticks // T)
    cpu = cpu_baseline_cars(args, workload, controls, args.warmup, cpu_steps, T)
    mix = {str(c): round(float((counts == c).mean()), 4) for c in (2, 4, 6)}
    line = {
        "metric": "arena-ticks/s", "value": value, "unit": "arena-ticks/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded car counts, bodies and held controls; kickoff start; procedural arena)",
        "config": {"workload": f"{n} mixed arenas (1v1/2v2/3v3 at {MIX[0]:.0%}/{MIX[1]:.0%}/{MIX[2]:.0%}, six "
                               f"stock bodies uniform per car), procedural arena PA(1, 256) 20,332 triangles, "
                               f"step_with_device_controls T={T}",
                   "arenas": n, "ticks_per_step": T, "cars_per_arena_mix": mix,
                   "cars": int(counts.sum()), "parallelism": "single GPU",
                   "kernel_launch": f"{lanes} lane{'s' if lanes > 1 else ''} per arena, {per_wave} arenas per "
                                    f"64-lane wavefront, per-tick data in "
                                    f"{'LDS' if stage == 2 else 'lane-private memory'}"
                                    + (", the arena's wheel-ray tree walks split between its lanes" if lanes > 1
                                       else ""),
                   "launch_shape_sweep_ms": sweep or None},
        "arenas_unsupported_frac_at_end": float((unsupported != 0).mean()),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "limiter": "latency and divergence (each arena's broadphase, contacts, solver and vehicle "
                                "dynamics run serially on its lanes, whose chain of dependent loads sets the time; "
                                "counters: 67% of wave cycles waiting on memory, a quarter of the lanes active per "
                                "VALU instruction, one wave per SIMD at most at 256 VGPRs; measured traffic 1.12x "
                                "the algorithmic bytes)",
                     "kernel": f"StepCarArenasKernel<{'true' if stage == 2 else 'false'}, {lanes}>",
                     "kernel_avg_ms": avg_kernel_s * 1e3,
                     "bytes_per_arena_tick": b_arena,
                     "traffic_gbs": traffic / avg_kernel_s / 1e9 if traffic else None,
                     "traffic_frac": traffic / avg_kernel_s / 1e9 / HBM_PEAK_GBS if traffic else None},
        "forms": forms,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (--gpus {args.gpus} requested); arenas are independent, "
              "so the multi-GPU form is one process per GPU on its share of the arenas.", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    import rocketgpu_amd as rga

    torch.cuda.set_device(0)
    n, T = args.arenas, args.ticks
    if args.calibrate_copy_mib > 0:
        src = torch.ones(args.calibrate_copy_mib << 18, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    stream = torch.cuda.Stream()
    if args.workload == "mixed":
        bench_cars(args, rga, np, torch, stream)
        return
    if args.split_sweep:
        split_sweep(args, rga, np, torch, stream, n, T)
        return
    bench_balls(args, rga, np, torch, stream)


def bench_balls(args, rga, np, torch, stream):
    """The ball-only workload: 16,384 arenas without cars, seeded ball launches."""
    n, T = args.arenas, args.ticks
    traffic_file = args.traffic or os.path.join(REPO, "profiles", "step_ball_arenas_16384_traffic.json")
    g = torch.Generator(device="cuda")
    g.manual_seed(args.seed)

    # Launch shape (lanes per workgroup, state staging): the tuner times every shape on a fresh copy of the workload
    # over the warm-up window plus ten steps and keeps the fastest.
    shape = (args.block, args.stage, args.per_wave) if args.block else None
    sweep = {}
    if shape is None:
        key = lambda c: f"block{c[0]}_stage{c[1]}_perwave{c[2]}"
        cands = ([(64, s, w) for s in (0, 1) for w in (64, 32, 16, 8, 4)] + [(128, 0, 64), (256, 0, 64)] +
                 [(64, 2, w) for w in (64, 32, 16, 8)] + [(64, 3, w) for w in (16, 8, 4)] + [(128, 3, 4), (256, 3, 4)] +
                 [(b, st, w) for st in (4, 5) for b in (64, 128, 256) for w in (16, 8, 4)] +
                 [(b, st, 8) for st in (6, 7) for b in (64, 128, 256)])
        for cand in cands:
            sim_b, _, _ = make_batch(rga, np, n, args.max_cars, args.seed)
            try:
                sim_b.set_launch_shape(*cand)
            except rga.RocketGpuError as err:  # the shape's LDS does not fit this device
                sweep[key(cand)] = str(err)
                sim_b.close()
                continue
            for _ in range(args.warmup):
                sim_b.step(T, stream=stream)
            sweep[key(cand)] = timed(torch, stream, lambda: sim_b.step(T, stream=stream), 10)
            sim_b.close()
            if shape is None or sweep[key(cand)] < sweep[key(shape)]:  # shape is always a timed candidate
                shape = cand

    sim, init_ball, init_env = make_batch(rga, np, n, args.max_cars, args.seed)
    sim.set_launch_shape(*shape)
    ctrl = torch.randint(0, 256, (T, sim.car_capacity * 24), dtype=torch.uint8, device="cuda", generator=g)
    torch.cuda.synchronize()
    for _ in range(args.warmup):
        sim.step_with_device_controls(ctrl.data_ptr(), T, sim.car_capacity, stream=stream)
    stream.synchronize()
    torch.cuda.synchronize()

    events = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
              for _ in range(args.steps)]
    t0 = time.perf_counter()
    for a, b in events:
        a.record(stream)
        sim.step_with_device_controls(ctrl.data_ptr(), T, sim.car_capacity, stream=stream)
        b.record(stream)
    stream.synchronize()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_ms = [a.elapsed_time(b) for a, b in events]
    avg_kernel_s = sum(kernel_ms) / len(kernel_ms) / 1e3

    arena_ticks = n * T * args.steps
    value = arena_ticks / wall
    per_launch_bytes = n * T * bytes_per_arena_tick(0)
    achieved = per_launch_bytes / avg_kernel_s / 1e9
    traffic = None
    if os.path.exists(traffic_file):
        with open(traffic_file) as fh:
            tr = json.load(fh)
        # Measured per kernel form: the staging and lanes per arena decide the traffic, the block size does not.
        for entry in tr.get("entries", []):
            if (tr.get("arenas") == n and tr.get("ticks") == T and entry.get("stage") == shape[1]
                    and entry.get("arenas_per_wave") == shape[2]):
                traffic = entry.get("hbm_bytes_per_launch")

    envs = sim.download_state_array("env_states")
    ticks_done = envs[:, 0:4].copy().view(np.uint32)[:, 0]
    assert np.all(ticks_done == (args.warmup + args.steps) * T), "tick counters do not match the steps run"

    sim.close()

    forms = None
    if not args.no_forms:
        # Each form on a fresh batch from the same initial states over the same tick window as the headline
        # (warm-up, then the timed steps): T single-tick launches, one T-tick launch, and hipGraph replays of both.
        def form(run):
            fsim, _, _ = make_batch(rga, np, n, args.max_cars, args.seed)
            fsim.set_launch_shape(*shape)
            for _ in range(args.warmup):
                run(fsim)
            ms = timed(torch, stream, lambda: run(fsim), args.steps)
            fsim.close()
            return ms

        forms = {"T": T, "window_ticks": [args.warmup * T, (args.warmup + args.steps) * T], "ms_per_T_ticks": {
            "plain_T_launches_of_1_tick": form(lambda s: [s.step(1, stream=stream) for _ in range(T)]),
            "one_launch_of_T_ticks": form(lambda s: s.step(T, stream=stream)),
            "hipgraph_T_launches_of_1_tick": form(lambda s: s.step_graph(1, T, stream=stream)),
            "hipgraph_one_launch_of_T_ticks": form(lambda s: s.step_graph(T, 1, stream=stream))}}

    serving = None
    if not args.no_serve:
        ssim, _, _ = make_batch(rga, np, args.serve_arenas, 2, args.seed + 1)
        ssim.set_launch_shape(*shape)
        serving = ssim.serve(clients=256, min_arenas=16, max_arenas=256, horizons=(1, 4, 8),
                             max_arenas_per_launch=65536, max_in_flight=2, request_rate_hz=2000.0, duration_s=2.0,
                             seed=args.seed)
        serving["config"] = ("256 clients x U{16..256} resident ball arenas, Poisson 2,000 req/s aggregate for 2 s, "
                             "k in {1, 4, 8} ticks, <= 65,536 arenas per launch, 2 launches in flight; "
                             f"{args.serve_arenas} arenas resident in HBM")
        ssim.close()

    cpu_ticks = args.steps * T if args.cpu_ticks < 0 else args.cpu_ticks
    cpu = cpu_baseline(args, init_ball, init_env, args.warmup * T, cpu_ticks)
    line = {
        "metric": "arena-ticks/s", "value": value, "unit": "arena-ticks/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded ball launches, procedural arena)",
        "config": {"workload": f"{n} ball arenas (no cars), procedural arena PA(1, 256) 20,332 triangles, "
                               f"step_with_device_controls T={T}",
                   "arenas": n, "ticks_per_step": T, "cars_per_arena": 0, "parallelism": "single GPU",
                   "launch_shape": {"block": shape[0], "stage": shape[1], "arenas_per_wave": shape[2]}, "launch_shape_sweep_ms": sweep or None},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "limiter": "latency (dependent BVH and state loads at low occupancy; measured traffic is a "
                                "small fraction of the algorithmic bytes)",
                     "kernel": kernel_name(shape), "kernel_avg_ms": avg_kernel_s * 1e3,
                     "bytes_per_arena_tick": bytes_per_arena_tick(0)},
        "forms": forms,
        "serving": serving,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
