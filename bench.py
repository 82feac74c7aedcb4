"""Benchmark of the batched arena step on one MI355X: arena-ticks/s, one JSON line.

A step is one call of step_with_device_controls(controls[T, car_capacity], T ticks) over the whole batch of arenas,
with the state resident in HBM.  Timing: HIP events recorded on the stream the step kernel is launched on, around
each step; the timed region (K steps) is bracketed by device synchronization.

The workload is 16,384 ball arenas on the procedural arena PA(1, 256) (20,332 triangles) with seeded ball launches
and T = 8 ticks per step.  Cars are not simulated by this build yet, so the arenas carry no cars; config.workload
says so.  The roofline object prices the step kernel's algorithmic bytes per arena-tick against the 8 TB/s HBM
peak: 2 x (BallState 288 + EnvState 44 + 34 pads x 16 + world solver state 11,712) + 4 = 25,180 B at 0 cars.
cpu_baseline times the reference CPU build (oracle/_ref) on this machine's host cores over the same tick window of
the same workload.

Beside the headline:
- the launch shape, which the tuner picks over lanes per workgroup, where the arena state and the per-tick contact
  scratch live (HBM in place, lane-private memory, LDS), and arenas per wavefront (64 down to 4: more waves per CU,
  less divergence);
- the other step forms, each on a fresh batch over the same window: T launches of one tick, and hipGraph replays;
- the serving loop (simulated clients, Poisson arrivals, continuous batching over 262,144 resident arenas), with
  p50/p99 request latency.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--arenas N] [--ticks T] [--block B --stage S --per-wave A]
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
    p.add_argument("--steps", type=int, default=100)
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
    p.add_argument("--traffic", default=os.path.join(REPO, "profiles", "step_ball_arenas_16384_traffic.json"))
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


def cpu_baseline(args, init_ball, init_env, start_tick, ticks):
    """The reference CPU build on this host: the same arenas over the same tick window as the GPU's timed steps."""
    if ticks <= 0:
        return None
    cores = min(16, os.cpu_count() or 1)
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
    if args.split_sweep:
        split_sweep(args, rga, np, torch, stream, n, T)
        return
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
    if os.path.exists(args.traffic):
        with open(args.traffic) as fh:
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
                     "kernel": kernel_name(shape), "kernel_avg_ms": avg_kernel_s * 1e3,
                     "bytes_per_arena_tick": bytes_per_arena_tick(0)},
        "forms": forms,
        "serving": serving,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
