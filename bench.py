"""Benchmark of the batched arena step on one MI355X: arena-ticks/s, one JSON line.

A step is one call of step_with_device_controls(controls[T, car_capacity], T ticks) over the whole batch of arenas,
with the state resident in HBM.  Timing: HIP events recorded on the stream the step kernel is launched on, around
each step; the timed region (K steps) is bracketed by device synchronization.

The workload is 16,384 ball arenas on the procedural arena PA(1, 256) (20,332 triangles) with seeded ball launches
and T = 8 ticks per step; cars are not simulated by this build yet, so the arenas carry no cars (named as such in
config.workload).  The roofline object prices the step kernel's algorithmic bytes per arena-tick
(2 x (BallState 288 + EnvState 44 + 34 pads x 16 + world solver state 11,712) + 4 = 25,180 B at 0 cars) against
the 8 TB/s HBM peak.  cpu_baseline times the reference CPU build (oracle/_ref) on a bounded sample of the same
workload on this machine's host cores.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--arenas N] [--ticks T] [--cpu-ticks C]
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
    p.add_argument("--cpu-ticks", type=int, default=240, help="ticks of the reference CPU sample (0: skip)")
    p.add_argument("--traffic", default=os.path.join(REPO, "profiles", "step_ball_arenas_16384_traffic.json"))
    return p.parse_args()


def cpu_baseline(args, init_ball, init_env):
    """The reference CPU build on this host: the same arenas, args.cpu_ticks ticks after an 8-tick warm-up."""
    if args.cpu_ticks <= 0:
        return None
    cores = min(16, os.cpu_count() or 1)
    os.environ["ROCKETGPU_CPU_THREADS"] = str(cores)
    from oracle import refsim as R

    if not R.available():
        return {"value": None, "unit": "arena-ticks/s", "cores": cores, "kind": "reference",
                "sample": "reference CPU build (oracle/_ref) not present on this machine"}
    n = init_ball.shape[0]
    ref = R.RefSim(n, 2, 1.0, 256.0)
    ref.set_env_car_count(__import__("numpy").zeros(n, "int32"))
    ref.upload(R.ENV, init_env)
    ref.upload(R.BALL, init_ball)
    ref.step(8)
    t0 = time.perf_counter()
    ref.step(args.cpu_ticks)
    dt = time.perf_counter() - t0
    return {"value": n * args.cpu_ticks / dt, "unit": "arena-ticks/s", "cores": cores, "kind": "reference",
            "sample": f"{n} ball arenas x {args.cpu_ticks} ticks (after 8 warm-up ticks) in {dt:.2f} s, "
                      f"ROCKETGPU_CPU_THREADS={cores}"}


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
    sim = rga.BatchSim(n, args.max_cars)
    sim.set_env_car_count(np.zeros(n, np.int32))
    init_env = np.zeros((n, 44), np.uint8)
    # Default EnvState records (tick 0, no score, last_scored_team -1, frame fraction 0.5).
    init_env[:] = sim.download_state_array("env_states")[0]
    init_ball = rga.workload_ball_states(args.seed, 0, n, 1.0)
    sim.upload("env_states", init_env)
    sim.upload("ball_states", init_ball)

    stream = torch.cuda.Stream()
    g = torch.Generator(device="cuda")
    g.manual_seed(args.seed)
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
        if tr.get("arenas") == n and tr.get("ticks") == T:
            traffic = tr.get("hbm_bytes_per_launch")

    envs = sim.download_state_array("env_states")
    ticks_done = envs[:, 0:4].copy().view(np.uint32)[:, 0]
    assert np.all(ticks_done == (args.warmup + args.steps) * T), "tick counters do not match the steps run"
    sim.close()

    cpu = cpu_baseline(args, init_ball, init_env)
    line = {
        "metric": "arena-ticks/s", "value": value, "unit": "arena-ticks/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded ball launches, procedural arena)",
        "config": {"workload": f"{n} ball arenas (no cars), procedural arena PA(1, 256) 20,332 triangles, "
                               f"step_with_device_controls T={T}",
                   "arenas": n, "ticks_per_step": T, "cars_per_arena": 0, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "StepBallArenasKernel", "kernel_avg_ms": avg_kernel_s * 1e3,
                     "bytes_per_arena_tick": bytes_per_arena_tick(0)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
