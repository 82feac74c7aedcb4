"""Car arenas on the GPU (the product's car kernel through the C-ABI) against the reference CPU build and the host
twin, tick by tick.

The reference is the master (tests/car_cases.py): its kickoff records are uploaded (PreserveInternal), then the GPU
batch, the reference and the host twin step the same seeded held controls.  On every tick:
  - every arena's records on the GPU equal the host twin's, byte for byte (the same core compiled for gfx950 and
    for the host), flagged or not;
  - every arena whose unsupported word is 0 equals the reference byte for byte.
A T-tick launch (step_with_device_controls with T rows) equals the reference's T-tick step.
"""
import sys

import numpy as np
import pytest

from car_cases import arena_mismatches, reference_at_kickoff, reference_records, teammate_chase_controls
from conftest import oracle_or_skip, twin_cars

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _torch():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _gpu_records(sim):
    get = sim.download_state_array
    return {"cars": get("car_states"), "wheels": get("wheel_states"), "balls": get("ball_states"),
            "envs": get("env_states"), "pads": get("boost_pad_states")}, get("unsupported").view(np.uint32).ravel()


def _twin_records(T, h, n_env, cars):
    cap = n_env * cars
    def get(which, shape, dtype=np.uint8):
        a = np.zeros(shape, dtype)
        assert T.rgt_cars_download(h, which, a.ctypes.data, shape[0]) == 0
        return a
    return {"cars": get(0, (cap, 880)), "wheels": get(8, (cap, 292)), "balls": get(1, (n_env, 288)),
            "envs": get(2, (n_env, 44)), "pads": get(3, (n_env * 34, 16))}


def _setup(n_env, cars, seed, presets=None):
    R = oracle_or_skip()
    import rocketgpu_amd as rga

    ref, counts, teams, cfgs = reference_at_kickoff(R, n_env, cars, seed, presets)
    rec = reference_records(R, ref, n_env, cars)
    sim = rga.BatchSim(n_env, cars)
    sim.set_env_car_count(counts)
    sim.set_teams(teams)
    sim.set_car_configs(cfgs)
    for name, key in (("car_states", "cars"), ("wheel_states", "wheels"), ("ball_states", "balls"),
                      ("env_states", "envs"), ("boost_pad_states", "pads")):
        sim.upload(name, rec[key], preserve_internal=True)
    T = twin_cars()
    h = T.rgt_cars_create(n_env, cars, 1.0, 256.0, seed)
    for which, arr in ((6, counts), (7, teams), (5, cfgs), (0, rec["cars"]), (8, rec["wheels"]), (1, rec["balls"]),
                       (2, rec["envs"]), (3, rec["pads"])):
        a = np.ascontiguousarray(arr)
        assert T.rgt_cars_upload(h, which, a.ctypes.data, a.shape[0]) == 0
    return R, ref, sim, T, h


@pytest.mark.parametrize("n_env,cars,ticks,seed,mixed,min_frac",
                         [(96, 2, 600, 2024, False, 0.95), (48, 6, 240, 77, True, 0.9)])
def test_gpu_cars_match_twin_and_reference(n_env, cars, ticks, seed, mixed, min_frac):
    torch = _torch()
    import rocketgpu_amd as rga

    say(f"case: {n_env} arenas x {cars} cars, {ticks} ticks, seed {seed}, mixed bodies {mixed}")
    presets = np.arange(n_env * cars) % 6 if mixed else None
    R, ref, sim, T, h = _setup(n_env, cars, seed, presets)
    cap = n_env * cars
    controls = rga.workload_controls(seed, 0, n_env, cars, ticks)
    dev = torch.from_numpy(controls.reshape(ticks, -1)).to("cuda")
    compared = 0
    for t in range(ticks):
        ctl = np.ascontiguousarray(controls[t])
        sim.step_with_device_controls(dev[t].data_ptr(), 1, cap)
        ref.step_with_controls(ctl, 1, cap)
        assert T.rgt_cars_upload(h, 4, ctl.ctypes.data, cap) == 0
        assert T.rgt_cars_step(h, 1) == 0
        gpu, uns = _gpu_records(sim)
        twin = _twin_records(T, h, n_env, cars)
        diff_twin = arena_mismatches(gpu, twin, n_env, cars)
        assert diff_twin.size == 0, f"tick {t}: GPU differs from the host twin in arenas {diff_twin[:10]}"
        bad = arena_mismatches(gpu, reference_records(R, ref, n_env, cars), n_env, cars)
        live = uns == 0
        assert not live[bad].any(), f"tick {t}: unflagged arenas {bad[live[bad]][:10]} differ from the reference"
        compared += int(live.sum())
        if t % 100 == 99:
            say(f"  tick {t + 1}: {int(live.sum())} of {n_env} arenas unflagged")
    say(f"  arena-ticks compared with the reference: {compared} of {n_env * ticks}")
    assert compared >= min_frac * n_env * ticks
    T.rgt_cars_destroy(h)
    sim.close()
    ref.close()


def test_gpu_cars_multi_tick_launch_matches_reference():
    torch = _torch()
    import rocketgpu_amd as rga

    n_env, cars, seed, T_ticks = 64, 2, 31, 8
    say(f"case: {n_env} arenas x {cars} cars, launches of {T_ticks} ticks")
    R, ref, sim, T, h = _setup(n_env, cars, seed)
    cap = n_env * cars
    launches = 10
    controls = rga.workload_controls(seed, 0, n_env, cars, launches * T_ticks)
    dev = torch.from_numpy(controls.reshape(launches * T_ticks, -1)).to("cuda")
    for l in range(launches):
        rows = np.ascontiguousarray(controls[l * T_ticks:(l + 1) * T_ticks])
        sim.step_with_device_controls(dev[l * T_ticks].data_ptr(), T_ticks, cap)
        ref.step_with_controls(rows, T_ticks, cap)
    gpu, uns = _gpu_records(sim)
    bad = arena_mismatches(gpu, reference_records(R, ref, n_env, cars), n_env, cars)
    live = uns == 0
    say(f"  {int(live.sum())} of {n_env} arenas unflagged after {launches * T_ticks} ticks")
    assert not live[bad].any()
    assert live.sum() >= n_env // 4
    T.rgt_cars_destroy(h)
    sim.close()
    ref.close()


def test_gpu_cars_teammates_collide_match_twin_and_reference():
    """Teammates steered into each other (hitbox-hitbox contacts, bumps, wheels on other cars, the joint solve of
    touching cars): the GPU equals the host twin on every tick and the reference while unflagged."""
    torch = _torch()

    n_env, cars, ticks, seed = 16, 6, 1000, 9
    say(f"case: teammate chase, {n_env} arenas x {cars} cars, {ticks} ticks, seed {seed}")
    R, ref, sim, T, h = _setup(n_env, cars, seed, np.arange(n_env * cars) % 6)
    cap = n_env * cars
    theirs = reference_records(R, ref, n_env, cars)
    ever_flagged = np.zeros(n_env, bool)
    for t in range(ticks):
        ctl = teammate_chase_controls(theirs["cars"], n_env, cars)
        dev = torch.from_numpy(ctl).to("cuda")
        sim.step_with_device_controls(dev.data_ptr(), 1, cap)
        ref.step_with_controls(ctl, 1, cap)
        assert T.rgt_cars_upload(h, 4, ctl.ctypes.data, cap) == 0
        assert T.rgt_cars_step(h, 1) == 0
        gpu, uns = _gpu_records(sim)
        diff_twin = arena_mismatches(gpu, _twin_records(T, h, n_env, cars), n_env, cars)
        assert diff_twin.size == 0, f"tick {t}: GPU differs from the host twin in arenas {diff_twin[:10]}"
        theirs = reference_records(R, ref, n_env, cars)
        bad = arena_mismatches(gpu, theirs, n_env, cars)
        live = uns == 0
        assert not live[bad].any(), f"tick {t}: unflagged arenas {bad[live[bad]][:10]} differ from the reference"
        ever_flagged |= ~live
        if t % 250 == 249:
            say(f"  tick {t + 1}: {int(live.sum())} of {n_env} arenas unflagged")
    from rocketgpu_amd import views

    bumps = gpu["cars"].view(views.state_dtype("car_states"))["bump_event_bumps"].reshape(n_env, -1).sum(axis=1)
    say(f"  bumps in arenas never flagged: {int(bumps[~ever_flagged].sum())}")
    assert bumps[~ever_flagged].sum() >= 100
    assert (~ever_flagged).sum() >= n_env // 2
    T.rgt_cars_destroy(h)
    sim.close()
    ref.close()
