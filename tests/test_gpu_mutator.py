"""MutatorConfig at the boundary (set_mutator_config, rocketgpu.h:1003) against the reference CPU build with the
same non-default mutator: gravity, ball drag and restitution for ball arenas; gravity, car mass, jump and boost
accelerations for car arenas (settled kickoff rows recorded under the new physics, then seeded controls).  Fields
this build keeps fixed are refused by name."""
import sys

import numpy as np
import pytest

from car_cases import arena_mismatches, batch_records, reference_records
from conftest import oracle_or_skip

pytestmark = pytest.mark.gpu

BALL_MUT = dict(gravity_z=-500.0, ball_drag=0.05, ball_restitution=0.5, ball_friction=0.3)
CAR_MUT = dict(gravity_z=-540.0, car_mass=200.0, jump_accel=1500.0, boost_accel_ground=1100.0,
               car_air_control_pitch=150.0)


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _torch():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def test_gpu_mutator_ball_arenas_match_reference():
    _torch()
    R = oracle_or_skip()
    import rocketgpu_amd as rga

    n, ticks = 256, 600
    say(f"case: {n} ball arenas, {ticks} ticks, mutator {BALL_MUT}")
    balls = rga.workload_ball_states(11, 0, n, 1.0)
    ref = R.RefSim(n, 2)
    ref.set_env_car_count(np.zeros(n, np.int32))
    ref.set_mutator_floats(**BALL_MUT)
    ref.reset_to_kickoff(0)
    ref.upload(R.BALL, balls)
    sim = rga.BatchSim(n, 2)
    sim.set_env_car_count(np.zeros(n, np.int32))
    sim.set_mutator_config(**BALL_MUT)
    sim.reset_to_kickoff(0)
    sim.upload("ball_states", balls)
    assert np.array_equal(sim.download_state_array("env_states"), ref.download(R.ENV, n))
    for t in range(0, ticks, 100):
        sim.step(100)
        ref.step(100)
        assert np.array_equal(sim.download_state_array("ball_states"), ref.download(R.BALL, n)), f"tick {t + 100}"
    got = sim.get_mutator_config()[0]
    assert got["gravity_z"] == np.float32(-500.0) and got["ball_drag"] == np.float32(0.05)
    sim.close()
    ref.close()


def test_gpu_mutator_car_arenas_match_reference():
    torch = _torch()
    R = oracle_or_skip()
    import rocketgpu_amd as rga

    n, cars, seed = 48, 4, 21
    say(f"case: {n} arenas x {cars} cars, mutator {CAR_MUT}")
    counts = np.full(n, cars, np.int32)
    teams = (np.arange(n * cars) % cars >= cars // 2).astype(np.uint8)
    presets = (np.arange(n * cars) % 6).astype(np.uint8)
    ref = R.RefSim(n, cars, respawn_seed=seed)
    ref.set_env_car_count(counts)
    ref.set_teams(teams)
    ref.set_car_presets(presets)
    ref.set_mutator_floats(**CAR_MUT)
    ref.reset_to_kickoff(seed)
    sim = rga.BatchSim(n, cars)
    sim.set_env_car_count(counts)
    sim.set_teams(teams)
    sim.set_car_configs(ref.download_car_configs(n * cars))
    sim.set_mutator_config(**CAR_MUT)
    sim.reset_to_kickoff(seed)
    got, unsupported = batch_records(sim)
    bad = arena_mismatches(reference_records(R, ref, n, cars), got, n, cars)
    assert bad.size == 0, f"kickoff under the mutator: arenas {bad.tolist()} differ"
    ctrl = rga.workload_controls(seed, 0, n, cars, 240)
    dctrl = torch.from_numpy(ctrl.reshape(240, -1)).cuda()
    flagged = unsupported != 0
    for k in range(30):
        rows = ctrl[8 * k:8 * k + 8].reshape(8, -1)
        sim.step_with_device_controls(dctrl.data_ptr() + 8 * k * dctrl.shape[1], 8, n * cars)
        ref.step_with_controls(rows, 8, n * cars)
        got, unsupported = batch_records(sim)
        flagged |= unsupported != 0
        bad = arena_mismatches(reference_records(R, ref, n, cars), got, n, cars)
        assert not np.isin(bad, np.nonzero(~flagged)[0]).any(), f"tick {8 * (k + 1)}: unflagged arenas differ"
    say(f"  {int((~flagged).sum())} of {n} arenas unflagged after 240 ticks")
    assert (~flagged).sum() >= n // 4
    sim.close()
    ref.close()


def test_gpu_mutator_refuses_fixed_fields():
    _torch()
    import rocketgpu_amd as rga

    sim = rga.BatchSim(4, 2)
    for field, value in (("ball_radius", 100.0), ("sticky_force", 300.0), ("game_mode", 1), ("demo_respawn_time", 2.0)):
        with pytest.raises(rga.RocketGpuError, match=field):
            sim.set_mutator_config(**{field: value})
    with pytest.raises(rga.RocketGpuError, match="ball_drag"):
        sim.set_mutator_config(ball_drag=1.5)
    sim.set_mutator_config(kickoff_settled=0)
    assert sim.get_mutator_config()[0]["kickoff_settled"] == 0
    sim.close()
