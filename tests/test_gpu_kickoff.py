"""Kickoff resets and env commands of car arenas on the GPU against the reference CPU build.

reset_to_kickoff with cars places every car at its kickoff spot (the spot order drawn from the seed) and, with
settled kickoffs on (the default), replaces it by the settled row of its (CarConfig, team) group, which this
library records by running a countdown kickoff on the GPU.  apply_env_commands with cars covers countdown resets
(cars re-created at their spot on the first tick), settled resets, follow-up kickoffs that keep the score (the
stored spot order rotated, the short team's last car moved), kept pads and new rounds.  Every record must equal
the reference's byte for byte, and so must the arenas after stepping on from there while no unsupported body
contact has been met.
"""
import sys

import numpy as np
import pytest

from car_cases import arena_mismatches, batch_records, reference_at_kickoff, reference_records
from conftest import oracle_or_skip

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _torch():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _batch(n_env, cars, counts, teams, cfgs):
    import rocketgpu_amd as rga

    sim = rga.BatchSim(n_env, cars)
    sim.set_env_car_count(counts)
    sim.set_teams(teams)
    sim.set_car_configs(cfgs)
    return sim


def _same(ref_rec, got, n_env, cars, what):
    bad = arena_mismatches(ref_rec, got, n_env, cars)
    assert bad.size == 0, f"{what}: arenas {bad[:10].tolist()} differ from the reference"


@pytest.mark.parametrize("n_env,cars,seed,mixed", [(96, 2, 7, False), (24, 6, 11, True), (16, 3, 0, False),
                                                    (12, 5, 123456789, True)])
def test_gpu_car_kickoff_matches_reference(n_env, cars, seed, mixed):
    _torch()
    R = oracle_or_skip()
    say(f"case: kickoff of {n_env} arenas x {cars} cars, seed {seed}, mixed bodies {mixed}")
    presets = np.arange(n_env * cars) % 6 if mixed else None
    ref, counts, teams, cfgs = reference_at_kickoff(R, n_env, cars, seed, presets)
    want = reference_records(R, ref, n_env, cars)
    sim = _batch(n_env, cars, counts, teams, cfgs)
    sim.reset_to_kickoff(seed)
    got, unsupported = batch_records(sim)
    _same(want, got, n_env, cars, "after reset_to_kickoff")
    assert not unsupported.any()
    ctrl = sim.download_state_array("controls")
    assert not ctrl.any()
    # Stepping on from the kickoff (held zero controls) stays equal while nothing unsupported is met.
    for t in range(4):
        sim.step(30)
        ref.step(30)
        got, unsupported = batch_records(sim)
        ok = unsupported == 0
        bad = arena_mismatches(reference_records(R, ref, n_env, cars), got, n_env, cars)
        assert not np.isin(bad, np.nonzero(ok)[0]).any(), f"tick {30 * (t + 1)}: unflagged arenas differ"
    say(f"  {int((unsupported == 0).sum())} of {n_env} arenas unflagged after 120 ticks")
    sim.close()
    ref.close()


def _commands(rga, n_env, round_):
    c = np.zeros(n_env, rga.ENV_COMMAND_DTYPE)
    for e in range(n_env):
        kind = (e + round_) % 6
        if kind == 0:      # countdown reset, new match
            c[e]["reset_to_kickoff"], c[e]["reset_scoreboard"] = 1, 1
            c[e]["kickoff_countdown_ticks"] = 0xFFFF
            c[e]["kickoff_seed"] = 1000 + e
        elif kind == 1:    # settled reset keeping the score: a follow-up kickoff (seed 0)
            c[e]["reset_to_kickoff"] = 1
        elif kind == 2:    # countdown reset keeping score and pads
            c[e]["reset_to_kickoff"] = 1
            c[e]["keep_boost_pads"] = 1
            c[e]["kickoff_countdown_ticks"] = 0xFFFF
        elif kind == 3:    # settled reset, seed mode 3 (spots in order)
            c[e]["reset_to_kickoff"], c[e]["reset_scoreboard"] = 1, 1
            c[e]["kickoff_seed_mode"] = 3
        elif kind == 4:    # new round of the bodies only
            c[e]["new_round_bodies"] = 1
            c[e]["reset_tick_count"] = 1
        # kind 5: no command
    return c


@pytest.mark.parametrize("n_env,cars,mixed", [(36, 4, False), (30, 3, True)])
def test_gpu_car_env_commands_match_reference(n_env, cars, mixed):
    _torch()
    R = oracle_or_skip()
    import rocketgpu_amd as rga

    say(f"case: env commands on {n_env} arenas x {cars} cars, mixed bodies {mixed}")
    presets = np.arange(n_env * cars) % 6 if mixed else None
    ref, counts, teams, cfgs = reference_at_kickoff(R, n_env, cars, 5, presets)
    sim = _batch(n_env, cars, counts, teams, cfgs)
    sim.reset_to_kickoff(5)
    flagged = np.zeros(n_env, bool)
    for round_ in range(3):
        cmds = _commands(rga, n_env, round_)
        sim.apply_env_commands(cmds)
        ref.apply_env_commands(cmds.view(np.uint8).reshape(n_env, -1))
        got, unsupported = batch_records(sim)
        flagged |= unsupported != 0
        bad = arena_mismatches(reference_records(R, ref, n_env, cars), got, n_env, cars)
        assert not np.isin(bad, np.nonzero(~flagged)[0]).any(), f"round {round_}: arenas {bad.tolist()} differ"
        for _ in range(5):
            sim.step(40)
            ref.step(40)
            got, unsupported = batch_records(sim)
            flagged |= unsupported != 0
            bad = arena_mismatches(reference_records(R, ref, n_env, cars), got, n_env, cars)
            assert not np.isin(bad, np.nonzero(~flagged)[0]).any(), f"round {round_}: stepped arenas differ"
        say(f"  round {round_}: {int((~flagged).sum())} of {n_env} arenas unflagged")
    assert (~flagged).sum() >= n_env // 2
    sim.close()
    ref.close()


@pytest.mark.parametrize("stage,per_wave", [(0, 32), (0, 8), (0, 4), (2, 8), (2, 4), (4, 32), (4, 16), (4, 8)])
def test_gpu_car_arenas_per_wave_agree(stage, per_wave):
    """The car kernel's launch knobs (arenas per wavefront; per-tick data in lane-private memory or LDS; 2 or 4 lanes
    per arena sharing its wheel rays at stage 4) change nothing in the results."""
    torch = _torch()
    import rocketgpu_amd as rga

    say(f"case: 96 mixed arenas, arenas per wave 64 vs stage {stage}, {per_wave} per wave")
    n = 96
    counts = np.array([2, 4, 6] * (n // 3), np.int32)
    presets = np.arange(n * 6) % 6
    teams = (np.arange(n * 6) % 6 >= np.repeat(counts, 6) // 2).astype(np.uint8)
    cfgs = np.stack([rga.car_preset_config(int(p)) for p in presets])
    ctrl = torch.from_numpy(rga.workload_controls(9, 0, n, 6, 64).reshape(64, -1)).cuda()
    out = []
    for st, pw in ((0, 64), (stage, per_wave)):
        sim = _batch(n, 6, counts, teams, cfgs)
        sim.set_launch_shape(64, st, pw)
        sim.reset_to_kickoff(9)
        for k in range(8):
            sim.step_with_device_controls(ctrl.data_ptr() + k * 8 * ctrl.shape[1], 8, n * 6)
        out.append(batch_records(sim))
        sim.close()
    (a, ua), (b, ub) = out
    for key in a:
        assert np.array_equal(a[key], b[key]), key
    assert np.array_equal(ua, ub)


def _mixed_batch(n, seed):
    import rocketgpu_amd as rga

    counts = np.array([2, 4, 6, 3] * (n // 4), np.int32)
    presets = np.arange(n * 6) % 6
    teams = (np.arange(n * 6) % 6 >= np.repeat(counts, 6) // 2).astype(np.uint8)
    cfgs = np.stack([rga.car_preset_config(int(p)) for p in presets])
    sim = _batch(n, 6, counts, teams, cfgs)
    sim.reset_to_kickoff(seed)
    return sim


def test_gpu_copy_state_snapshot_and_async_download_with_cars():
    """copy_state_from, a snapshot save/load and an asynchronous download each carry a car batch (wheel runtime
    and internal state included): the copies then step exactly as the original."""
    torch = _torch()
    import rocketgpu_amd as rga

    n = 64
    say(f"case: {n} mixed arenas, state copy / snapshot / async download")
    ctrl = torch.from_numpy(rga.workload_controls(3, 0, n, 6, 64).reshape(64, -1)).cuda()
    a = _mixed_batch(n, 3)
    for k in range(4):
        a.step_with_device_controls(ctrl.data_ptr() + k * 8 * ctrl.shape[1], 8, n * 6)
    b = _mixed_batch(n, 99)
    b.copy_state_from(a)
    c = _mixed_batch(n, 98)
    c.load_state(a.save_state())
    for k in range(4, 8):
        for s in (a, b, c):
            s.step_with_device_controls(ctrl.data_ptr() + k * 8 * ctrl.shape[1], 8, n * 6)
    ra, ua = batch_records(a)
    for other in (b, c):
        ro, uo = batch_records(other)
        for key in ra:
            assert np.array_equal(ra[key], ro[key]), key
        assert np.array_equal(ua, uo)
    stream = torch.cuda.Stream()
    out = torch.empty((n * 6, 880), dtype=torch.uint8, pin_memory=True).numpy()
    a.download_state_array_async("car_states", out, stream=stream)
    stream.synchronize()
    assert np.array_equal(out, ra["cars"])
    other = rga.BatchSim(n + 1, 6)
    with pytest.raises(rga.RocketGpuError, match="same shape"):
        other.copy_state_from(a)
    for s in (a, b, c, other):
        s.close()
