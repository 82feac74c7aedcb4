"""The serving loop over arenas with cars: byte-budget admission, per-request held controls, periodic kickoff
resets.  Two checks per served client:
- alone vs in the mix: the client's requests replayed alone on a fresh batch (its resets, controls and steps in
  order) give its arenas' records byte for byte;
- against the reference CPU build: the same requests replayed on a reference batch of the same shape give the
  client's arenas' records byte for byte while this build's unsupported word for the arena is 0.
"""
import sys

import numpy as np
import pytest

from car_cases import arena_mismatches
from conftest import oracle_or_skip

pytestmark = pytest.mark.gpu

N, CARS, SEED = 512, 6, 41


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _setup(rga):
    counts = np.array([2, 4, 6, 2, 2, 4, 2, 6] * (N // 8), np.int32)
    presets = (np.arange(N * CARS) * 7 % 6).astype(np.uint8)
    teams = (np.arange(N * CARS) % CARS >= np.repeat(counts, CARS) // 2).astype(np.uint8)
    cfgs = np.stack([rga.car_preset_config(int(p)) for p in range(6)])[presets]
    sim = rga.BatchSim(N, CARS)
    sim.set_env_car_count(counts)
    sim.set_teams(teams)
    sim.set_car_configs(cfgs)
    sim.reset_to_kickoff(SEED)
    return sim, counts, presets, teams


def _reset_commands(rga, first, count, n):
    c = np.zeros(n, rga.ENV_COMMAND_DTYPE)
    c["reset_to_kickoff"][first:first + count] = 1
    c["reset_scoreboard"][first:first + count] = 1
    c["kickoff_seed_mode"][first:first + count] = 3
    return c


def _records(get, first, count):
    sl = slice(first * CARS, (first + count) * CARS)
    return {"cars": get("car_states")[sl], "wheels": get("wheel_states")[sl],
            "balls": get("ball_states")[first:first + count], "envs": get("env_states")[first:first + count],
            "pads": get("boost_pad_states")[first * 34:(first + count) * 34]}


def test_gpu_serving_car_arenas_alone_in_mix_and_reference():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import rocketgpu_amd as rga

    kw = dict(clients=8, min_arenas=16, max_arenas=64, horizons=(1, 4, 8), max_arenas_per_launch=256,
              max_in_flight=2, request_rate_hz=300.0, duration_s=1.0, seed=SEED, reset_every=3,
              fresh_controls=True, p_boost=0.3)
    sim, counts, presets, teams = _setup(rga)
    budget = int(sum(rga.lib().rg_batch_arena_bytes(sim._h, int(c)) for c in counts[:96]))
    kw["launch_budget_bytes"] = budget
    rep, per_arena = sim.serve(ticks_per_arena=True, **kw)
    say(f"served: {rep}")
    assert rep["requests"] > 50 and rep["p99_ms"] >= rep["p50_ms"] > 0.0 and rep["resident_bytes"] > 0
    assert rep["arena_ticks"] == int(per_arena.sum())
    cfg = rga.BatchSim.serve_config(**kw)
    served_unsupported = sim.download_state_array("unsupported").view(np.uint32).ravel()

    R = oracle_or_skip()
    compared = 0
    for client in (0, 5):
        first, count, ticks = rga.serve_client_plan(cfg, client)
        say(f"client {client}: arenas {first}..{first + count}, {ticks.size} requests, {int(ticks.sum())} ticks")
        assert int(per_arena[first:first + count].min()) == int(ticks.sum())
        served = _records(sim.download_state_array, first, count)
        # Alone: a fresh batch, the client's requests one after another.
        alone, *_ = _setup(rga)
        ref = R.RefSim(N, CARS, respawn_seed=SEED)
        ref.set_env_car_count(counts)
        ref.set_teams(teams)
        ref.set_car_presets(presets)
        ref.reset_to_kickoff(SEED)
        ids = torch.arange(first, first + count, dtype=torch.int32, device="cuda")
        ctrl_all = np.zeros((N * CARS, 24), np.uint8)
        for j, k in enumerate(ticks.tolist()):
            if (j + 1) % kw["reset_every"] == 0:
                cmds = _reset_commands(rga, first, count, N)
                alone.apply_env_commands(cmds)
                ref.apply_env_commands(cmds.view(np.uint8).reshape(N, -1))
            rows = rga.request_controls(SEED, j, first, count, CARS, kw["p_boost"])
            ctrl_all[first * CARS:(first + count) * CARS] = rows
            alone.upload("controls", ctrl_all)
            ref.upload(R.CONTROLS, ctrl_all)
            alone.step_arenas(ids.data_ptr(), count, k)
            ref.step(k)
        got = _records(alone.download_state_array, first, count)
        for key in served:
            assert np.array_equal(served[key], got[key]), f"client {client}: {key} differ alone vs in the mix"
        want = {"cars": ref.download(R.CAR, N * CARS)[first * CARS:(first + count) * CARS],
                "wheels": ref.download_wheels(N * CARS)[first * CARS:(first + count) * CARS],
                "balls": ref.download(R.BALL, N)[first:first + count],
                "envs": ref.download(R.ENV, N)[first:first + count],
                "pads": ref.download(R.PAD, N * 34)[first * 34:(first + count) * 34]}
        bad = arena_mismatches(want, served, count, CARS)
        ok = served_unsupported[first:first + count] == 0
        assert not np.isin(bad, np.nonzero(ok)[0]).any(), f"client {client}: unflagged arenas differ"
        compared += int(ok.sum())
        say(f"  {int(ok.sum())} of {count} arenas unflagged and equal to the reference")
        alone.close()
        ref.close()
    assert compared > 0
    sim.close()
