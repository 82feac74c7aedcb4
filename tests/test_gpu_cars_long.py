"""A long mixed run on the GPU against the oracle: 4,096 arenas of 1v1 / 2v2 / 3v3 (55 / 27 / 18 %, stock bodies),
10,000 ticks from the kickoff under seeded held controls.

tests/golden/cars_mixed_4096_t10000.npz (written by tests/golden/make_car_fixtures.py with the reference CPU build)
holds, per arena, a running FNV-1a chain over its body digest after every tick, at every 1,000 ticks.  The GPU run
(one tick per launch through step_with_device_controls, two lanes per arena) keeps the same chains from the
device's per-arena digests.
At each checkpoint, every arena whose unsupported word has stayed 0 up to it must have the oracle's chain: all its
bodies' positions, velocities, angular velocities and orientations equal, bit for bit, on every tick so far.
"""
import os
import sys

import numpy as np
import pytest

from digests import FNV_OFFSET, arena_body_digests, chain_step

pytestmark = pytest.mark.gpu

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cars_mixed_4096_t10000.npz")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_gpu_cars_mixed_4096_10000_ticks_match_oracle_chains():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import rocketgpu_amd as rga

    fx = np.load(FIXTURE)
    counts, presets, teams = fx["counts"], fx["presets"], fx["teams"]
    seed, period, every = int(fx["seed"]), int(fx["period"]), int(fx["every"])
    ref_chain = fx["chain"]
    n = counts.size
    ticks = every * (ref_chain.shape[0] - 1)
    say(f"case: {n} mixed arenas, {ticks} ticks, checkpoints every {every}")
    table = np.stack([rga.car_preset_config(i) for i in range(6)])
    sim = rga.BatchSim(n, 6)
    sim.set_env_car_count(counts)
    sim.set_teams(teams)
    sim.set_car_configs(table[presets])
    # The headline launch shape: two lanes per arena, which split each car's wheel-ray tree walks.
    sim.set_launch_shape(64, 4, 32)
    sim.reset_to_kickoff(seed)
    controls = rga.workload_controls(seed, 0, n, 6, period)
    dev = torch.from_numpy(controls.reshape(period, -1)).to("cuda")
    row_bytes = dev.shape[1]
    digests = torch.zeros(n, dtype=torch.int64, device="cuda")

    def device_digests():
        sim.env_digests_into(digests.data_ptr())
        torch.cuda.synchronize()
        return digests.cpu().numpy().view(np.uint64)

    # The host digest of the downloaded records equals the device's (the chain below is the fixture's).
    d0 = device_digests()
    host = arena_body_digests(sim.download_state_array("ball_states"), sim.download_state_array("car_states"),
                              counts, 6)
    assert np.array_equal(d0, host)
    chain = chain_step(np.full(n, FNV_OFFSET, dtype=np.uint64), d0)
    assert np.array_equal(chain, ref_chain[0]), "kickoff differs from the oracle's"
    flagged = np.zeros(n, bool)
    compared = 0
    for t in range(ticks):
        sim.step_with_device_controls(dev.data_ptr() + (t % period) * row_bytes, 1, n * 6)
        chain = chain_step(chain, device_digests())
        flagged |= sim.download_state_array("unsupported").view(np.uint32).ravel() != 0
        if (t + 1) % every == 0:
            k = (t + 1) // every
            live = ~flagged
            bad = np.nonzero(live & (chain != ref_chain[k]))[0]
            assert bad.size == 0, f"tick {t + 1}: unflagged arenas {bad[:10]} differ from the oracle"
            compared += int(live.sum())
            say(f"  tick {t + 1}: {int(live.sum())} of {n} arenas unflagged, all equal to the oracle")
    say(f"  arenas never flagged: {int((~flagged).sum())} of {n}")
    # Most arenas run the whole 10,000 ticks without meeting physics this build does not step.
    assert (~flagged).sum() >= 0.8 * n
    sim.close()
