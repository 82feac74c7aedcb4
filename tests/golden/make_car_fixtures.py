"""Writes the long-run car parity fixture by running the oracle (the reference CPU build, oracle/_ref).

Runs only where the reference sources were built (this container): `python tests/golden/make_car_fixtures.py`.

cars_mixed_4096_t10000.npz holds data only:

  counts [N] i32, presets [N * 6] u8, teams [N * 6] u8   the arenas: 1v1 / 2v2 / 3v3 at 55 / 27 / 18 %, a uniform
                                                          stock body (0-5) per car slot, first half Blue
  seed, period, every                                     kickoff seed (reset_to_kickoff and the respawn draws), the
                                                          control period (controls of tick t are row t % period of
                                                          rocketgpu_amd.workload_controls(seed, 0, N, 6, period)),
                                                          the checkpoint interval
  chain [K + 1, N] u64                                    per arena, a running FNV-1a 64 over the little-endian bytes
                                                          of its body digest after every tick; row 0 after the kickoff,
                                                          row k after tick k * every
  meta (json)

Body digest of an arena (rocketgpu_amd/workload.h BodyDigest): FNV-1a 64 over the raw bits of pos, vel, ang_vel,
rot (w, x, y, z) of the ball, then of each of its cars in slot order.  The chain starts at the FNV offset basis.
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
sys.path.insert(0, os.path.dirname(HERE))

from oracle import refsim as R  # noqa: E402
from digests import arena_body_digests, chain_step, FNV_OFFSET  # noqa: E402

N, TICKS, EVERY, SEED, PERIOD = 4096, 10000, 1000, 2024, 1000
MIX = (0.55, 0.27, 0.18)


def mixed_arenas(n, seed):
    rng = np.random.default_rng(seed)
    u = rng.random(n)
    counts = np.where(u < MIX[0], 2, np.where(u < MIX[0] + MIX[1], 4, 6)).astype(np.int32)
    presets = rng.integers(0, 6, n * 6).astype(np.uint8)
    slot = np.arange(n * 6) % 6
    teams = (slot >= np.repeat(counts, 6) // 2).astype(np.uint8)
    return counts, presets, teams


def main():
    import rocketgpu_amd as rga

    counts, presets, teams = mixed_arenas(N, SEED)
    controls = rga.workload_controls(SEED, 0, N, 6, PERIOD)
    ref = R.RefSim(N, 6, respawn_seed=SEED)
    ref.set_env_car_count(counts)
    ref.set_teams(teams)
    ref.set_car_presets(presets)
    ref.reset_to_kickoff(SEED)
    chain = np.full(N, FNV_OFFSET, dtype=np.uint64)

    def digest():
        return arena_body_digests(ref.download(R.BALL, N), ref.download(R.CAR, N * 6), counts, 6)

    chain = chain_step(chain, digest())
    rows = [chain.copy()]
    t0 = time.time()
    for t in range(TICKS):
        ref.step_with_controls(np.ascontiguousarray(controls[t % PERIOD]), 1, N * 6)
        chain = chain_step(chain, digest())
        if (t + 1) % EVERY == 0:
            rows.append(chain.copy())
            print(f"tick {t + 1}: {time.time() - t0:.0f} s", flush=True)
    ref.close()
    meta = {"workload": f"{N} mixed arenas (1v1/2v2/3v3 at 55/27/18 %, stock bodies 0-5), procedural arena "
                        "PA(1, 256), kickoff, seeded held controls with period 1000", "arenas": N, "ticks": TICKS}
    np.savez_compressed(os.path.join(HERE, f"cars_mixed_{N}_t{TICKS}.npz"), counts=counts, presets=presets,
                        teams=teams, seed=np.uint64(SEED), period=np.int32(PERIOD), every=np.int32(EVERY),
                        chain=np.stack(rows), meta=np.frombuffer(json.dumps(meta).encode(), np.uint8))


if __name__ == "__main__":
    main()
