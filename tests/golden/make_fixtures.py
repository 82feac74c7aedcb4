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


def main(names):
    for name in names or list(FIXTURES) + list(BALL_PATH_FIXTURES):
        if name in FIXTURES:
            make(name, *FIXTURES[name])
        else:
            make_ball_path(name, *BALL_PATH_FIXTURES[name])


if __name__ == "__main__":
    main(sys.argv[1:])
