"""The HIP path on the GPU against the reference CPU build: bit-exact, every arena, every tick.

Each fixture (tests/golden/*.npz, written by the oracle) holds the uploaded states, the per-tick batch digest, per-env
digests at checkpoints and every byte of the final ball and env states.  The HIP path runs through the C-ABI
(rocketgpu_amd -> librocketgpu_amd.so), one step(1) per tick, with the per-env digests computed on the device after
each tick and compared on the host at the end.
"""
import os
import sys

import numpy as np
import pytest

from conftest import fixture_names, load_fixture
from digests import ball_env_digests, fold

pytestmark = pytest.mark.gpu


def _torch():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def make_sim(f, **kw):
    import rocketgpu_amd as rga

    m = f["meta"]
    n = m["arenas"]
    sim = rga.BatchSim(n, kw.pop("max_cars_per_env", 2), m["scale"], m["fillet_radius"], **kw)
    sim.set_env_car_count(np.zeros(n, np.int32))
    sim.upload("env_states", f["init_env"])
    sim.upload("ball_states", f["init_ball"])
    return sim


def run_ticks(torch, sim, n, ticks, step=None):
    """Step tick by tick; per-env digests of every tick on the device ([ticks, n] uint64 on the host)."""
    buf = torch.zeros((ticks, n), dtype=torch.int64, device="cuda")
    for t in range(ticks):
        (step or sim.step)(1)
        sim.env_digests_into(buf[t].data_ptr())
    torch.cuda.synchronize()
    return buf.cpu().numpy().view(np.uint64)


@pytest.mark.parametrize("name", fixture_names())
def test_gpu_matches_reference_every_tick(name):
    torch = _torch()
    f = load_fixture(name)
    m = f["meta"]
    n, ticks, every = m["arenas"], m["ticks"], int(f["checkpoint_every"])
    say(f"[gpu parity] {name}: {n} arenas x {ticks} ticks")
    sim = make_sim(f)
    d = run_ticks(torch, sim, n, ticks)
    batch = fold(d)
    bad = np.flatnonzero(batch != f["batch_digest"])
    if bad.size:
        t = int(bad[0])
        arenas = np.flatnonzero(d[t] != f["env_digest"][t // every]) if (t + 1) % every == 0 else []
        pytest.fail(f"first differing tick {t} of {ticks} ({bad.size} ticks differ); arenas {list(arenas)[:8]}")
    assert np.array_equal(d[every - 1::every], f["env_digest"])
    balls = sim.download_state_array("ball_states")
    envs = sim.download_state_array("env_states")
    diff = np.flatnonzero((balls != f["final_ball"]).any(axis=1))
    assert diff.size == 0, f"final BallState bytes differ in arenas {list(diff[:8])}"
    assert np.array_equal(envs, f["final_env"]), "final EnvState bytes differ"
    assert np.array_equal(ball_env_digests(balls), f["env_digest"][-1])
    sim.close()


def test_gpu_multi_tick_launch_equals_single_ticks():
    """step(k) keeps the arena state across ticks inside one launch; it must equal k launches of step(1)."""
    _torch()
    f = load_fixture("ball_arena_96_t1000")
    finals = []
    for chunks in ([1] * 240, [240], [8] * 30, [3, 200, 37]):
        sim = make_sim(f)
        for c in chunks:
            sim.step(c)
        finals.append(sim.download_state_array("ball_states"))
        sim.close()
    for x in finals[1:]:
        assert np.array_equal(finals[0], x)
    ticks_at = 239
    assert np.array_equal(fold(ball_env_digests(finals[0])), f["batch_digest"][ticks_at])


def test_gpu_device_controls_entry_point():
    """step_with_device_controls on ball-only arenas reads no car control and matches step()."""
    torch = _torch()
    f = load_fixture("ball_arena_small_scale_64_t1500")
    n = f["meta"]["arenas"]
    sim = make_sim(f)
    ticks = 16
    ctrl = torch.randint(0, 255, (ticks, sim.car_capacity * 24), dtype=torch.uint8, device="cuda")
    for _ in range(4):
        sim.step_with_device_controls(ctrl.data_ptr(), ticks, sim.car_capacity)
    sim.synchronize()
    got = sim.download_state_array("ball_states")
    assert np.array_equal(fold(ball_env_digests(got)), f["batch_digest"][4 * ticks - 1])
    import rocketgpu_amd as rga

    with pytest.raises(rga.RocketGpuError, match="control_stride"):
        sim.step_with_device_controls(ctrl.data_ptr(), ticks, sim.car_capacity - 1)
    with pytest.raises(rga.RocketGpuError, match="null"):
        sim.step_with_device_controls(0, ticks, sim.car_capacity)
    sim.close()
    del n


def test_gpu_stream_argument():
    """A caller stream (torch side stream) orders the step, digest and download like the null stream."""
    torch = _torch()
    f = load_fixture("ball_arena_96_t1000")
    n = f["meta"]["arenas"]
    sim = make_sim(f)
    s = torch.cuda.Stream()
    buf = torch.zeros(n, dtype=torch.int64, device="cuda")
    with torch.cuda.stream(s):
        sim.step(100, stream=s)
        sim.env_digests_into(buf.data_ptr(), stream=s)
    s.synchronize()
    assert np.array_equal(buf.cpu().numpy().view(np.uint64), f["env_digest"][0])
    sim.close()


def test_gpu_error_paths():
    _torch()
    import rocketgpu_amd as rga

    f = load_fixture("ball_arena_96_t1000")
    n = f["meta"]["arenas"]
    sim = make_sim(f)
    with pytest.raises(rga.RocketGpuError, match="outside"):
        sim.set_env_car_count(np.full(n, 3, np.int32))
    with pytest.raises(rga.RocketGpuError):
        sim.set_env_car_count(np.zeros(n - 1, np.int32))
    with pytest.raises(rga.RocketGpuError, match="count"):
        sim.upload("ball_states", f["init_ball"][:-1])
    with pytest.raises(rga.RocketGpuError):
        sim.step(-1)
    with pytest.raises(rga.RocketGpuError):
        rga.BatchSim(0)
    # A failed call leaves the batch usable.
    sim.step(1)
    balls = sim.download_state_array("ball_states")
    assert fold(ball_env_digests(balls)) == f["batch_digest"][0]
    sim.close()


def test_gpu_full_size_against_reference_build():
    """16,384 ball arenas (the benchmark's batch) against the reference CPU build run beside it, when that build is
    present; 240 ticks, per-tick batch digests."""
    torch = _torch()
    from oracle import refsim as R

    if not R.available():
        pytest.skip("reference CPU build (oracle/_ref) not present")
    import rocketgpu_amd as rga

    n, ticks, seed = 16384, 240, 2024
    say(f"[gpu parity] full size: {n} arenas x {ticks} ticks vs reference build")
    init_ball = rga.workload_ball_states(seed, 0, n, 1.0)
    init_env = np.tile(R.default_state(R.ENV), (n, 1))
    ref = R.RefSim(n, 2, 1.0, 256.0)
    ref.set_env_car_count(np.zeros(n, np.int32))
    ref.upload(R.ENV, init_env)
    ref.upload(R.BALL, init_ball)
    rbatch, _ = ref.run(np.zeros(n, np.int32), 0, ticks, 0)
    rball = ref.download(R.BALL, n)
    sim = rga.BatchSim(n, 2)
    sim.set_env_car_count(np.zeros(n, np.int32))
    sim.upload("env_states", init_env)
    sim.upload("ball_states", init_ball)
    d = run_ticks(torch, sim, n, ticks)
    bad = np.flatnonzero(fold(d) != rbatch)
    assert bad.size == 0, f"first differing tick {bad[0]}"
    assert np.array_equal(sim.download_state_array("ball_states"), rball)
    sim.close()
