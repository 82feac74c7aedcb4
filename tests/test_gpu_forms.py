"""The batched and served forms of the step on the GPU: arena subsets with per-arena horizons, hipGraph replays and
the serving loop, each checked bit-exact against the reference CPU build's fixtures or against the plain step."""
import numpy as np
import pytest

from conftest import load_fixture
from digests import ball_env_digests, fold

pytestmark = pytest.mark.gpu


def _torch():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def make_sim(f, n=None):
    import rocketgpu_amd as rga

    m = f["meta"]
    n = n or m["arenas"]
    sim = rga.BatchSim(n, 2, m["scale"], m["fillet_radius"])
    sim.set_env_car_count(np.zeros(n, np.int32))
    sim.upload("env_states", f["init_env"][:n])
    sim.upload("ball_states", f["init_ball"][:n])
    return sim


def test_step_arenas_per_arena_horizons():
    """Arena i stepped (i % 3 + 1) * 100 ticks in one launch equals the reference at that tick."""
    torch = _torch()
    f = load_fixture("ball_arena_96_t1000")
    n = f["meta"]["arenas"]
    sim = make_sim(f)
    ids = torch.arange(n, dtype=torch.int32, device="cuda").flip(0).contiguous()
    horizon = ((ids.cpu().numpy() % 3) + 1) * 100
    ticks = torch.from_numpy(horizon.astype(np.int32)).cuda()
    sim.step_arenas(ids.data_ptr(), n, 0, ticks.data_ptr())
    sim.synchronize()
    got = ball_env_digests(sim.download_state_array("ball_states"))
    want = np.array([f["env_digest"][(e % 3), e] for e in range(n)], np.uint64)
    assert np.array_equal(got, want)
    sim.close()


def test_step_arenas_subset_leaves_others_untouched():
    torch = _torch()
    f = load_fixture("ball_arena_small_scale_64_t1500")
    n = f["meta"]["arenas"]
    sim = make_sim(f)
    before = sim.download_state_array("ball_states")
    sub = np.arange(0, n, 2, dtype=np.int32)
    ids = torch.from_numpy(sub).cuda()
    for _ in range(5):
        sim.step_arenas(ids.data_ptr(), sub.size, 100)
    sim.synchronize()
    after = sim.download_state_array("ball_states")
    odd = np.arange(1, n, 2)
    assert np.array_equal(after[odd], before[odd])
    assert np.array_equal(ball_env_digests(after[sub]), f["env_digest"][0][sub])
    sim.close()


@pytest.mark.parametrize("ticks_per_launch,launches", [(4, 5), (1, 20), (20, 1)])
def test_graph_replays_match_reference(ticks_per_launch, launches):
    """Replayed several times, with eager work and synchronizations between replays; every replay is checked."""
    torch = _torch()
    f = load_fixture("ball_arena_96_t1000")
    sim = make_sim(f)
    s = torch.cuda.Stream()
    per_replay = ticks_per_launch * launches
    for r in range(5):
        sim.step_graph(ticks_per_launch, launches, stream=s)
        s.synchronize()
        balls = sim.download_state_array("ball_states")
        assert fold(ball_env_digests(balls)) == f["batch_digest"][(r + 1) * per_replay - 1], f"replay {r}"
        torch.zeros(1024, device="cuda").sum().item()  # eager work between replays
    sim.close()


@pytest.mark.parametrize("block,stage,per_wave", [(64, 0, 64), (128, 0, 64), (256, 0, 32), (64, 1, 64),
                                                  (128, 1, 16), (256, 1, 8), (64, 0, 4), (64, 1, 8),
                                                  (64, 2, 16), (128, 2, 8), (64, 3, 8), (128, 3, 4), (64, 3, 4),
                                                  (64, 4, 16), (64, 4, 8), (128, 4, 8), (256, 4, 4),
                                                  (64, 5, 16), (64, 5, 8), (128, 5, 8), (256, 5, 4)])
def test_launch_shapes_match_reference(block, stage, per_wave):
    """Every launch shape (lanes per workgroup, state staging, arenas per wave) gives the reference's result."""
    _torch()
    f = load_fixture("ball_arena_1024_t3000")
    sim = make_sim(f)
    sim.set_launch_shape(block, stage, per_wave)
    for _ in range(6):
        sim.step(500)
    balls = sim.download_state_array("ball_states")
    envs = sim.download_state_array("env_states")
    sim.close()
    assert np.array_equal(balls, f["final_ball"])
    assert np.array_equal(envs, f["final_env"])


@pytest.mark.parametrize("block,stage", [(64, 6), (128, 6), (256, 6), (64, 7), (256, 7)])
def test_split_launches_match_reference(block, stage):
    """Stages 6 and 7 sort the arenas into heavy and light lists before every launch (by the windows their last
    contact walk loaded) and give the heavy ones 32 or 16 lanes: 375 launches of 8 ticks give the reference's result,
    and so does a subset launch (which runs as stage 5)."""
    _torch()
    f = load_fixture("ball_arena_1024_t3000")
    sim = make_sim(f)
    sim.set_launch_shape(block, stage, 8)
    for _ in range(375):
        sim.step(8)
    balls = sim.download_state_array("ball_states")
    envs = sim.download_state_array("env_states")
    sim.close()
    assert np.array_equal(balls, f["final_ball"])
    assert np.array_equal(envs, f["final_env"])


def test_split_stage_rejects_other_group_sizes():
    import rocketgpu_amd as rga
    f = load_fixture("ball_arena_96_t1000")
    sim = make_sim(f)
    with pytest.raises(rga.RocketGpuError, match="stages 6 and 7 need 8 arenas per wave"):
        sim.set_launch_shape(64, 6, 4)
    with pytest.raises(rga.RocketGpuError, match="stage must be 0 to 7"):
        sim.set_launch_shape(64, 8, 8)
    sim.close()


def test_serving_results_same_alone_and_in_the_mix():
    """Each client's arenas after the served mix equal the same arenas stepped alone by the same tick count."""
    torch = _torch()
    import rocketgpu_amd as rga

    n = 4096
    balls = rga.workload_ball_states(77, 0, n, 1.0)
    sim = rga.BatchSim(n, 2)
    sim.set_env_car_count(np.zeros(n, np.int32))
    env0 = sim.download_state_array("env_states")
    sim.upload("ball_states", balls)
    rep, per_arena = sim.serve(clients=32, min_arenas=16, max_arenas=96, horizons=(1, 4, 8),
                               max_arenas_per_launch=1024, max_in_flight=2, request_rate_hz=800.0, duration_s=0.5,
                               seed=5, ticks_per_arena=True)
    assert rep["requests"] > 100 and rep["p99_ms"] >= rep["p50_ms"] > 0.0
    assert rep["arena_ticks"] == int(per_arena.sum())
    served = sim.download_state_array("ball_states")
    sim.close()

    alone = rga.BatchSim(n, 2)
    alone.set_env_car_count(np.zeros(n, np.int32))
    alone.upload("env_states", env0)
    alone.upload("ball_states", balls)
    ids = torch.arange(n, dtype=torch.int32, device="cuda")
    ticks = torch.from_numpy(per_arena).cuda()
    alone.step_arenas(ids.data_ptr(), n, 0, ticks.data_ptr())
    alone.synchronize()
    assert np.array_equal(alone.download_state_array("ball_states"), served)
    alone.close()


def test_launch_shape_rejects_oversized_lds():
    _torch()
    import rocketgpu_amd as rga

    sim = rga.BatchSim(64, 2)
    with pytest.raises(rga.RocketGpuError, match="LDS"):
        sim.set_launch_shape(256, 3, 64)
    with pytest.raises(rga.RocketGpuError):
        sim.set_launch_shape(64, 5, 64)
    sim.step(2)  # the previous shape still works
    sim.close()


def test_serving_rejects_bad_configs():
    _torch()
    import rocketgpu_amd as rga

    sim = rga.BatchSim(256, 2)
    with pytest.raises(rga.RocketGpuError, match="more arenas"):
        sim.serve(clients=64, min_arenas=16, max_arenas=16, duration_s=0.1)
    with pytest.raises(rga.RocketGpuError, match="budget|below"):
        sim.serve(clients=4, min_arenas=16, max_arenas=32, max_arenas_per_launch=8, duration_s=0.1)
    sim.close()
