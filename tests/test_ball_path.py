"""ball_path (include/rocketgpu/rocketgpu.h:1069-1073): every arena's ball stepped alone for samples x ticks ticks,
pos/vel sampled every `ticks` ticks, stepping stops when |y| reaches goal_y.  Checked bit-exact against the
reference CPU build's fixtures, on the host twin (CPU) and through the C-ABI on the GPU."""
import numpy as np
import pytest

from conftest import fixture_names, load_fixture, oracle_or_skip, twin

NAMES = fixture_names("ball_path_")


def test_ball_path_fixtures_present_and_consistent():
    assert len(NAMES) >= 2
    for name in NAMES:
        f = load_fixture(name)
        m = f["meta"]
        assert f["out"].shape == (m["arenas"], m["samples"], 6)
        res = f["result"]
        assert set(np.unique(res[:, 0])) <= {0, 1, 2}
        scored = res[:, 0] > 0
        assert scored.any(), "each fixture should contain arenas that reach a goal"
        # The goal side matches the sign of y, and a scored arena's samples stop changing after its goal tick.
        last = f["out"][:, -1, 1]
        assert np.all((last[scored] > 0) == (res[scored, 0] == 1))
        for e in np.flatnonzero(scored):
            s_goal = (res[e, 1] - 1) // m["ticks"]
            assert np.all(f["out"][e, s_goal:] == f["out"][e, s_goal])
            assert abs(f["out"][e, s_goal, 1]) >= m["goal_y"]


def test_oracle_reproduces_ball_path_fixture():
    R = oracle_or_skip()
    f = load_fixture("ball_path_small_scale_64_s25_t9")
    m = f["meta"]
    s = R.RefSim(m["arenas"], 2, m["scale"], m["fillet_radius"])
    s.set_env_car_count(np.zeros(m["arenas"], np.int32))
    s.upload(R.ENV, f["init_env"])
    s.upload(R.BALL, f["init_ball"])
    out, res = s.ball_path(m["samples"], m["ticks"], m["goal_y"])
    assert np.array_equal(out.view(np.uint32), f["out"].view(np.uint32))
    assert np.array_equal(res, f["result"])


@pytest.mark.parametrize("name", NAMES)
def test_twin_ball_path_matches_fixture(name):
    T = twin()
    f = load_fixture(name)
    m = f["meta"]
    n = m["arenas"]
    h = T.rgt_create(n, m["scale"], m["fillet_radius"])
    try:
        T.rgt_upload(h, 2, f["init_env"].ctypes.data, n)
        T.rgt_upload(h, 1, f["init_ball"].ctypes.data, n)
        out = np.zeros((n, m["samples"], 6), np.float32)
        res = np.zeros((n, 2), np.int32)
        T.rgt_ball_path(h, m["samples"], m["ticks"], m["goal_y"], out.ctypes.data, res.ctypes.data)
        assert np.array_equal(out.view(np.uint32), f["out"].view(np.uint32))
        assert np.array_equal(res, f["result"])
        balls = np.zeros((n, 288), np.uint8)
        T.rgt_download(h, 1, balls.ctypes.data, n)
        assert np.array_equal(balls, f["final_ball"])
    finally:
        T.rgt_destroy(h)


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_gpu_ball_path_matches_fixture(name):
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import rocketgpu_amd as rga

    f = load_fixture(name)
    m = f["meta"]
    n = m["arenas"]
    sim = rga.BatchSim(n, 2, m["scale"], m["fillet_radius"])
    sim.set_env_car_count(np.zeros(n, np.int32))
    sim.upload("env_states", f["init_env"])
    sim.upload("ball_states", f["init_ball"])
    out, res = sim.ball_path(m["samples"], m["ticks"], m["goal_y"])
    torch.cuda.synchronize()
    assert np.array_equal(out.cpu().numpy().view(np.uint32), f["out"].view(np.uint32))
    assert np.array_equal(res.cpu().numpy(), f["result"])
    assert np.array_equal(sim.download_state_array("ball_states"), f["final_ball"])
    assert np.array_equal(sim.download_state_array("env_states"), f["final_env"])
    # Null outputs are rejected before any launch; non-positive sizes are a no-op, as in the reference.
    assert rga.lib().rg_batch_ball_path(sim._h, 2, 2, 1.0, None, None, None) == 1
    assert "null" in rga.lib().rg_last_error().decode()
    assert rga.lib().rg_batch_ball_path(sim._h, 0, 2, 1.0, None, None, None) == 0
    sim.close()
