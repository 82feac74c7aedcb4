"""Env commands on ball arenas: reset_to_kickoff, apply_env_commands(_device) and step_with_device_commands
(include/rocketgpu/rocketgpu.h:1065-1083), with the kickoff reset of detail/step/arena_ball.hpp:2845-3083 and the
per-tick command path of detail/step/env_step.hpp:6700-6740.

Checked bit-exact, every byte of the balls, envs and pads after each phase, against the reference CPU build's fixture
(tests/golden/env_commands_48.npz): on the host twin (CPU) and through the C-ABI on the GPU.  A fresh batch starts as
the reference's constructor leaves it (a settled kickoff, src/rocketgpu.cpp:1281)."""
import sys

import numpy as np
import pytest

from conftest import GOLDEN, load_fixture, oracle_or_skip, twin

FIX = "env_commands_48"
BALL, ENV, PAD = 1, 2, 3
PADS_PER_ENV = 34


def _offsets(array):
    import rocketgpu_amd as rga

    return {name: off for name, _t, off, _s in rga.state_layouts()[array]["fields"]}


_B, _E = _offsets("ball_states"), _offsets("env_states")
BALL_POS, BALL_VEL, BALL_POS_BT, BALL_SLEEPING = _B["pos"], _B["vel"], _B["pos_bt"], _B["sleeping"]
ENV_KICKOFF_RESETS, ENV_MATCH_ORDER, ENV_RESPAWN_ORDER = (_E["kickoff_resets"], _E["kickoff_match_order"],
                                                          _E["respawn_match_order"])


def _f32(rows, off, n):
    return np.ascontiguousarray(rows[:, off:off + 4 * n]).view(np.float32)


def _phases(f):
    return f["meta"]["phases"]


def test_fixture_consistent():
    f = load_fixture(FIX)
    n = f["meta"]["arenas"]
    k = len(_phases(f))
    assert f["ball"].shape == (k, n, 288) and f["env"].shape == (k, n, 44)
    assert f["pads"].shape == (k, n * PADS_PER_ENV, 16)
    assert f["commands"].shape == (n, 16) and f["command_stream"].shape == (f["meta"]["stream_ticks"] * n, 16)
    # A fresh batch is at a settled kickoff: one ball for every env, at rest on the floor at the spawn.
    fresh = f["fresh_ball"]
    assert (fresh == fresh[0]).all()
    assert np.array_equal(_f32(fresh[:1], BALL_POS, 3)[0], np.float32([0.0, 0.0, 92.75]))
    assert np.array_equal(_f32(fresh[:1], BALL_VEL, 3)[0], np.zeros(3, np.float32))
    assert int(fresh[0, BALL_SLEEPING]) == 1
    # Fresh pads: active, no cooldown, the last six of the 34 big (the soccar table order).
    pads = f["fresh_pads"].reshape(n, PADS_PER_ENV, 16)
    assert (pads[:, :, 4] == 1).all() and (pads[:, :, 0:4] == 0).all()
    assert (pads[:, :, 5] == (np.arange(PADS_PER_ENV) >= 28)).all()


def _orders_ok(env_rows):
    """Every drawn kickoff order is a permutation of the 5 spots and every respawn order one of the 4 spots."""
    mo = np.ascontiguousarray(env_rows[:, ENV_MATCH_ORDER:ENV_MATCH_ORDER + 2]).view(np.uint16)[:, 0].astype(int)
    ro = env_rows[:, ENV_RESPAWN_ORDER].astype(int)
    drawn = (mo & 0x8000) != 0
    for m, r in zip(mo[drawn], ro[drawn]):
        assert sorted((m >> (3 * j)) & 7 for j in range(5)) == [0, 1, 2, 3, 4]
        assert sorted((r >> (2 * j)) & 3 for j in range(4)) == [0, 1, 2, 3]
    return int(drawn.sum())


def test_kickoff_and_respawn_orders_are_permutations():
    f = load_fixture(FIX)
    drawn = _orders_ok(f["fresh_env"]) + sum(_orders_ok(e) for e in f["env"])
    assert drawn > 100
    # Different seeds draw different orders across envs (the draw depends on the env index).
    orders = np.ascontiguousarray(f["fresh_env"][:, ENV_MATCH_ORDER:ENV_MATCH_ORDER + 2]).view(np.uint16)[:, 0]
    assert len(set(orders)) > 10


def test_oracle_reproduces_fixture():
    oracle_or_skip()
    sys.path.insert(0, GOLDEN)
    import make_fixtures as M

    f = load_fixture(FIX)
    m = f["meta"]
    run = M.env_command_run(m["arenas"], m["seed"], m["scale"], m["fillet_radius"], m["stream_ticks"])
    for key, want in run.items():
        assert np.array_equal(want, f[key]), key


def _twin_phases(T, h, f):
    """Runs the fixture's phases on the twin; yields (phase, balls, envs, pads) after each."""
    n = f["meta"]["arenas"]
    ticks = f["meta"]["stream_ticks"]
    for phase in _phases(f):
        if phase.startswith("step ") and phase[5:].isdigit():
            T.rgt_step(h, int(phase[5:]))
        elif phase in ("apply_env_commands", "apply_env_commands_device"):
            T.rgt_apply_env_commands(h, f["commands"].ctypes.data, n)
        elif phase == "step_with_device_commands":
            T.rgt_step_with_commands(h, f["command_stream"].ctypes.data, n, ticks)
        else:
            T.rgt_reset_to_kickoff(h, int(phase.split()[-1]))
        out = []
        for which, count, size in ((BALL, n, 288), (ENV, n, 44), (PAD, n * PADS_PER_ENV, 16)):
            a = np.zeros((count, size), np.uint8)
            T.rgt_download(h, which, a.ctypes.data, count)
            out.append(a)
        yield (phase, *out)


def test_twin_matches_fixture():
    T = twin()
    f = load_fixture(FIX)
    m = f["meta"]
    n = m["arenas"]
    h = T.rgt_create(n, m["scale"], m["fillet_radius"])
    try:
        for which, key, count in ((ENV, "init_env", n), (BALL, "init_ball", n), (PAD, "init_pads", n * PADS_PER_ENV)):
            T.rgt_upload(h, which, np.ascontiguousarray(f[key]).ctypes.data, count)
        for i, (phase, ball, env, pads) in enumerate(_twin_phases(T, h, f)):
            assert np.array_equal(ball, f["ball"][i]), phase
            assert np.array_equal(env, f["env"][i]), phase
            assert np.array_equal(pads, f["pads"][i]), phase
    finally:
        T.rgt_destroy(h)


def test_twin_settled_ball_is_the_references():
    T = twin()
    f = load_fixture(FIX)
    h = T.rgt_create(4, f["meta"]["scale"], f["meta"]["fillet_radius"])
    try:
        ball = np.zeros(288, np.uint8)
        assert T.rgt_settled_ball(h, ball.ctypes.data) == 0
        assert np.array_equal(ball, f["fresh_ball"][0])
    finally:
        T.rgt_destroy(h)


def test_twin_unsettled_reset_spawns_the_ball_asleep():
    """kickoff_settled off: the ball is reset to the layout's spawn, asleep, at rest (ResetBall with rules)."""
    T = twin()
    h = T.rgt_create(3, 1.0, 256.0)
    try:
        T.rgt_set_kickoff_settled(h, 0)
        T.rgt_step(h, 5)
        T.rgt_reset_to_kickoff(h, 0)
        balls = np.zeros((3, 288), np.uint8)
        envs = np.zeros((3, 44), np.uint8)
        T.rgt_download(h, BALL, balls.ctypes.data, 3)
        T.rgt_download(h, ENV, envs.ctypes.data, 3)
        # pos_bt = spawn x (1/50 as a float), pos = pos_bt x 50, both rounded to float (SetBallOriginBt).
        pos_bt = np.float32([0.0, 0.0, 92.75]) * np.float32(1.0 / 50.0)
        assert (_f32(balls, BALL_POS_BT, 3) == pos_bt).all()
        assert (_f32(balls, BALL_POS, 3) == pos_bt * np.float32(50.0)).all()
        assert (_f32(balls, BALL_VEL, 3) == 0).all() and (balls[:, BALL_SLEEPING] == 1).all()
        assert (envs[:, ENV_KICKOFF_RESETS] == 1).all()
        _orders_ok(envs)
    finally:
        T.rgt_destroy(h)


# --------------------------------------------------------------------------------------------------------------
# GPU, through the C-ABI.

def _gpu():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import rocketgpu_amd as rga

    return torch, rga


def _dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


@pytest.mark.gpu
def test_gpu_fresh_batch_is_the_references():
    torch, rga = _gpu()
    f = load_fixture(FIX)
    m = f["meta"]
    sim = rga.BatchSim(m["arenas"], 2, m["scale"], m["fillet_radius"])
    assert sim.boost_pad_capacity == m["arenas"] * PADS_PER_ENV
    assert np.array_equal(sim.download_state_array("ball_states"), f["fresh_ball"])
    assert np.array_equal(sim.download_state_array("env_states"), f["fresh_env"])
    assert np.array_equal(sim.download_state_array("boost_pad_states"), f["fresh_pads"])
    sim.close()


@pytest.mark.gpu
def test_gpu_env_commands_match_fixture():
    torch, rga = _gpu()
    f = load_fixture(FIX)
    m = f["meta"]
    n = m["arenas"]
    ticks = m["stream_ticks"]
    sim = rga.BatchSim(n, 2, m["scale"], m["fillet_radius"])
    sim.set_env_car_count(np.zeros(n, np.int32))
    sim.upload("env_states", f["init_env"])
    sim.upload("ball_states", f["init_ball"])
    sim.upload("boost_pad_states", f["init_pads"])
    d_cmds = _dev(torch, f["commands"])
    d_stream = _dev(torch, f["command_stream"])
    d_controls = _dev(torch, f["control_stream"])
    for i, phase in enumerate(_phases(f)):
        print(f"[env commands] phase {phase}", flush=True)
        if phase.startswith("step ") and phase[5:].isdigit():
            sim.step(int(phase[5:]))
        elif phase == "apply_env_commands":
            sim.apply_env_commands(f["commands"])
        elif phase == "apply_env_commands_device":
            sim.apply_env_commands_device(d_cmds.data_ptr(), n)
        elif phase == "step_with_device_commands":
            sim.step_with_device_commands(d_controls.data_ptr(), 2 * n, d_stream.data_ptr(), n, ticks)
            # The last tick's controls become the batch's controls (src/rocketgpu.cpp:2627-2629).
            assert np.array_equal(sim.download_state_array("controls"), f["control_stream"][-2 * n:])
        else:
            sim.reset_to_kickoff(int(phase.split()[-1]))
        torch.cuda.synchronize()
        assert np.array_equal(sim.download_state_array("ball_states"), f["ball"][i]), phase
        assert np.array_equal(sim.download_state_array("env_states"), f["env"][i]), phase
        assert np.array_equal(sim.download_state_array("boost_pad_states"), f["pads"][i]), phase
    sim.close()


@pytest.mark.gpu
def test_gpu_env_command_argument_errors():
    """The reference's checks and messages (src/rocketgpu.cpp:2136-2144, 2221-2227, 2561-2576)."""
    torch, rga = _gpu()
    L = rga.lib()
    sim = rga.BatchSim(8, 2)
    h = sim._h
    cmds = np.zeros((9, 16), np.uint8)
    assert L.rg_batch_apply_env_commands(h, None, 1, None) == 1
    assert "host_commands must not be null" in L.rg_last_error().decode()
    assert L.rg_batch_apply_env_commands(h, cmds.ctypes.data, 9, None) == 2
    assert L.rg_batch_apply_env_commands_device(h, None, 1, None) == 1
    assert "device_commands must not be null" in L.rg_last_error().decode()
    d = _dev(torch, np.zeros((64, 16), np.uint8))
    c = _dev(torch, np.zeros((64, 24), np.uint8))
    assert L.rg_batch_apply_env_commands_device(h, d.data_ptr(), 9, None) == 2
    assert L.rg_batch_step_with_device_commands(h, None, 0, d.data_ptr(), 8, 1, None) == 1
    assert "device_controls_by_tick must not be null" in L.rg_last_error().decode()
    assert L.rg_batch_step_with_device_commands(h, c.data_ptr(), 0, None, 8, 1, None) == 1
    assert "device_env_commands_by_tick must not be null" in L.rg_last_error().decode()
    assert L.rg_batch_step_with_device_commands(h, c.data_ptr(), 3, d.data_ptr(), 8, 1, None) == 1
    assert "control_stride" in L.rg_last_error().decode()
    assert L.rg_batch_step_with_device_commands(h, c.data_ptr(), 0, d.data_ptr(), 7, 1, None) == 1
    assert "env_command_stride must be at least env_count" in L.rg_last_error().decode()
    # Non-positive ticks are a no-op, before any check, as in the reference.
    assert L.rg_batch_step_with_device_commands(h, None, 0, None, 0, 0, None) == 0
    assert L.rg_batch_set_teams(h, np.zeros(17, np.uint8).ctypes.data, 17, None) == 2
    torch.cuda.synchronize()
    sim.close()


@pytest.mark.gpu
def test_gpu_unsettled_reset_matches_twin():
    torch, rga = _gpu()
    T = twin()
    n = 16
    sim = rga.BatchSim(n, 2)
    sim.set_kickoff_settled(False)
    balls = rga.workload_ball_states(5, 0, n)
    sim.upload("ball_states", balls)
    sim.step(40)
    sim.reset_to_kickoff(99)
    sim.step(3)
    h = T.rgt_create(n, 1.0, 256.0)
    try:
        T.rgt_set_kickoff_settled(h, 0)
        T.rgt_reset_to_kickoff(h, 0)
        T.rgt_upload(h, BALL, balls.ctypes.data, n)
        T.rgt_step(h, 40)
        T.rgt_reset_to_kickoff(h, 99)
        T.rgt_step(h, 3)
        for which, name, count, size in ((BALL, "ball_states", n, 288), (ENV, "env_states", n, 44)):
            want = np.zeros((count, size), np.uint8)
            T.rgt_download(h, which, want.ctypes.data, count)
            assert np.array_equal(sim.download_state_array(name), want), name
    finally:
        T.rgt_destroy(h)
        sim.close()
