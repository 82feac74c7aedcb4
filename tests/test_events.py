"""Physics event stream (include/rocketgpu/rocketgpu.h:45-98, BatchSim::set_physics_event_stream rocketgpu.h:1015).

In arenas without cars the step emits two kinds of events: BallWorldContact for every mesh triangle the ball's
contact query touches (arena_ball.hpp:1927-1936, in the order the query finds them) and Goal on a tick that scores
(env_step.hpp:7025-7032).  The env pre-tick stamps each window with the env's tick (env_step.hpp:6695-6697); the
caller clears count and overflow.  Checked field by field (padding bytes are unspecified) against the reference CPU
build stepping the same arenas: the host twin on the CPU, the C-ABI library on the GPU across launch shapes and step
forms.
"""
import ctypes

import numpy as np
import pytest

from conftest import oracle_or_skip, twin

EVENT_FIELDS = ["tick", "type", "car_id", "other_car_id", "team", "pad_index", "car_on_ground", "is_big_pad",
                "is_demo", "position", "normal", "car_velocity", "ball_position", "ball_velocity", "boost_before",
                "boost_after"]
BALL_WORLD_CONTACT, GOAL = 1, 5


def event_dtype():
    import rocketgpu_amd as rg
    return rg.PHYSICS_EVENT_DTYPE, rg.PHYSICS_EVENT_WINDOW_DTYPE


def arena_balls(R, n, seed=11):
    """First half: the benchmark's seeded ball launches (wall and corner contacts).  Second half: balls 300-900 uu
    in front of a goal mouth, moving into it at 1400-2200 uu/s (goals within the first ~80 ticks)."""
    balls = R.workload_ball_states(seed, 0, n, 1.0)
    rng = np.random.default_rng(seed)
    pos_off, vel_off = R.offsetof(R.BALL, "pos"), R.offsetof(R.BALL, "vel")
    for e in range(n // 2, n):
        side = 1.0 if e % 2 == 0 else -1.0
        pos = np.array([rng.uniform(-650, 650), side * rng.uniform(4220, 4820), rng.uniform(100, 500)], np.float32)
        vel = np.array([rng.uniform(-200, 200), side * rng.uniform(1400, 2200), rng.uniform(-300, 300)], np.float32)
        balls[e, pos_off:pos_off + 12] = pos.view(np.uint8)
        balls[e, vel_off:vel_off + 12] = vel.view(np.uint8)
    return balls


def same_events(ref_records, ref_windows, got_records, got_windows, cap):
    """Windows equal byte for byte; each env's first `count` records equal in every field."""
    ev, win = event_dtype()
    rw = np.ascontiguousarray(ref_windows).view(win).reshape(-1)
    gw = np.ascontiguousarray(got_windows).view(win).reshape(-1)
    assert np.array_equal(rw["tick"], gw["tick"])
    assert np.array_equal(rw["count"], gw["count"])
    assert np.array_equal(rw["overflow"], gw["overflow"])
    rr = np.ascontiguousarray(ref_records).view(ev).reshape(len(rw), cap)
    gr = np.ascontiguousarray(got_records).view(ev).reshape(len(gw), cap)
    for e in np.flatnonzero(rw["count"]):
        c = min(int(rw["count"][e]), cap)
        for k in EVENT_FIELDS:
            a = np.ascontiguousarray(rr[e, :c][k])
            b = np.ascontiguousarray(gr[e, :c][k])
            assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), (e, k)
    return rr, rw


def count_types(records, windows, cap, counts):
    ev, win = event_dtype()
    rw = np.ascontiguousarray(windows).view(win).reshape(-1)
    rr = np.ascontiguousarray(records).view(ev).reshape(len(rw), cap)
    for e in np.flatnonzero(rw["count"]):
        for t in rr[e, :min(int(rw["count"][e]), cap)]["type"]:
            counts[int(t)] = counts.get(int(t), 0) + 1


class TwinEvents:
    def __init__(self, T, n, balls, cap, goal_only=False):
        self.T, self.n, self.cap = T, n, cap
        self.h = T.rgt_create(n, 1.0, 256.0)
        T.rgt_upload(self.h, 1, balls.ctypes.data, n)
        self.records = np.zeros((n * cap, 104), np.uint8)
        self.windows = np.zeros((n, 16), np.uint8)
        T.rgt_set_event_stream(self.h, self.records.ctypes.data, self.windows.ctypes.data, cap, 1 if goal_only else 0)

    def step(self, ticks):
        self.windows[:, 8:16] = 0
        self.T.rgt_step(self.h, ticks)

    def close(self):
        self.T.rgt_destroy(self.h)


class RefEvents:
    def __init__(self, R, n, balls, cap, goal_only=False):
        self.s = R.RefSim(n, 2)
        self.s.set_env_car_count(np.zeros(n, np.int32))
        self.s.upload(R.BALL, balls)
        self.cap = cap
        self.records = np.zeros((n * cap, 104), np.uint8)
        self.windows = np.zeros((n, 16), np.uint8)
        self.s.set_event_stream(self.records, self.windows, cap, goal_only)

    def step(self, ticks):
        self.windows[:, 8:16] = 0
        self.s.step(ticks)


def test_event_records_match_reference_layout():
    R = oracle_or_skip()
    ev, win = event_dtype()
    assert R.sizeof(7) == ev.itemsize == 104
    assert R.sizeof(8) == win.itemsize == 16
    for k in EVENT_FIELDS:
        assert R.offsetof(7, k) == ev.fields[k][1], k
    for k in ("tick", "count", "overflow"):
        assert R.offsetof(8, k) == win.fields[k][1], k
    import rocketgpu_amd as rg
    lay = rg.physics_event_layout()
    assert lay == {"record_size": 104, "window_size": 16, "type_offset": 8, "velocity_offset": 80,
                   "count_offset": 8, "overflow_offset": 12, "goal_type": GOAL}


@pytest.mark.parametrize("ticks_per_step", [1, 4])
def test_twin_events_match_oracle(ticks_per_step):
    R = oracle_or_skip()
    T = twin()
    n, cap = 48, 24
    balls = arena_balls(R, n)
    ref, tw = RefEvents(R, n, balls, cap), TwinEvents(T, n, balls, cap)
    counts = {}
    try:
        for _ in range(240 // ticks_per_step):
            ref.step(ticks_per_step)
            tw.step(ticks_per_step)
            same_events(ref.records, ref.windows, tw.records, tw.windows, cap)
            count_types(ref.records, ref.windows, cap, counts)
    finally:
        tw.close()
    assert counts.get(BALL_WORLD_CONTACT, 0) > 200 and counts.get(GOAL, 0) > 20, counts


def test_twin_events_overflow_and_goal_only():
    R = oracle_or_skip()
    T = twin()
    n = 32
    balls = arena_balls(R, n, seed=5)
    for cap, goal_only in ((2, False), (3, True)):
        ref, tw = RefEvents(R, n, balls, cap, goal_only), TwinEvents(T, n, balls, cap, goal_only)
        overflowed = 0
        counts = {}
        try:
            for _ in range(30):
                ref.step(8)
                tw.step(8)
                _, rw = same_events(ref.records, ref.windows, tw.records, tw.windows, cap)
                overflowed += int(rw["overflow"].sum())
                count_types(ref.records, ref.windows, cap, counts)
        finally:
            tw.close()
        assert overflowed > 0
        if goal_only:
            assert set(counts) == {GOAL}


def test_twin_events_leave_the_state_unchanged():
    R = oracle_or_skip()
    T = twin()
    n = 32
    balls = arena_balls(R, n, seed=3)
    on = TwinEvents(T, n, balls, 16)
    off = T.rgt_create(n, 1.0, 256.0)
    try:
        T.rgt_upload(off, 1, balls.ctypes.data, n)
        for _ in range(20):
            on.step(6)
            T.rgt_step(off, 6)
        a = np.zeros((n, 288), np.uint8)
        b = np.zeros((n, 288), np.uint8)
        T.rgt_download(on.h, 1, a.ctypes.data, n)
        T.rgt_download(off, 1, b.ctypes.data, n)
        assert np.array_equal(a, b)
    finally:
        on.close()
        T.rgt_destroy(off)


# ----------------------------------------------------------------------------------------------------------------
# GPU: the C-ABI library.

class GpuEvents:
    def __init__(self, n, balls, cap, goal_only=False, shape=None):
        import torch
        import rocketgpu_amd as rg
        self.torch = torch
        self.sim = rg.BatchSim(n, 2)
        self.sim.set_env_car_count(np.zeros(n, np.int32))
        self.sim.upload("ball_states", balls)
        if shape is not None:
            self.sim.set_launch_shape(*shape)
        self.cap = cap
        self.records = torch.zeros(n * cap * 104, dtype=torch.uint8, device="cuda")
        self.windows = torch.zeros(n * 16, dtype=torch.uint8, device="cuda")
        self.sim.set_physics_event_stream(self.records.data_ptr(), self.windows.data_ptr(), cap, goal_only)

    def clear(self):
        self.windows.view(-1, 16)[:, 8:16].zero_()

    def host(self):
        self.torch.cuda.synchronize()
        return self.records.cpu().numpy().reshape(-1, 104), self.windows.cpu().numpy().reshape(-1, 16)

    def close(self):
        self.sim.set_physics_event_stream(0, 0, 0)
        self.sim.close()


SHAPES = [(64, 0, 64), (64, 3, 8), (64, 4, 8), (64, 5, 8), (128, 5, 4), (64, 6, 8), (128, 7, 8)]


@pytest.mark.gpu
@pytest.mark.parametrize("shape", SHAPES)
def test_gpu_events_match_oracle(shape):
    R = oracle_or_skip()
    n, cap = 96, 24
    balls = arena_balls(R, n, seed=17)
    ref = RefEvents(R, n, balls, cap)
    g = GpuEvents(n, balls, cap, shape=shape)
    counts = {}
    try:
        for k in range(60):
            ticks = 1 if k < 30 else 4
            ref.step(ticks)
            g.clear()
            g.sim.step(ticks)
            rec, win = g.host()
            same_events(ref.records, ref.windows, rec, win, cap)
            count_types(ref.records, ref.windows, cap, counts)
    finally:
        g.close()
    assert counts.get(BALL_WORLD_CONTACT, 0) > 200 and counts.get(GOAL, 0) > 20, counts


@pytest.mark.gpu
def test_gpu_events_other_step_forms():
    """step_graph, step_with_device_commands (no commands active) and ball_path write the same events as step."""
    import torch
    R = oracle_or_skip()
    n, cap = 64, 32
    balls = arena_balls(R, n, seed=23)
    ref = RefEvents(R, n, balls, cap)
    g = GpuEvents(n, balls, cap, shape=(64, 5, 8))
    try:
        # Graph of two 3-tick launches: one window spans both.
        for _ in range(10):
            ref.step(6)
            g.clear()
            g.sim.step_graph(3, 2)
            rec, win = g.host()
            same_events(ref.records, ref.windows, rec, win, cap)
        # Device commands, all inactive.
        controls = torch.zeros(4 * 2 * n * 24, dtype=torch.uint8, device="cuda")
        commands = torch.zeros(4 * n * 16, dtype=torch.uint8, device="cuda")
        for _ in range(5):
            ref.step(4)
            g.clear()
            g.sim.step_with_device_commands(controls.data_ptr(), 2 * n, commands.data_ptr(), n, 4)
            rec, win = g.host()
            same_events(ref.records, ref.windows, rec, win, cap)
        # ball_path: 3 samples of 4 ticks, stopping at |y| >= 4000.
        ref.windows[:, 8:16] = 0
        ref.s.ball_path(3, 4, 4000.0)
        g.clear()
        out = torch.zeros(n * 3 * 6, dtype=torch.float32, device="cuda")
        res = torch.zeros(n * 2, dtype=torch.int32, device="cuda")
        g.sim.ball_path(3, 4, 4000.0, out.data_ptr(), res.data_ptr())
        rec, win = g.host()
        same_events(ref.records, ref.windows, rec, win, cap)
    finally:
        g.close()


@pytest.mark.gpu
def test_gpu_event_stream_validation():
    import rocketgpu_amd as rg
    sim = rg.BatchSim(8, 2)
    L = rg.lib()
    try:
        buf = ctypes.c_void_p(0x1000)  # never dereferenced: every call below is rejected or disables the stream
        for records, windows, cap, envs, cars in ((buf, buf, 4, 7, 2), (buf, buf, 4, 8, 3), (None, buf, 4, 8, 2),
                                                  (buf, None, 4, 8, 2), (buf, buf, 0, 8, 2)):
            rc = L.rg_batch_set_physics_event_stream(sim._h, records, windows, cap, envs, cars, 0)
            assert rc == 1
            assert L.rg_last_error().decode() == "physics event stream must match the complete batch"
        assert L.rg_batch_set_physics_event_stream(sim._h, None, None, 0, 0, 0, 0) == 0
        sim.step(2)
        sim.synchronize()
    finally:
        sim.close()
