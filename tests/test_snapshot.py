"""Full-state snapshot (the checkpoint form of the reference's Arena::GetFullState / SetFullState / Serialize,
src/rocketsim_compat_dev.cpp:369-447): a batch saved at tick t and loaded into a fresh batch of the same shape
continues bit for bit; its public sections equal the reference CPU build's records at tick t; damaged or foreign
snapshots are rejected before anything is written."""
import ctypes

import numpy as np
import pytest

from conftest import oracle_or_skip
from test_events import arena_balls

HEADER_BYTES = 72
PUBLIC = ("car_states", "controls", "ball_states", "env_states", "boost_pad_states")


def fnv1a(data: bytes) -> int:
    h = 0xcbf29ce484222325
    arr = np.frombuffer(data, np.uint8)
    for b in arr.tolist():
        h = ((h ^ b) * 0x100000001b3) & 0xFFFFFFFFFFFFFFFF
    return h


def make_batch(n, balls, shape=(64, 5, 8)):
    import rocketgpu_amd as rg
    sim = rg.BatchSim(n, 2)
    sim.set_env_car_count(np.zeros(n, np.int32))
    sim.upload("ball_states", balls)
    sim.set_launch_shape(*shape)
    return sim


def public_state(sim):
    return {a: sim.download_state_array(a) for a in PUBLIC}


@pytest.mark.gpu
def test_snapshot_resume_is_bit_exact():
    R = oracle_or_skip()
    n = 64
    balls = arena_balls(R, n, seed=29)
    a = make_batch(n, balls)
    b = make_batch(n, R.workload_ball_states(3, 0, n, 1.0))
    try:
        a.step(37)
        snap = a.save_state()
        a.step(50)
        after_a = public_state(a)
        snap_a = a.save_state()
        b.step(11)  # b's own history is overwritten by the load
        b.load_state(snap)
        b.step(50)
        assert all(np.array_equal(after_a[k], v) for k, v in public_state(b).items())
        assert b.save_state() == snap_a
        # Loading into the batch it came from rewinds it.
        a.step(9)
        a.load_state(snap)
        a.step(50)
        assert a.save_state() == snap_a
    finally:
        a.close()
        b.close()


@pytest.mark.gpu
def test_snapshot_public_sections_match_reference():
    R = oracle_or_skip()
    n = 48
    balls = arena_balls(R, n, seed=31)
    sim = make_batch(n, balls)
    ref = R.RefSim(n, 2)
    ref.set_env_car_count(np.zeros(n, np.int32))
    ref.upload(R.BALL, balls)
    try:
        sim.step(60)
        ref.step(60)
        snap = sim.save_state()
        cap = n * 2
        off = HEADER_BYTES + cap * (880 + 24 + 120 + 1)
        ball_sec = np.frombuffer(snap, np.uint8, n * 288, off).reshape(n, 288)
        env_sec = np.frombuffer(snap, np.uint8, n * 44, off + n * 288).reshape(n, 44)
        assert np.array_equal(ball_sec, ref.download(R.BALL, n))
        assert np.array_equal(env_sec, ref.download(R.ENV, n))
        cars = np.frombuffer(snap, np.uint8, cap * 880, HEADER_BYTES).reshape(cap, 880)
        assert np.array_equal(cars, ref.download(R.CAR, cap))
        full = sim.get_full_state(["ball_states"])
        assert full["ball_states.pos"] == np.ascontiguousarray(ball_sec[:, 40:52]).tobytes()
        assert all(k.startswith("ball_states.") for k in full)
        with pytest.raises(ValueError):
            sim.get_full_state(["wheel_runtime"])
    finally:
        sim.close()


@pytest.mark.gpu
def test_snapshot_rejects_damaged_or_foreign_data():
    import rocketgpu_amd as rg
    R = oracle_or_skip()
    n = 16
    sim = make_batch(n, arena_balls(R, n, seed=37))
    other = make_batch(n + 1, arena_balls(R, n + 1, seed=37))
    try:
        sim.step(20)
        snap = bytearray(sim.save_state())
        before = sim.save_state()
        with pytest.raises(rg.RocketGpuError, match="this batch's snapshots are"):
            sim.load_state(bytes(snap[:-1]))
        damaged = bytearray(snap)
        damaged[HEADER_BYTES + 1000] ^= 1
        with pytest.raises(rg.RocketGpuError, match="digest mismatch"):
            sim.load_state(bytes(damaged))
        with pytest.raises(rg.RocketGpuError, match="this batch's snapshots are"):
            other.load_state(bytes(snap))
        wrong = bytearray(snap)
        wrong[:8] = b"NOTASNAP"
        with pytest.raises(rg.RocketGpuError, match="not a batch snapshot"):
            sim.load_state(bytes(wrong))
        # A car count above max_cars_per_env, with a digest that matches: rejected by the range checks.
        cap = n * 2
        off_counts = HEADER_BYTES + cap * (880 + 24 + 120 + 1) + n * (288 + 44) + n * 34 * 16
        bad = bytearray(snap)
        bad[off_counts:off_counts + 4] = np.int32(99).tobytes()
        bad[56:64] = np.uint64(fnv1a(bytes(bad[HEADER_BYTES:]))).tobytes()
        with pytest.raises(rg.RocketGpuError, match="invalid car count"):
            sim.load_state(bytes(bad))
        assert sim.save_state() == before  # nothing was written by the rejected loads
        n_bytes = ctypes.c_size_t()
        assert rg.lib().rg_batch_state_snapshot_size(sim._h, ctypes.byref(n_bytes)) == 0
        assert n_bytes.value == len(snap)
    finally:
        sim.close()
        other.close()
