"""The step core compiled for the host (the twin) against the golden fixtures, tick by tick.

The twin runs the same core sources as the gfx950 kernels; agreement here pins the core's algorithm to the
reference CPU build, so a GPU mismatch can only come from the device compilation or the device-side plumbing.
"""
import ctypes

import numpy as np
import pytest

from conftest import fixture_names, load_fixture, twin
from digests import ball_env_digests, fold


def run_twin(T, f, ticks, every=None):
    m = f["meta"]
    n = m["arenas"]
    h = T.rgt_create(n, m["scale"], m["fillet_radius"])
    assert h, "rgt_create failed"
    try:
        assert T.rgt_upload(h, 2, f["init_env"].ctypes.data, n) == 0
        assert T.rgt_upload(h, 1, f["init_ball"].ctypes.data, n) == 0
        batch = np.zeros(ticks, np.uint64)
        env = []
        d = np.zeros(n, np.uint64)
        b = np.zeros(1, np.uint64)
        for t in range(ticks):
            T.rgt_step(h, 1)
            T.rgt_env_digests(h, d.ctypes.data, b.ctypes.data)
            batch[t] = b[0]
            if every and (t + 1) % every == 0:
                env.append(d.copy())
        balls = np.zeros((n, 288), np.uint8)
        envs = np.zeros((n, 44), np.uint8)
        T.rgt_download(h, 1, balls.ctypes.data, n)
        T.rgt_download(h, 2, envs.ctypes.data, n)
        return batch, np.array(env), balls, envs
    finally:
        T.rgt_destroy(h)


@pytest.mark.parametrize("name", fixture_names())
def test_twin_matches_fixture(name):
    T = twin()
    f = load_fixture(name)
    ticks, every = f["meta"]["ticks"], int(f["checkpoint_every"])
    batch, env, balls, envs = run_twin(T, f, ticks, every)
    first_bad = np.flatnonzero(batch != f["batch_digest"])
    assert first_bad.size == 0, f"first differing tick {first_bad[0]}"
    assert np.array_equal(env, f["env_digest"])
    assert np.array_equal(balls, f["final_ball"]), "final BallState bytes differ"
    assert np.array_equal(envs, f["final_env"]), "final EnvState bytes differ"


def test_twin_digest_helper_agrees():
    """The C++ digest of the twin and the numpy digest of the tests agree on the same states."""
    T = twin()
    f = load_fixture("ball_arena_96_t1000")
    m = f["meta"]
    n = m["arenas"]
    h = T.rgt_create(n, m["scale"], m["fillet_radius"])
    try:
        T.rgt_upload(h, 2, f["init_env"].ctypes.data, n)
        T.rgt_upload(h, 1, f["init_ball"].ctypes.data, n)
        T.rgt_step(h, 37)
        d = np.zeros(n, np.uint64)
        b = np.zeros(1, np.uint64)
        T.rgt_env_digests(h, d.ctypes.data, b.ctypes.data)
        balls = np.zeros((n, 288), np.uint8)
        T.rgt_download(h, 1, balls.ctypes.data, n)
        assert np.array_equal(ball_env_digests(balls), d)
        assert fold(d) == b[0]
        assert b[0] == f["batch_digest"][36]
    finally:
        T.rgt_destroy(h)


def test_twin_multi_tick_equals_single_ticks():
    T = twin()
    f = load_fixture("ball_arena_small_scale_64_t1500")
    m = f["meta"]
    n = m["arenas"]
    out = []
    for chunks in ([1] * 120, [120], [7, 13, 100]):
        h = T.rgt_create(n, m["scale"], m["fillet_radius"])
        T.rgt_upload(h, 2, f["init_env"].ctypes.data, n)
        T.rgt_upload(h, 1, f["init_ball"].ctypes.data, n)
        for c in chunks:
            T.rgt_step(h, c)
        balls = np.zeros((n, 288), np.uint8)
        T.rgt_download(h, 1, balls.ctypes.data, n)
        T.rgt_destroy(h)
        out.append(balls)
    assert np.array_equal(out[0], out[1]) and np.array_equal(out[0], out[2])


def test_free_flight_ball_closed_form():
    """A ball far from every surface, with no spin, follows damped semi-implicit Euler under gravity.

    Per tick (tick_dt = 1/120): v <- v * damping (damping = (1 - 0.03)^(1/120) in float32, the ball drag of the
    game), then v.z += -650 * dt, then pos += v * dt, all in Bullet units (uu / 50).  Each tick the game also
    quantizes the ball's velocity toward zero on a 0.01 uu/s grid (moving it by at most 0.011 uu/s) and its
    position to 0.01 uu, so the bounds are that many ulps of the grid per tick.  Over 30 ticks the drag moves the
    velocity by about 2.3 uu/s and the position by about 0.3 uu, gravity by 162 uu/s: a dropped term, a wrong sign
    or a wrong tick_dt fails.
    """
    T = twin()
    n = 1
    h = T.rgt_create(n, 1.0, 256.0)
    try:
        from conftest import load_fixture as _lf

        f = _lf("ball_arena_96_t1000")
        ball = f["init_ball"][:1].copy()
        pos = np.array([0.0, 0.0, 1000.0], np.float32)
        vel = np.array([300.0, -200.0, 400.0], np.float32)
        ball[0, 40:52] = pos.view(np.uint8)
        ball[0, 68:80] = vel.view(np.uint8)
        ball[0, 80:92] = np.zeros(3, np.float32).view(np.uint8)
        T.rgt_upload(h, 2, f["init_env"][:1].ctypes.data, n)
        T.rgt_upload(h, 1, ball.ctypes.data, n)
        ticks = 30
        T.rgt_step(h, ticks)
        out = np.zeros((1, 288), np.uint8)
        T.rgt_download(h, 1, out.ctypes.data, 1)
        got_pos = out[0, 40:52].copy().view(np.float32)
        got_vel = out[0, 68:80].copy().view(np.float32)
        p = pos.astype(np.float64) / 50.0
        v = vel.astype(np.float64) / 50.0
        dt = 1.0 / 120.0
        damp = (1.0 - 0.03) ** dt
        for _ in range(ticks):
            v = v * damp
            v[2] += -650.0 / 50.0 * dt
            p = p + v * dt
        assert np.allclose(got_vel, v * 50.0, rtol=0, atol=0.011 * ticks), (got_vel, v * 50)
        assert np.allclose(got_pos, p * 50.0, rtol=0, atol=0.005 + 0.011 * ticks * ticks / 120.0), (got_pos, p * 50)
        # Without spin the rotation stays the identity, up to the game's 18-bit smallest-three quaternion grid
        # (whose zero sits half a step, about 2.7e-6, off zero), and stays a unit quaternion.
        q = out[0, 52:68].copy().view(np.float32).astype(np.float64)
        assert np.allclose(q, [1, 0, 0, 0], rtol=0, atol=1e-5), q
        assert abs(np.dot(q, q) - 1.0) < 1e-6
    finally:
        T.rgt_destroy(h)
