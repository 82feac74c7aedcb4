"""Car pre-tick (StepEnvPreTickCar, include/rocketgpu/detail/step/env_step.hpp:6751-6873): field restore, internal
timers, tick-start caps and quantization, the jump / double jump / dodge / auto-flip state machine, the body
velocity clamp.  Checked byte for byte (CarState 880 B and WheelRuntimeState 292 B) against the reference's own
function on car states the reference's car arenas produce: the committed fixture, the live oracle, the host twin on
the CPU and the gfx950 kernel on the GPU.  A fifth of the cars are demoed with timers ending on, after and well
after the tick and every kind of respawn override, so respawns are covered.  Three-wheel bodies are not stepped by
this phase; the data holds none (the fixture test checks that)."""
import ctypes
import os

import numpy as np
import pytest

from conftest import REPO, load_fixture, oracle_or_skip, twin

FIX = "car_pretick_3v3"
PHASES_PATH = os.path.join(REPO, "rocketgpu_amd", "_build", "librg_phases.so")
SCALE, FILLET = 1.0, 256.0


def off(array, field):
    """Byte offset of a record field, from the library's state layouts."""
    import rocketgpu_amd as rg
    for name, _fmt, offset, _shape in rg.state_layouts()[array]["fields"]:
        if name == field:
            return offset
    raise KeyError(field)


def f32(rows, off, n=3):
    return np.ascontiguousarray(rows[:, off:off + 4 * n]).view(np.float32)


def twin_pretick(T, n_env, max_cars, seed, cars, controls, configs, wheels, envs, teams):
    c = np.ascontiguousarray(cars).copy()
    w = np.ascontiguousarray(wheels).copy()
    k, g, e, t = (np.ascontiguousarray(a) for a in (controls, configs, envs, teams))
    skipped = T.rgt_car_pretick(n_env, max_cars, SCALE, FILLET, seed, c.ctypes.data, k.ctypes.data, g.ctypes.data,
                                w.ctypes.data, e.ctypes.data, t.ctypes.data)
    return c, w, skipped


def batches(f):
    """(n_env, max_cars, seed, [(cars, controls, wheels, envs, out_cars, out_wheels) per batch], configs, teams)."""
    m = f["meta"]
    rows = [(f["cars"][k], f["controls"][k], f["wheels"][k], f["envs"][k], f["out_cars"][k], f["out_wheels"][k])
            for k in range(len(f["cars"]))]
    return m["arenas"], m["cars_per_env"], m["seed"], rows, f["configs"], f["teams"]


def test_fixture_covers_the_state_machine():
    f = load_fixture(FIX)
    n_env, cpe, _seed, rows, configs, teams = batches(f)
    cap = n_env * cpe
    assert f["cars"].shape[1:] == (cap, 880) and f["wheels"].shape[1:] == (cap, 292)
    assert configs.shape == (cap, 120) and teams.shape == (cap,)
    out = f["out_cars"].reshape(-1, 880)
    inp = f["cars"].reshape(-1, 880)
    flags = out[:, off("car_states", "cal_pre_flags")]
    assert ((flags & 1) != 0).sum() >= 5, "ground jumps started"
    assert ((flags & 2) != 0).sum() >= 5, "dodges started"
    # Respawns on the tick: demoed on the way in, not on the way out, with each kind of override.
    demoed_in = inp[:, off("car_states", "is_demoed")] != 0
    respawned = demoed_in & (out[:, off("car_states", "is_demoed")] == 0)
    assert respawned.sum() >= 20
    ov = inp[respawned, off("car_states", "respawn_slot_override_plus1")]
    assert (ov == 0).any() and ((ov >= 1) & (ov <= 4)).any() and ((ov >= 5) & (ov < 10)).any() and (ov >= 128).any()
    assert not configs[:, off("car_configs", "three_wheels")].any()


def test_oracle_reproduces_fixture():
    R = oracle_or_skip()
    n_env, cpe, seed, rows, configs, teams = batches(load_fixture(FIX))
    for cars, controls, wheels, envs, out_cars, out_wheels in rows:
        c, w = R.car_pretick(n_env, cpe, seed, cars, controls, configs, wheels, envs, teams)
        assert np.array_equal(c, out_cars) and np.array_equal(w, out_wheels)


def test_car_params_match_reference_defaults():
    R = oracle_or_skip()
    T = twin()
    ours = np.zeros(64, np.float32)
    n = T.rgt_car_params(ours.ctypes.data)
    ref = R.car_params()
    assert n == len(ref)
    assert np.array_equal(ours[:n].view(np.uint32), ref.view(np.uint32))


def test_axis_quantization():
    T = twin()
    x = np.concatenate([np.linspace(-1.2, 1.2, 20001, dtype=np.float32),
                        np.arange(-128, 128, dtype=np.float32) / 127.5, [-0.0, 0.0, 1e-30, -1e-30]]).astype(np.float32)
    out = np.zeros_like(x)
    T.rgt_quantize_axis(x.ctypes.data, out.ctypes.data, x.size)
    # Closed form: one byte per axis; values s / 128 below zero (exact) and s / 127 above (correctly rounded).
    vals = np.unique(out)
    neg = (np.arange(-128, 0, dtype=np.float64) / 128.0).astype(np.float32)
    pos = (np.arange(0, 128, dtype=np.float64) / 127.0).astype(np.float32)
    assert set(vals.tolist()) <= set(neg.tolist()) | set(pos.tolist())
    assert len(vals) == 256
    assert out[x >= 1.0].min() == 1.0 and out[x <= -1.0].max() == -1.0
    R = oracle_or_skip()
    assert np.array_equal(R.quantize_axis(x).view(np.uint32), out.view(np.uint32))


def test_twin_matches_fixture():
    T = twin()
    n_env, cpe, seed, rows, configs, teams = batches(load_fixture(FIX))
    for b, (cars, controls, wheels, envs, out_cars, out_wheels) in enumerate(rows):
        c, w, skipped = twin_pretick(T, n_env, cpe, seed, cars, controls, configs, wheels, envs, teams)
        assert skipped == 0
        bad = np.flatnonzero((c != out_cars).any(1) | (w != out_wheels).any(1))
        assert bad.size == 0, (b, bad[:10])


def test_jump_start_adds_the_jump_speed_along_up():
    """A car that starts a ground jump gets jump_immediate_speed (875/3 uu/s) along its up axis, unless the body
    velocity clamp at the car's top speed takes part of it back."""
    f = load_fixture(FIX)
    out = f["out_cars"].reshape(-1, 880)
    inp = f["cars"].reshape(-1, 880)
    # Cars that respawn on the tick start from rest (their jump speed is all of their velocity): left out here.
    started = np.flatnonzero(((out[:, off("car_states", "cal_pre_flags")] & 3) == 1) &
                             (inp[:, off("car_states", "is_demoed")] == 0))
    assert started.size >= 5
    vel_bt = off("car_states", "vel_bt")
    # Cars at the 2300 uu/s cap after the impulse are clamped back; the rest gain exactly the jump speed.
    v_out = f32(out[started], vel_bt).astype(np.float64) * 50.0
    free = np.linalg.norm(v_out, axis=1) < 2299.0
    assert free.sum() >= 5
    started = started[free]
    dv = f32(out[started], vel_bt) * 50.0 - f32(inp[started], vel_bt) * 50.0
    speed = np.linalg.norm(dv.astype(np.float64), axis=1)
    # Quantization of the tick-start velocity moves it by less than 0.02 uu/s per axis.
    assert np.all(np.abs(speed - 875.0 / 3.0) < 0.1), speed


def test_respawned_cars_stand_on_their_spot():
    """A car respawned without an offset or kickoff override stands on one of its team's four respawn spots, its
    spawn height above the spot, undemoed, with the spawn boost, at rest unless it jumped or dodged."""
    import rocketgpu_amd as rg
    f = load_fixture(FIX)
    n_env, cpe, _seed, _rows, configs, teams = batches(f)
    inp = f["cars"].reshape(-1, 880)
    out = f["out_cars"].reshape(-1, 880)
    cfg = np.tile(configs, (len(f["cars"]), 1))
    team = np.tile(teams, len(f["cars"]))
    ov = inp[:, off("car_states", "respawn_slot_override_plus1")]
    done = (inp[:, off("car_states", "is_demoed")] != 0) & (out[:, off("car_states", "is_demoed")] == 0)
    plain = np.flatnonzero(done & (ov < 5))
    assert plain.size >= 5
    spots = {0: [(-2304.0, -4608.0), (-2688.0, -4608.0), (2304.0, -4608.0), (2688.0, -4608.0)]}
    spots[1] = [(x, -y) for x, y in spots[0]]
    pos = f32(out[plain], off("car_states", "pos"))
    spawn_z = cfg[plain, off("car_configs", "spawn_z"):off("car_configs", "spawn_z") + 4].view(np.float32)[:, 0]
    for p, z, t in zip(pos, spawn_z, team[plain]):
        assert min(abs(p[0] - x) + abs(p[1] - y) for x, y in spots[int(t)]) < 0.02, p
        assert abs(p[2] - z) < 0.02
    # At rest unless the same pre-tick already jumped or dodged (a respawned car can dodge on its first tick).
    rest = (out[plain, off("car_states", "cal_pre_flags")] & 7) == 0
    assert rest.sum() >= 3
    assert np.all(f32(out[plain[rest]], off("car_states", "vel")) == 0.0)
    assert np.allclose(f32(out[plain], off("car_states", "boost"), 1)[:, 0], 33.3)


def live_batches(R, n_env, cpe, ticks, seed, every):
    import rocketgpu_amd as rg
    from tests.golden.make_fixtures import demo_some_cars
    lay = {f[0]: f[2] for f in rg.state_layouts()["car_states"]["fields"]}
    run = R.car_arena_run(n_env, cpe, ticks, seed)
    for t in range(0, ticks, every):
        cars = demo_some_cars(run["cars"][t], seed + t, lay)
        rc, rw = R.car_pretick(n_env, cpe, seed, cars, run["controls"][t], run["configs"], run["wheels"][t],
                               run["envs"][t], run["teams"])
        yield cars, run["controls"][t], run["wheels"][t], run["envs"][t], rc, rw, run["configs"], run["teams"]


def test_twin_matches_oracle_on_fresh_car_arenas():
    R = oracle_or_skip()
    T = twin()
    for n_env, cpe, ticks, seed in ((8, 2, 200, 101), (3, 6, 160, 102)):
        for cars, controls, wheels, envs, rc, rw, configs, teams in live_batches(R, n_env, cpe, ticks, seed, 1):
            c, w, skipped = twin_pretick(T, n_env, cpe, seed, cars, controls, configs, wheels, envs, teams)
            assert skipped == 0
            bad = np.flatnonzero((c != rc).any(1) | (w != rw).any(1))
            assert bad.size == 0, (n_env, cpe, bad[:10])


# ----------------------------------------------------------------------------------------------------------------
# GPU: the gfx950 kernel of the phase.

def gpu_pretick(n_env, max_cars, seed, cars, controls, configs, wheels, envs, teams):
    import torch
    if not os.path.exists(PHASES_PATH):
        pytest.fail("librg_phases.so is not built")
    P = ctypes.CDLL(PHASES_PATH)
    vp = ctypes.c_void_p
    P.rgp_car_pretick.argtypes = [vp, vp, vp, vp, vp, vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_double,
                                  ctypes.c_double, ctypes.c_uint64, vp, vp]
    n = len(cars)
    dev = {k: torch.from_numpy(np.ascontiguousarray(v)).to("cuda")
           for k, v in (("cars", cars), ("controls", controls), ("configs", configs), ("wheels", wheels),
                        ("envs", envs), ("teams", teams))}
    unsupported = torch.zeros(n, dtype=torch.uint8, device="cuda")
    rc = P.rgp_car_pretick(dev["cars"].data_ptr(), dev["controls"].data_ptr(), dev["configs"].data_ptr(),
                           dev["wheels"].data_ptr(), dev["envs"].data_ptr(), dev["teams"].data_ptr(), n_env,
                           max_cars, SCALE, FILLET, seed, unsupported.data_ptr(), None)
    assert rc == 0
    torch.cuda.synchronize()
    return dev["cars"].cpu().numpy(), dev["wheels"].cpu().numpy(), int(unsupported.sum())


@pytest.mark.gpu
def test_gpu_matches_fixture():
    n_env, cpe, seed, rows, configs, teams = batches(load_fixture(FIX))
    for b, (cars, controls, wheels, envs, out_cars, out_wheels) in enumerate(rows):
        c, w, unsupported = gpu_pretick(n_env, cpe, seed, cars, controls, configs, wheels, envs, teams)
        assert unsupported == 0
        bad = np.flatnonzero((c != out_cars).any(1) | (w != out_wheels).any(1))
        assert bad.size == 0, (b, bad[:10])


@pytest.mark.gpu
def test_gpu_matches_oracle_on_fresh_car_arenas():
    R = oracle_or_skip()
    for n_env, cpe, ticks, seed in ((16, 2, 240, 201), (6, 6, 160, 202)):
        for cars, controls, wheels, envs, rc, rw, configs, teams in live_batches(R, n_env, cpe, ticks, seed, 3):
            c, w, unsupported = gpu_pretick(n_env, cpe, seed, cars, controls, configs, wheels, envs, teams)
            assert unsupported == 0
            bad = np.flatnonzero((c != rc).any(1) | (w != rw).any(1))
            assert bad.size == 0, (n_env, cpe, bad[:10])
