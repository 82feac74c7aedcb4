"""Car pre-tick (StepEnvPreTickCar, include/rocketgpu/detail/step/env_step.hpp:6751-6873): field restore, internal
timers, tick-start caps and quantization, the jump / double jump / dodge / auto-flip state machine, the body
velocity clamp.  Checked byte for byte (CarState 880 B and WheelRuntimeState 292 B) against the reference's own
function on car states the reference's car arenas produce: the committed fixture, the live oracle, the host twin on
the CPU and the gfx950 kernel on the GPU.  A demoed car whose respawn falls on the tick and three-wheel bodies are
not stepped by this phase; the data below holds neither (the fixture test checks that)."""
import ctypes
import os

import numpy as np
import pytest

from conftest import REPO, load_fixture, oracle_or_skip, twin

FIX = "car_pretick_1v1_3v3"
PHASES_PATH = os.path.join(REPO, "rocketgpu_amd", "_build", "librg_phases.so")


def off(array, field):
    """Byte offset of a record field, from the library's state layouts."""
    import rocketgpu_amd as rg
    for name, _fmt, offset, _shape in rg.state_layouts()[array]["fields"]:
        if name == field:
            return offset
    raise KeyError(field)


def f32(rows, off, n=3):
    return np.ascontiguousarray(rows[:, off:off + 4 * n]).view(np.float32)


def twin_pretick(T, cars, controls, configs, wheels):
    c = np.ascontiguousarray(cars).copy()
    w = np.ascontiguousarray(wheels).copy()
    k = np.ascontiguousarray(controls)
    g = np.ascontiguousarray(configs)
    skipped = T.rgt_car_pretick(len(c), c.ctypes.data, k.ctypes.data, g.ctypes.data, w.ctypes.data)
    return c, w, skipped


def test_fixture_covers_the_state_machine():
    f = load_fixture(FIX)
    n = f["meta"]["rows"]
    assert f["cars"].shape == (n, 880) and f["wheels"].shape == (n, 292)
    assert f["controls"].shape == (n, 24) and f["configs"].shape == (n, 120)
    flags = f["out_cars"][:, off("car_states", "cal_pre_flags")]
    assert ((flags & 1) != 0).sum() >= 20, "ground jumps started"
    assert ((flags & 2) != 0).sum() >= 20, "dodges started"
    assert ((flags & 32) != 0).sum() >= 20, "cars mid-dodge at the tick start"
    # No three-wheel body, no demoed car.
    assert not f["configs"][:, off("car_configs", "three_wheels")].any()
    assert not f["cars"][:, off("car_states", "is_demoed")].any()


def test_oracle_reproduces_fixture():
    R = oracle_or_skip()
    f = load_fixture(FIX)
    c, w = R.car_pretick(f["cars"], f["controls"], f["configs"], f["wheels"])
    assert np.array_equal(c, f["out_cars"]) and np.array_equal(w, f["out_wheels"])


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
    f = load_fixture(FIX)
    c, w, skipped = twin_pretick(T, f["cars"], f["controls"], f["configs"], f["wheels"])
    assert skipped == 0
    bad = np.flatnonzero((c != f["out_cars"]).any(1) | (w != f["out_wheels"]).any(1))
    assert bad.size == 0, bad[:10]


def test_jump_start_adds_the_jump_speed_along_up():
    """A car that starts a ground jump gets jump_immediate_speed (875/3 uu/s) along its up axis, and nothing else
    changes its velocity in the pre-tick when it is not clamped."""
    f = load_fixture(FIX)
    started = np.flatnonzero((f["out_cars"][:, off("car_states", "cal_pre_flags")] & 3) == 1)
    assert started.size >= 10
    vel_bt = off("car_states", "vel_bt")
    dv = f32(f["out_cars"][started], vel_bt) * 50.0 - f32(f["cars"][started], vel_bt) * 50.0
    speed = np.linalg.norm(dv.astype(np.float64), axis=1)
    # Quantization of the tick-start velocity moves it by less than 0.02 uu/s per axis.
    assert np.all(np.abs(speed - 875.0 / 3.0) < 0.1), speed


def test_twin_matches_oracle_on_fresh_car_arenas():
    R = oracle_or_skip()
    T = twin()
    for n_env, cpe, ticks, seed in ((8, 2, 240, 101), (3, 6, 160, 102)):
        cars, wheels, controls, configs = R.car_arena_inputs(n_env, cpe, ticks, seed)
        rc, rw = R.car_pretick(cars, controls, configs, wheels)
        c, w, skipped = twin_pretick(T, cars, controls, configs, wheels)
        assert skipped == 0
        bad = np.flatnonzero((c != rc).any(1) | (w != rw).any(1))
        assert bad.size == 0, (n_env, cpe, bad[:10])


# ----------------------------------------------------------------------------------------------------------------
# GPU: the gfx950 kernel of the phase.

def gpu_pretick(cars, controls, configs, wheels):
    import torch
    if not os.path.exists(PHASES_PATH):
        pytest.fail("librg_phases.so is not built")
    P = ctypes.CDLL(PHASES_PATH)
    vp = ctypes.c_void_p
    P.rgp_car_pretick.argtypes = [vp, vp, vp, vp, ctypes.c_int32, vp, vp]
    n = len(cars)
    dev = {k: torch.from_numpy(np.ascontiguousarray(v)).to("cuda")
           for k, v in (("cars", cars), ("controls", controls), ("configs", configs), ("wheels", wheels))}
    unsupported = torch.zeros(n, dtype=torch.uint8, device="cuda")
    rc = P.rgp_car_pretick(dev["cars"].data_ptr(), dev["controls"].data_ptr(), dev["configs"].data_ptr(),
                           dev["wheels"].data_ptr(), n, unsupported.data_ptr(), None)
    assert rc == 0
    torch.cuda.synchronize()
    return dev["cars"].cpu().numpy(), dev["wheels"].cpu().numpy(), int(unsupported.sum())


@pytest.mark.gpu
def test_gpu_matches_fixture():
    f = load_fixture(FIX)
    c, w, unsupported = gpu_pretick(f["cars"], f["controls"], f["configs"], f["wheels"])
    assert unsupported == 0
    bad = np.flatnonzero((c != f["out_cars"]).any(1) | (w != f["out_wheels"]).any(1))
    assert bad.size == 0, bad[:10]


@pytest.mark.gpu
def test_gpu_matches_oracle_on_fresh_car_arenas():
    R = oracle_or_skip()
    for n_env, cpe, ticks, seed in ((16, 2, 300, 201), (6, 6, 200, 202)):
        cars, wheels, controls, configs = R.car_arena_inputs(n_env, cpe, ticks, seed)
        rc, rw = R.car_pretick(cars, controls, configs, wheels)
        c, w, unsupported = gpu_pretick(cars, controls, configs, wheels)
        assert unsupported == 0
        bad = np.flatnonzero((c != rc).any(1) | (w != rw).any(1))
        assert bad.size == 0, (n_env, cpe, bad[:10])
