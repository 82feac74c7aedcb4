"""Car post-tick (StepEnvPostCar after the world step, include/rocketgpu/detail/step/env_step.hpp:6902-6926):
cooldowns and last controls, pending bumps and velocity caches, the tick counter, and the reported fields, with
and without the report.  Checked byte for byte against the reference's own function: the committed fixture (car-arena
states with seeded bumps, caches and demoed cars), the live oracle, the host twin on the CPU and the gfx950 kernel on
the GPU."""
import ctypes
import os

import numpy as np
import pytest

from conftest import REPO, load_fixture, oracle_or_skip, twin

FIX = "car_posttick_1v1_3v3"
PHASES_PATH = os.path.join(REPO, "rocketgpu_amd", "_build", "librg_phases.so")


def layout():
    import rocketgpu_amd as rg
    return {f[0]: f[2] for f in rg.state_layouts()["car_states"]["fields"]}


def twin_posttick(T, cars, controls, wheels, report):
    c = np.ascontiguousarray(cars).copy()
    k = np.ascontiguousarray(controls)
    w = np.ascontiguousarray(wheels)
    T.rgt_car_posttick(len(c), c.ctypes.data, k.ctypes.data, w.ctypes.data, 1 if report else 0)
    return c


def test_fixture_covers_the_branches():
    f = load_fixture(FIX)
    lay = layout()
    c = f["cars"]
    assert (c[:, lay["pending_bump_count"]] > 0).sum() >= 40
    cache = c[:, lay["vel_impulse_cache"]:lay["vel_impulse_cache"] + 12].view(np.float32)
    assert (np.abs(cache).sum(1) > 0).sum() >= 40
    assert c[:, lay["is_demoed"]].sum() >= 3


def test_oracle_reproduces_fixture():
    R = oracle_or_skip()
    f = load_fixture(FIX)
    assert np.array_equal(R.car_posttick(f["cars"], f["controls"], f["wheels"], True), f["out_report"])
    assert np.array_equal(R.car_posttick(f["cars"], f["controls"], f["wheels"], False), f["out_advance"])


@pytest.mark.parametrize("report", [True, False])
def test_twin_matches_fixture(report):
    T = twin()
    f = load_fixture(FIX)
    c = twin_posttick(T, f["cars"], f["controls"], f["wheels"], report)
    want = f["out_report"] if report else f["out_advance"]
    bad = np.flatnonzero((c != want).any(1))
    assert bad.size == 0, bad[:10]


def test_pending_bumps_sum_into_the_velocity():
    """A car with k pending bumps leaves the post-tick with vel_bt = vel_bt + the k bumps (summed in order), no
    bumps and no velocity caches, and its tick counter advanced by one."""
    f = load_fixture(FIX)
    lay = layout()
    c, out = f["cars"], f["out_report"]
    rows = np.flatnonzero((c[:, lay["pending_bump_count"]] > 0) & (c[:, lay["is_demoed"]] == 0))
    assert rows.size > 0
    vb = lay["vel_bt"]
    for i in rows:
        k = int(c[i, lay["pending_bump_count"]])
        v = c[i, vb:vb + 12].view(np.float32).copy()
        bumps = c[i, lay["pending_bump_bt"]:lay["pending_bump_bt"] + 84].view(np.float32).reshape(7, 3)
        for j in range(k):
            v = (v + bumps[j]).astype(np.float32)
        assert np.array_equal(out[i, vb:vb + 12].view(np.float32), v)
        assert out[i, lay["pending_bump_count"]] == 0
        t0 = c[i, lay["tick_count_since_update"]:lay["tick_count_since_update"] + 8].view(np.uint64)[0]
        t1 = out[i, lay["tick_count_since_update"]:lay["tick_count_since_update"] + 8].view(np.uint64)[0]
        assert t1 == t0 + 1


def test_twin_matches_oracle_on_fresh_car_arenas():
    R = oracle_or_skip()
    T = twin()
    from tests.golden.make_fixtures import perturb_post_inputs
    run = R.car_arena_run(6, 2, 150, 301)
    for t in range(0, 150, 5):
        mid, mid_wheels = R.car_pretick(6, 2, 301, run["cars"][t], run["controls"][t], run["configs"],
                                        run["wheels"][t], run["envs"][t], run["teams"])
        inputs = perturb_post_inputs(mid, 302 + t)
        for report in (True, False):
            want = R.car_posttick(inputs, run["controls"][t], mid_wheels, report)
            got = twin_posttick(T, inputs, run["controls"][t], mid_wheels, report)
            bad = np.flatnonzero((got != want).any(1))
            assert bad.size == 0, (t, report, bad[:10])


@pytest.mark.gpu
@pytest.mark.parametrize("report", [True, False])
def test_gpu_matches_fixture(report):
    import torch
    if not os.path.exists(PHASES_PATH):
        pytest.fail("librg_phases.so is not built")
    P = ctypes.CDLL(PHASES_PATH)
    vp = ctypes.c_void_p
    P.rgp_car_posttick.argtypes = [vp, vp, vp, ctypes.c_int32, ctypes.c_int32, vp]
    f = load_fixture(FIX)
    n = len(f["cars"])
    cars = torch.from_numpy(f["cars"].copy()).to("cuda")
    controls = torch.from_numpy(f["controls"]).to("cuda")
    wheels = torch.from_numpy(f["wheels"]).to("cuda")
    assert P.rgp_car_posttick(cars.data_ptr(), controls.data_ptr(), wheels.data_ptr(), n, 1 if report else 0,
                              None) == 0
    torch.cuda.synchronize()
    want = f["out_report"] if report else f["out_advance"]
    bad = np.flatnonzero((cars.cpu().numpy() != want).any(1))
    assert bad.size == 0, bad[:10]
