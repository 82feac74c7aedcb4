"""Car force phases of the controls step: the jump hold (UpdateJumpState, include/rocketgpu/detail/step/
raycast_wheels_and_contact_prep.hpp:3629-3652 with ContinueJump :3399-3431) and boost (ApplyBoost, :3252-3300), with
the forces queued on the car's frame (the batch form) or applied to the velocity at once.  Checked byte for byte
against the reference's own functions: the committed fixture (six cases: each phase with and without a queue, boost
also with recharge on), the live oracle, the host twin on the CPU and the gfx950 kernel on the GPU.  The forces are
also pinned against the textbook rotation matrix of the car's quaternion, and the boost bookkeeping against its
closed forms."""
import ctypes
import os

import numpy as np
import pytest

from conftest import REPO, load_fixture, oracle_or_skip, twin

FIX = "car_force_phases"
PHASES_PATH = os.path.join(REPO, "rocketgpu_amd", "_build", "librg_phases.so")
QUEUE_BYTES = 156


def layout():
    import rocketgpu_amd as rg
    return {f[0]: f[2] for f in rg.state_layouts()["car_states"]["fields"]}


def fixture():
    f = load_fixture(FIX)
    out_cars = f["out_cars_xor"] ^ f["cars"][None]
    return f, out_cars


def f32(c, off, n=1):
    return c[:, off:off + 4 * n].view(np.float32)


def queue_parts(q):
    qf, qi = q.view(np.float32), q.view(np.int32)
    return {"force": qf[:, 0:3], "jump_force": qf[:, 9:12], "has_jump": qi[:, 12], "n_force": qi[:, 13],
            "n_torque": qi[:, 14], "force_items": qf[:, 15:27].reshape(-1, 4, 3)}


def rotation_matrix(rot):
    """Rows of the rotation matrices of unit quaternions (w, x, y, z), in float64 (the textbook form)."""
    w, x, y, z = (rot[:, i].astype(np.float64) for i in range(4))
    return np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)], 1),
        np.stack([2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)], 1),
        np.stack([2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)], 1)], 1)


def twin_phase(T, phase, cars, controls, queues, params, configs=None, flags=3):
    c = np.ascontiguousarray(cars).copy()
    q = None if queues is None else np.ascontiguousarray(queues).copy()
    k = np.ascontiguousarray(controls)
    g = None if configs is None else np.ascontiguousarray(configs)
    p = np.ascontiguousarray(params, dtype=np.float32)
    assert T.rgt_car_force_phase(len(c), c.ctypes.data, k.ctypes.data, None if g is None else g.ctypes.data,
                                 None if q is None else q.ctypes.data, phase, flags, p.ctypes.data) == 0
    return c, q


def test_fixture_covers_the_branches():
    f, out = fixture()
    lay = layout()
    c = f["cars"]
    cases = [tuple(x) for x in f["cases"]]
    alive = c[:, lay["is_demoed"]] == 0
    # Jump hold (queued): jumps that go on, jumps that end, grounded cars whose jump is cleared.
    j = out[cases.index((0, 1, 0, 3))]
    jq = queue_parts(f["out_queues"][cases.index((0, 1, 0, 3))])
    was = c[:, lay["is_jumping"]] == 1
    assert ((jq["has_jump"] == 1) & (f["queues"].view(np.int32)[:, 12] == 0)).sum() >= 10
    assert (was & (j[:, lay["is_jumping"]] == 0)).sum() >= 10
    assert ((c[:, lay["has_jumped"]] == 1) & (j[:, lay["has_jumped"]] == 0)).sum() >= 3
    # Boost (queued, defaults): pushes, latches that expire, tanks that run dry, clamps.
    b = out[cases.index((1, 1, 0, 3))]
    bq = queue_parts(f["out_queues"][cases.index((1, 1, 0, 3))])
    pushed = bq["n_force"] > f["queues"].view(np.int32)[:, 13]
    assert (pushed & alive).sum() >= 30
    assert ((c[:, lay["is_boosting"]] == 1) & (b[:, lay["is_boosting"]] == 0) & alive).sum() >= 10
    assert ((f32(c, lay["boost"])[:, 0] > 100) & alive).sum() >= 5
    assert ((f32(c, lay["boost"])[:, 0] < 0) & alive).sum() >= 5
    # Recharge on: some idle cars refill.
    r = out[cases.index((1, 1, 1, 3))]
    assert ((f32(r, lay["boost"])[:, 0] > f32(b, lay["boost"])[:, 0]) & alive).sum() >= 10
    assert (~alive).sum() >= 5
    # Air control (queued, full): pitch locked by a flip or its restart time with pitch input, authoritative bases
    # (no angular acceleration), restored inertia rows, and both torque gates.
    a0 = queue_parts(f["queues"])
    a1 = queue_parts(f["out_queues"][cases.index((2, 1, 0, 3))])
    qf0, qf1 = f["queues"].view(np.float32), f["out_queues"][cases.index((2, 1, 0, 3))].view(np.float32)
    turned = (qf1[:, 3:6] != qf0[:, 3:6]).any(1)
    auth = c[:, lay["basis_authoritative"]] == 1
    assert turned[~auth].sum() >= 100 and not turned[auth].any()
    restart = f32(c, lay["cal_pre_pitchlock_restart_time"])[:, 0]
    pitch = f["controls"][:, 8:12].view(np.float32)[:, 0]
    locked = (c[:, lay["is_flipping"]] == 1) | ((restart >= 0) & (restart < 0.3)) | \
        ((c[:, lay["has_flipped"]] == 1) & (f32(c, lay["internal_flip_time"])[:, 0] < 0.95))
    assert (locked & (pitch != 0)).sum() >= 20 and (~locked & (pitch != 0)).sum() >= 20
    assert (c[:, lay["restored_inv_inertia_valid"]] == 1).sum() >= 5
    assert ((a1["n_torque"] - a0["n_torque"]) == 2).sum() >= 50
    assert ((a1["n_torque"] - a0["n_torque"]) == 1).sum() >= 1
    # Flip torque (queued): flips that run on and queue a torque, flips that end, cancelling pitch.
    fl = out[cases.index((3, 1, 0, 1))]
    fq = queue_parts(f["out_queues"][cases.index((3, 1, 0, 1))])
    on = fl[:, lay["is_flipping"]] == 1
    assert on.sum() >= 10 and ((c[:, lay["is_flipping"]] == 1) & ~on).sum() >= 5
    assert (fq["n_torque"][on] == a0["n_torque"][on] + 1).sum() >= 8
    late = on & (f32(c, lay["internal_flip_time"])[:, 0] >= 0.041) & (pitch != 0)
    assert late.sum() >= 3


def test_oracle_reproduces_fixture():
    R = oracle_or_skip()
    f, out = fixture()
    for i, (phase, queued, ps, flags) in enumerate(f["cases"]):
        c, q = R.car_force_phase(int(phase), f["cars"], f["controls"], f["queues"] if queued else None,
                                 f["params"][ps], f["configs"], int(flags))
        assert np.array_equal(c, out[i]), i
        if queued:
            assert np.array_equal(q, f["out_queues"][i]), i


def test_oracle_default_parameters_match_the_twin():
    R = oracle_or_skip()
    f = load_fixture(FIX)
    assert np.array_equal(R.force_params(), f["params"][0])


@pytest.mark.parametrize("case", range(13))
def test_twin_matches_fixture(case):
    T = twin()
    f, out = fixture()
    phase, queued, ps, flags = (int(x) for x in f["cases"][case])
    c, q = twin_phase(T, phase, f["cars"], f["controls"], f["queues"] if queued else None, f["params"][ps],
                      f["configs"], flags)
    bad = np.flatnonzero((c != out[case]).any(1))
    assert bad.size == 0, bad[:10]
    if queued:
        bad = np.flatnonzero((q != f["out_queues"][case]).any(1))
        assert bad.size == 0, bad[:10]


def test_boost_force_is_the_forward_axis():
    """A car that boosts queues mass * boost_accel_ground / 50 along the first column of its rotation matrix."""
    f, out = fixture()
    lay = layout()
    cases = [tuple(x) for x in f["cases"]]
    p = dict(zip(("tick_dt", "car_mass", "jump_min_time", "jump_reset_time_pad", "jump_max_time", "jump_accel",
                  "boost_max", "boost_used_per_second", "boost_min_time", "boost_accel_ground"), f["params"][0]))
    q0 = queue_parts(f["queues"])
    q1 = queue_parts(f["out_queues"][cases.index((1, 1, 0, 3))])
    rows = np.flatnonzero(q1["n_force"] == q0["n_force"] + 1)
    assert rows.size >= 30
    delta = q1["force"][rows].astype(np.float64) - q0["force"][rows]
    rot = f32(f["cars"], lay["rot"], 4)[rows]
    want = rotation_matrix(rot)[:, :, 0] * (p["boost_accel_ground"] * p["car_mass"] / 50.0)
    assert np.allclose(delta, want, rtol=0, atol=2e-3)
    # The item slot takes the force while the queue holds fewer than four.
    few = q0["n_force"][rows] < 4
    items = q1["force_items"][rows[few], q0["n_force"][rows[few]]]
    # float32 rounding is relative to the whole vector (|F| ~ 3570), so the tolerance is absolute.
    assert np.allclose(items, want[few], rtol=0, atol=2e-3)


def test_jump_force_is_the_up_axis():
    """A jump that goes on sets the queued jump force to mass * jump_accel / 50 along the third column."""
    f, out = fixture()
    lay = layout()
    cases = [tuple(x) for x in f["cases"]]
    p = f["params"][0]
    j = out[cases.index((0, 1, 0, 3))]
    q1 = queue_parts(f["out_queues"][cases.index((0, 1, 0, 3))])
    rows = np.flatnonzero(j[:, lay["is_jumping"]] == 1)
    assert rows.size >= 10
    assert (q1["has_jump"][rows] == 1).all()
    rot = f32(f["cars"], lay["rot"], 4)[rows]
    want = rotation_matrix(rot)[:, :, 2] * (p[5] * p[1] / 50.0)
    assert np.allclose(q1["jump_force"][rows], want, rtol=0, atol=3e-3)  # |F| ~ 5250


def test_unqueued_boost_changes_the_velocity():
    """Without a queue a boosting car's velocity gains boost_accel_ground * dt along its forward axis (game units),
    its Bullet-unit copy follows, and cars that do not push keep their velocity."""
    f, out = fixture()
    lay = layout()
    cases = [tuple(x) for x in f["cases"]]
    p = f["params"][0]
    b = out[cases.index((1, 0, 0, 3))]
    bq = queue_parts(f["out_queues"][cases.index((1, 1, 0, 3))])
    pushed = (bq["n_force"] > queue_parts(f["queues"])["n_force"]) & (f["cars"][:, lay["is_demoed"]] == 0)
    v0 = f32(f["cars"], lay["vel"], 3).astype(np.float64)
    v1 = f32(b, lay["vel"], 3)
    fwd = rotation_matrix(f32(f["cars"], lay["rot"], 4))[:, :, 0]
    assert np.allclose(v1[pushed] - v0[pushed], fwd[pushed] * (p[9] * p[0]), rtol=0, atol=2e-3)
    assert np.array_equal(f32(b, lay["vel_bt"], 3)[pushed], (v1[pushed] * np.float32(0.02)).astype(np.float32))
    assert np.array_equal(v1[~pushed], v0[~pushed].astype(np.float32))


def test_boost_spend_and_recharge():
    """A car that keeps boosting spends boost_used_per_second * dt; with recharge on an idle car past the delay
    gains recharge_boost_per_second * dt; every live car ends within 0..boost_max."""
    f, out = fixture()
    lay = layout()
    cases = [tuple(x) for x in f["cases"]]
    p = f["params"][1]
    dt = float(p[0])
    c = f["cars"]
    alive = c[:, lay["is_demoed"]] == 0
    b0 = f32(c, lay["boost"])[:, 0].astype(np.float64)
    frac_ok = f32(c, lay["boost_amount"])[:, 0] * np.float32(100) == f32(c, lay["boost"])[:, 0]
    for case in ((1, 1, 0, 3), (1, 1, 1, 3)):
        b1 = f32(out[cases.index(case)], lay["boost"])[:, 0]
        assert ((b1[alive] >= 0) & (b1[alive] <= 100)).all()
    b1 = f32(out[cases.index((1, 1, 1, 3))], lay["boost"])[:, 0].astype(np.float64)
    keep = alive & frac_ok & (c[:, lay["is_boosting"]] == 1) & (b0 > 1) & (b0 <= 100) & \
        (out[cases.index((1, 1, 1, 3))][:, lay["is_boosting"]] == 1)
    assert keep.sum() >= 5
    assert np.allclose(b1[keep], b0[keep] - p[7] * dt, rtol=0, atol=2e-4)
    t1 = f32(out[cases.index((1, 1, 1, 3))], lay["time_since_boosted"])[:, 0]
    idle = alive & frac_ok & (c[:, lay["is_boosting"]] == 0) & (t1 >= p[11]) & (b0 >= 0) & (b0 < 99)
    idle &= out[cases.index((1, 1, 1, 3))][:, lay["is_boosting"]] == 0
    assert idle.sum() >= 5
    assert np.allclose(b1[idle], b0[idle] + p[10] * dt, rtol=0, atol=2e-4)


def test_twin_matches_oracle_on_fresh_inputs():
    """Fresh perturbations of live car-arena states, under the defaults, with recharge on and with other masses,
    accelerations and timings, every phase with and without a queue."""
    R = oracle_or_skip()
    T = twin()
    from tests.golden.make_fixtures import perturb_force_inputs
    run = R.car_arena_run(4, 4, 60, 411)
    sets = [R.force_params(), R.force_params(recharge_boost_enabled=1.0, recharge_boost_delay=0.1),
            R.force_params(car_mass=200.0, boost_accel_ground=1000.0, jump_accel=1500.0, boost_min_time=0.2,
                           jump_max_time=0.1, boost_max=80.0, recharge_boost_enabled=1.0,
                           car_air_throttle_accel=80.0, flip_pitchlock_extra_time=0.5, car_air_control_pitch=100.0,
                           car_air_control_yaw=120.0, car_air_control_roll=300.0, car_air_damp_pitch=10.0,
                           car_air_damp_yaw=40.0, car_air_damp_roll=60.0)]
    for t in range(0, 60, 6):
        mid, _ = R.car_pretick(4, 4, 411, run["cars"][t], run["controls"][t], run["configs"], run["wheels"][t],
                               run["envs"][t], run["teams"])
        c, k, q = perturb_force_inputs(mid, run["controls"][t], 500 + t)
        for ps in sets:
            for phase, flags in ((0, 3), (1, 3), (2, 3), (2, 1), (2, 2), (2, 0), (3, 1), (3, 0)):
                for queues in (q, None):
                    wc, wq = R.car_force_phase(phase, c, k, queues, ps, run["configs"], flags)
                    gc, gq = twin_phase(T, phase, c, k, queues, ps, run["configs"], flags)
                    bad = np.flatnonzero((gc != wc).any(1))
                    assert bad.size == 0, (t, phase, flags, queues is None, bad[:10])
                    if queues is not None:
                        assert np.array_equal(gq, wq), (t, phase, flags)


def test_air_throttle_and_direct_air_control():
    """Air throttle alone (flags 2, queued) queues mass * car_air_throttle_accel / 50 along the first column of the
    rotation matrix, scaled by the throttle (1 while boosting).  Air control alone (flags 1) queues no force.  The
    direct form (no queue) leaves the linear velocity of a car with no throttle and no boost alone and turns the
    angular velocity."""
    f, out = fixture()
    lay = layout()
    cases = [tuple(x) for x in f["cases"]]
    p = f["params"][0]
    c = f["cars"]
    kc = f["controls"].view([("throttle", "<f4"), ("steer", "<f4"), ("pitch", "<f4"), ("yaw", "<f4"),
                             ("roll", "<f4"), ("jump", "u1"), ("boost", "u1"), ("handbrake", "u1"),
                             ("pad", "u1")]).reshape(-1)
    thr = np.where(c[:, lay["is_boosting"]] == 1, 1.0, kc["throttle"].astype(np.float64))
    q0 = queue_parts(f["queues"])
    q1 = queue_parts(f["out_queues"][cases.index((2, 1, 0, 2))])
    rows = np.flatnonzero(thr != 0)
    assert rows.size >= 30
    assert (q1["n_force"][rows] == q0["n_force"][rows] + 1).all()
    assert (q1["n_force"][thr == 0] == q0["n_force"][thr == 0]).all()
    fwd = rotation_matrix(f32(c, lay["rot"], 4))[:, :, 0]
    want = fwd[rows] * (p[13] * p[1] / 50.0) * thr[rows, None]
    delta = q1["force"][rows].astype(np.float64) - q0["force"][rows]
    assert np.allclose(delta, want, rtol=0, atol=2e-3)
    # Update only (flags 1): no force is queued, and a torque item per gate that opens.
    q2 = queue_parts(f["out_queues"][cases.index((2, 1, 0, 1))])
    assert (q2["n_force"] == q0["n_force"]).all()
    assert ((q2["n_torque"] - q0["n_torque"]) >= 1).sum() >= 100
    # Direct form: a car with no throttle input and no boost keeps its linear velocity; angular velocity changes.
    d = out[cases.index((2, 0, 0, 3))]
    idle = thr == 0
    assert np.array_equal(f32(d, lay["vel"], 3)[idle], f32(c, lay["vel"], 3)[idle])
    assert (f32(d, lay["ang_vel"], 3) != f32(c, lay["ang_vel"], 3)).any(1).sum() >= 100


def test_flip_torque_turns_the_car():
    """A flip that runs on: without a queue the angular velocity gains R(q) (rint(x * flip_torque_x),
    rint(y * flip_torque_y), 0) dt for the relative torque (x, y) (rows with no cancelling pitch: under 0.041 s of
    flip, zero pitch or zero y), with R the car's basis columns when they are authoritative;
    a flip past its torque time ends and leaves the car alone."""
    f, out = fixture()
    lay = layout()
    cases = [tuple(x) for x in f["cases"]]
    p = f["params"][0]
    c = f["cars"]
    d = out[cases.index((3, 0, 0, 1))]
    ift = f32(c, lay["internal_flip_time"])[:, 0]
    running = (c[:, lay["is_flipping"]] == 1) & (c[:, lay["has_flipped"]] == 1) & (ift < p[14])
    assert (d[:, lay["is_flipping"]] == running).all()
    ended = (c[:, lay["is_flipping"]] == 1) & ~running
    assert ended.sum() >= 5 and np.array_equal(d[ended, lay["ang_vel"]:lay["ang_vel"] + 12],
                                                c[ended, lay["ang_vel"]:lay["ang_vel"] + 12])
    rel_all = f32(c, lay["flip_rel_torque"], 3)
    pitch = f["controls"][:, 8:12].view(np.float32)[:, 0]
    no_cancel = (ift < 0.041) | (pitch == 0) | (rel_all[:, 1] == 0)
    rows = running & no_cancel
    assert rows.sum() >= 5 and (rows & (c[:, lay["basis_authoritative"]] == 0)).sum() >= 3
    rel = rel_all[rows].astype(np.float32)
    local = np.stack([np.rint(rel[:, 0] * np.float32(p[22])), np.rint(rel[:, 1] * np.float32(p[23])),
                      np.zeros(rows.sum())], 1)
    # Quaternion cars turn by the rotation matrix, cars with an authoritative basis by their basis columns.
    m = rotation_matrix(f32(c, lay["rot"], 4)[rows])
    basis = np.stack([f32(c, lay[k], 3)[rows] for k in ("basis_forward", "basis_right", "basis_up")], 2)
    auth = c[rows, lay["basis_authoritative"]] == 1
    m[auth] = basis[auth]
    want = np.einsum("nij,nj->ni", m, local) * p[0]
    got = f32(d, lay["ang_vel"], 3)[rows].astype(np.float64) - f32(c, lay["ang_vel"], 3)[rows]
    assert np.allclose(got, want, rtol=0, atol=1e-4)


def test_twin_rejects_an_unknown_phase():
    T = twin()
    f = load_fixture(FIX)
    c = f["cars"][:1].copy()
    p = f["params"][0].copy()
    k = f["controls"][:1].copy()
    assert T.rgt_car_force_phase(1, c.ctypes.data, k.ctypes.data, None, None, 4, 3, p.ctypes.data) == 1
    assert T.rgt_car_force_phase(1, c.ctypes.data, k.ctypes.data, None, None, 3, 1, p.ctypes.data) == 1
    assert T.rgt_car_force_phase(1, c.ctypes.data, k.ctypes.data, None, None, 2, 3, p.ctypes.data) == 1
    assert np.array_equal(c, f["cars"][:1])


def test_phase_library_exports_every_entry_point():
    """The test-only phase library loads here and exports each phase's entry point (nothing is called)."""
    if not os.path.exists(PHASES_PATH):
        pytest.fail("librg_phases.so is not built")
    P = ctypes.CDLL(PHASES_PATH)
    for name in ("rgp_car_pretick", "rgp_car_posttick", "rgp_car_force_phase"):
        assert hasattr(P, name), name


def gpu_lib():
    if not os.path.exists(PHASES_PATH):
        pytest.fail("librg_phases.so is not built")
    P = ctypes.CDLL(PHASES_PATH)
    vp = ctypes.c_void_p
    P.rgp_car_force_phase.argtypes = [vp, vp, vp, vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, vp, vp]
    return P


@pytest.mark.gpu
@pytest.mark.parametrize("case", range(13))
def test_gpu_matches_fixture(case):
    import torch
    P = gpu_lib()
    f, out = fixture()
    phase, queued, ps, flags = (int(x) for x in f["cases"][case])
    n = len(f["cars"])
    cars = torch.from_numpy(f["cars"].copy()).to("cuda")
    controls = torch.from_numpy(f["controls"].copy()).to("cuda")
    configs = torch.from_numpy(f["configs"].copy()).to("cuda")
    queues = torch.from_numpy(f["queues"].copy()).to("cuda") if queued else None
    params = np.ascontiguousarray(f["params"][ps], dtype=np.float32)
    assert P.rgp_car_force_phase(cars.data_ptr(), controls.data_ptr(), configs.data_ptr(),
                                 None if queues is None else queues.data_ptr(), n, phase, flags, params.ctypes.data,
                                 None) == 0
    torch.cuda.synchronize()
    bad = np.flatnonzero((cars.cpu().numpy() != out[case]).any(1))
    assert bad.size == 0, bad[:10]
    if queued:
        bad = np.flatnonzero((queues.cpu().numpy() != f["out_queues"][case]).any(1))
        assert bad.size == 0, bad[:10]


@pytest.mark.gpu
def test_gpu_rejects_an_unknown_phase():
    P = gpu_lib()
    p = np.zeros(24, np.float32)
    assert P.rgp_car_force_phase(None, None, None, None, 1, 4, 3, p.ctypes.data, None) != 0
    assert P.rgp_car_force_phase(None, None, None, None, 1, 2, 3, p.ctypes.data, None) != 0
    assert P.rgp_car_force_phase(None, None, None, None, 1, 3, 1, p.ctypes.data, None) != 0
    assert P.rgp_car_force_phase(None, None, None, None, 1, 0, 3, None, None) != 0
