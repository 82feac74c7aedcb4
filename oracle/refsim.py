"""ctypes wrapper of the oracle: the reference CPU BatchSim built by oracle/Makefile into oracle/_ref/librgref.so.

TEST INFRASTRUCTURE ONLY.  Imported by tests/, tests/golden/make_fixtures.py, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker / CPU baseline; never by the product package.

The reference API it mirrors is include/rocketgpu/rocketgpu.h:988-1089 (BatchSim); the shim is oracle/ref_shim.cpp.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_ref", "librgref.so")

CAR, BALL, ENV, PAD, CONTROLS = 0, 1, 2, 3, 4

_lib = None


def available() -> bool:
    return os.path.exists(LIB_PATH)


def lib():
    global _lib
    if _lib is None:
        if not available():
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run `make -C oracle`)")
        L = ctypes.CDLL(LIB_PATH)
        L.rgref_last_error.restype = ctypes.c_char_p
        L.rgref_sizeof.restype = ctypes.c_long
        L.rgref_sizeof.argtypes = [ctypes.c_int]
        L.rgref_offsetof.restype = ctypes.c_long
        L.rgref_offsetof.argtypes = [ctypes.c_int, ctypes.c_char_p]
        L.rgref_create.restype = ctypes.c_void_p
        L.rgref_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_uint64]
        L.rgref_create_flags.restype = ctypes.c_void_p
        L.rgref_create_flags.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_uint64,
                                         ctypes.c_int]
        L.rgref_destroy.argtypes = [ctypes.c_void_p]
        L.rgref_set_mutator_floats.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
        vp, cp, lg = ctypes.c_void_p, ctypes.c_void_p, ctypes.c_long
        for name, args in [
            ("rgref_set_env_car_count", [vp, cp, lg]),
            ("rgref_set_teams", [vp, cp, lg]),
            ("rgref_set_car_presets", [vp, cp, lg]),
            ("rgref_reset_to_kickoff", [vp, ctypes.c_uint64]),
            ("rgref_upload", [vp, ctypes.c_int, cp, lg]),
            ("rgref_download", [vp, ctypes.c_int, cp, lg]),
            ("rgref_apply_env_commands", [vp, cp, lg]),
            ("rgref_step", [vp, ctypes.c_int]),
            ("rgref_step_with_controls", [vp, cp, ctypes.c_int, ctypes.c_int]),
            ("rgref_step_with_commands", [vp, cp, ctypes.c_int, cp, ctypes.c_int, ctypes.c_int]),
            ("rgref_apply_env_commands_device", [vp, cp, lg]),
            ("rgref_env_digests", [vp, cp, cp]),
            ("rgref_set_event_stream", [vp, cp, cp, ctypes.c_uint32, ctypes.c_int]),
            ("rgref_download_wheels", [vp, cp, lg]),
            ("rgref_download_car_configs", [vp, cp, lg]),
            ("rgref_car_pretick_batch", [vp, ctypes.c_uint64, cp, cp, cp, cp, cp, cp]),
            ("rgref_car_params", [cp]),
            ("rgref_car_posttick", [lg, cp, cp, cp, ctypes.c_int]),
            ("rgref_car_force_phase", [lg, cp, cp, cp, cp, ctypes.c_int, ctypes.c_int, cp]),
        ]:
            f = getattr(L, name)
            f.argtypes = args
            f.restype = ctypes.c_int
        L.rgref_ball_path.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_float, cp, cp]
        L.rgref_ball_path.restype = ctypes.c_int
        L.rgref_default_state.argtypes = [ctypes.c_int, cp]
        L.rgref_workload_ball_states.argtypes = [ctypes.c_uint64, ctypes.c_long, ctypes.c_long, ctypes.c_double, cp]
        L.rgref_math.argtypes = [ctypes.c_int, cp, cp, cp, ctypes.c_long]
        L.rgref_mesh_tables.argtypes = [vp, cp, cp, cp, cp, cp]
        L.rgref_quantize_axis.argtypes = [cp, cp, lg]
        L.rgref_quantize_axis.restype = None
        L.rgref_default_wheel_state.argtypes = [cp]
        L.rgref_default_wheel_state.restype = None
        L.rgref_run.restype = ctypes.c_long
        L.rgref_run.argtypes = [vp, cp, ctypes.c_uint64, ctypes.c_float, ctypes.c_int, cp, ctypes.c_int, cp]
        _lib = L
    return _lib


def sizeof(which: int) -> int:
    return int(lib().rgref_sizeof(which))


def offsetof(which: int, field: str) -> int:
    return int(lib().rgref_offsetof(which, field.encode()))


def default_state(which: int) -> np.ndarray:
    out = np.zeros(sizeof(which), dtype=np.uint8)
    if lib().rgref_default_state(which, out.ctypes.data) != 0:
        raise ValueError(which)
    return out


def workload_ball_states(seed: int, first_arena: int, count: int, scale: float = 1.0) -> np.ndarray:
    out = np.zeros((count, sizeof(BALL)), dtype=np.uint8)
    lib().rgref_workload_ball_states(seed, first_arena, count, scale, out.ctypes.data)
    return out


def math(fn: int, x, y=None) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.ascontiguousarray(x if y is None else y, dtype=np.float32)
    out = np.zeros_like(x)
    lib().rgref_math(fn, x.ctypes.data, y.ctypes.data, out.ctypes.data, x.size)
    return out


def _ptr(a: np.ndarray):
    return ctypes.c_void_p(a.ctypes.data)


class RefSim:
    """One reference BatchSim(cfg, procedural layout)."""

    def __init__(self, env_count: int, max_cars: int, scale: float = 1.0, fillet_radius: float = 256.0,
                 respawn_seed: int = 0, flags: int = 15):
        """flags: bit 0 boost pads, bit 1 car-car collisions, bit 2 demos, bit 3 scoring."""
        self._L = lib()
        self.h = self._L.rgref_create_flags(env_count, max_cars, scale, fillet_radius, respawn_seed, flags)
        if not self.h:
            raise RuntimeError(self._L.rgref_last_error().decode())
        self.env_count = env_count
        self.max_cars = max_cars

    def close(self):
        if self.h:
            self._L.rgref_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _check(self, rc):
        if rc != 0:
            raise RuntimeError(self._L.rgref_last_error().decode())

    def set_env_car_count(self, counts):
        a = np.ascontiguousarray(counts, dtype=np.int32)
        self._check(self._L.rgref_set_env_car_count(self.h, _ptr(a), a.size))

    def set_teams(self, teams):
        a = np.ascontiguousarray(teams, dtype=np.uint8)
        self._check(self._L.rgref_set_teams(self.h, _ptr(a), a.size))

    def set_car_presets(self, presets):
        a = np.ascontiguousarray(presets, dtype=np.uint8)
        self._check(self._L.rgref_set_car_presets(self.h, _ptr(a), a.size))

    def set_mutator_floats(self, **fields):
        """set_mutator_config with the named float fields of the current MutatorConfig replaced."""
        names = (ctypes.c_char_p * len(fields))(*[k.encode() for k in fields])
        vals = np.array(list(fields.values()), np.float32)
        self._check(self._L.rgref_set_mutator_floats(self.h, names, vals.ctypes.data, len(fields)))

    def reset_to_kickoff(self, seed: int = 0):
        self._check(self._L.rgref_reset_to_kickoff(self.h, seed))

    def download(self, which: int, count: int) -> np.ndarray:
        out = np.zeros((count, sizeof(which)), dtype=np.uint8)
        self._check(self._L.rgref_download(self.h, which, _ptr(out), count))
        return out

    def upload(self, which: int, rows: np.ndarray):
        rows = np.ascontiguousarray(rows, dtype=np.uint8)
        assert rows.shape[1] == sizeof(which)
        self._check(self._L.rgref_upload(self.h, which, _ptr(rows), rows.shape[0]))

    def apply_env_commands(self, cmds: np.ndarray):
        cmds = np.ascontiguousarray(cmds, dtype=np.uint8)
        self._check(self._L.rgref_apply_env_commands(self.h, _ptr(cmds), cmds.shape[0]))

    def apply_env_commands_device(self, cmds: np.ndarray):
        cmds = np.ascontiguousarray(cmds, dtype=np.uint8)
        self._check(self._L.rgref_apply_env_commands_device(self.h, _ptr(cmds), cmds.shape[0]))

    def download_wheels(self, count: int) -> np.ndarray:
        """BatchSim::download_wheel_runtime_states: [count, 292] u8."""
        out = np.zeros((count, WHEEL_STATE_BYTES), dtype=np.uint8)
        self._check(self._L.rgref_download_wheels(self.h, _ptr(out), count))
        return out

    def download_car_configs(self, count: int) -> np.ndarray:
        out = np.zeros((count, sizeof(5)), dtype=np.uint8)
        self._check(self._L.rgref_download_car_configs(self.h, _ptr(out), count))
        return out

    def set_event_stream(self, records, windows, capacity: int, goal_only: bool = False):
        """BatchSim::set_physics_event_stream with host buffers: records [env_count * capacity, 104] u8, windows
        [env_count, 16] u8 (numpy arrays the caller keeps alive); None/None/0 disables."""
        self._events = (records, windows)
        rp = _ptr(records) if records is not None else None
        wp = _ptr(windows) if windows is not None else None
        self._check(self._L.rgref_set_event_stream(self.h, rp, wp, capacity, 1 if goal_only else 0))

    def step(self, ticks: int = 1):
        self._check(self._L.rgref_step(self.h, ticks))

    def step_with_commands(self, controls: np.ndarray, control_stride: int, commands: np.ndarray,
                           command_stride: int, ticks: int):
        """BatchSim::step_with_device_commands: commands [ticks * command_stride, 16] u8 EnvCommand records."""
        controls = np.ascontiguousarray(controls, dtype=np.uint8)
        commands = np.ascontiguousarray(commands, dtype=np.uint8)
        self._check(self._L.rgref_step_with_commands(self.h, _ptr(controls), control_stride, _ptr(commands),
                                                     command_stride, ticks))

    def step_with_controls(self, controls: np.ndarray, ticks: int, stride: int):
        controls = np.ascontiguousarray(controls, dtype=np.uint8)
        self._check(self._L.rgref_step_with_controls(self.h, _ptr(controls), ticks, stride))

    def env_digests(self, car_counts) -> np.ndarray:
        a = np.ascontiguousarray(car_counts, dtype=np.int32)
        out = np.zeros(self.env_count, dtype=np.uint64)
        self._check(self._L.rgref_env_digests(self.h, _ptr(a), _ptr(out)))
        return out

    def ball_path(self, samples: int, ticks: int, goal_y: float):
        """BatchSim::ball_path: (out [env, samples, 6] f32 pos+vel, result [env, 2] i32 goal side and tick)."""
        out = np.zeros((self.env_count, samples, 6), np.float32)
        res = np.zeros((self.env_count, 2), np.int32)
        if self._L.rgref_ball_path(self.h, samples, ticks, goal_y, _ptr(out), _ptr(res)) != 0:
            raise RuntimeError(self._L.rgref_last_error().decode())
        return out, res

    def run(self, car_counts, ctrl_seed: int, ticks: int, checkpoint_every: int = 0, p_boost: float = 0.3):
        a = np.ascontiguousarray(car_counts, dtype=np.int32)
        batch = np.zeros(ticks, dtype=np.uint64)
        n_ck = ticks // checkpoint_every if checkpoint_every > 0 else 0
        env = np.zeros((max(n_ck, 1), self.env_count), dtype=np.uint64)
        got = self._L.rgref_run(self.h, _ptr(a), ctrl_seed, p_boost, ticks, _ptr(batch), checkpoint_every, _ptr(env))
        if got < 0:
            raise RuntimeError(self._L.rgref_last_error().decode())
        return batch, env[:got]


# ----------------------------------------------------------------------------------------------------------------
# Single phases of the reference step (oracle/ref_phases.cpp).

WHEEL_STATE_BYTES = 292
CONTROLS_DTYPE = np.dtype([("throttle", "<f4"), ("steer", "<f4"), ("pitch", "<f4"), ("yaw", "<f4"), ("roll", "<f4"),
                           ("jump", "u1"), ("boost", "u1"), ("handbrake", "u1"), ("pad", "u1")])


def car_params() -> np.ndarray:
    """The reference's default MutatorConfig values the car pre-tick reads (ref_phases.cpp lists the order)."""
    out = np.zeros(64, np.float32)
    n = lib().rgref_car_params(_ptr(out))
    return out[:n]


def quantize_axis(x) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.float32)
    out = np.zeros_like(x)
    lib().rgref_quantize_axis(_ptr(x), _ptr(out), x.size)
    return out


_phase_sims = {}


def car_pretick(n_env: int, max_cars: int, seed: int, cars, controls, configs, wheels, envs, teams):
    """StepEnvPreTickCar on every car slot of a batch (slot k of env e at e * max_cars + k) with the procedural
    arena's rules and the respawn seed: returns (cars, wheels) after it (inputs are not changed)."""
    key = (n_env, max_cars)
    if key not in _phase_sims:
        _phase_sims[key] = RefSim(n_env, max_cars)
    sim = _phase_sims[key]
    c = np.ascontiguousarray(cars, dtype=np.uint8).copy()
    w = np.ascontiguousarray(wheels, dtype=np.uint8).copy()
    args = [np.ascontiguousarray(a, dtype=np.uint8) for a in (controls, configs, envs, teams)]
    if lib().rgref_car_pretick_batch(sim.h, seed, _ptr(c), _ptr(args[0]), _ptr(args[1]), _ptr(w), _ptr(args[2]),
                                     _ptr(args[3])) != 0:
        raise RuntimeError(lib().rgref_last_error().decode())
    return c, w


def car_posttick(cars, controls, wheels, report: bool):
    """StepEnvPostCar without the world-side finish on independent cars: returns the cars after it."""
    c = np.ascontiguousarray(cars, dtype=np.uint8).copy()
    w = np.ascontiguousarray(wheels, dtype=np.uint8).copy()
    ctl = np.ascontiguousarray(controls, dtype=np.uint8)
    if lib().rgref_car_posttick(c.shape[0], _ptr(c), _ptr(ctl), _ptr(w), 1 if report else 0) != 0:
        raise RuntimeError("rgref_car_posttick failed")
    return c


FORCE_QUEUE_BYTES = 156
# The order of a force-phase parameter list (ref_phases.cpp rgref_car_force_phase).
FORCE_PARAM_NAMES = ("tick_dt", "car_mass", "jump_min_time", "jump_reset_time_pad", "jump_max_time", "jump_accel",
                     "boost_max", "boost_used_per_second", "boost_min_time", "boost_accel_ground",
                     "recharge_boost_per_second", "recharge_boost_delay", "recharge_boost_enabled",
                     "car_air_throttle_accel", "flip_torque_time", "flip_pitchlock_extra_time",
                     "car_air_control_pitch", "car_air_control_yaw", "car_air_control_roll", "car_air_damp_pitch",
                     "car_air_damp_yaw", "car_air_damp_roll", "flip_torque_x", "flip_torque_y")


# The order of car_params() (ref_phases.cpp rgref_car_params).
CAR_PARAM_NAMES = ("tick_dt", "boost_pad_frame_frac", "car_mass", "car_max_speed", "car_max_ang_speed",
                   "supersonic_start_speed", "supersonic_maintain_min_speed", "supersonic_maintain_max_time",
                   "jump_immediate_speed", "jump_min_time", "jump_reset_time_pad", "jump_max_time",
                   "doublejump_max_delay", "flip_torque_time", "flip_initial_vel_scale",
                   "flip_forward_impulse_max_speed_scale", "flip_side_impulse_max_speed_scale",
                   "flip_backward_impulse_max_speed_scale", "flip_backward_impulse_scale_x", "boost_spawn_amount",
                   "unlimited_flips", "unlimited_double_jumps", "jump_accel", "boost_max", "boost_used_per_second",
                   "boost_min_time", "boost_accel_ground", "recharge_boost_per_second", "recharge_boost_delay",
                   "recharge_boost_enabled", "car_air_throttle_accel", "flip_pitchlock_extra_time",
                   "car_air_control_pitch", "car_air_control_yaw", "car_air_control_roll", "car_air_damp_pitch",
                   "car_air_damp_yaw", "car_air_damp_roll", "flip_torque_x", "flip_torque_y")


def force_params(**override) -> np.ndarray:
    """A force-phase parameter list: the reference's defaults with the named fields replaced."""
    d = dict(zip(CAR_PARAM_NAMES, (float(v) for v in car_params())))
    d.update(override)
    return np.array([d[k] for k in FORCE_PARAM_NAMES], np.float32)


def car_force_phase(phase: int, cars, controls, queues, params, configs=None, flags: int = 3):
    """UpdateJumpState (phase 0), ApplyBoost (phase 1) or ApplyAirControl (phase 2, with one CarConfig per car and
    flags bit 0 update / bit 1 throttle) on independent cars, queueing on the given force queues (None: the form
    that changes the velocity).  Returns (cars, queues) after it."""
    c = np.ascontiguousarray(cars, dtype=np.uint8).copy()
    q = None if queues is None else np.ascontiguousarray(queues, dtype=np.uint8).copy()
    ctl = np.ascontiguousarray(controls, dtype=np.uint8)
    cfg = None if configs is None else np.ascontiguousarray(configs, dtype=np.uint8)
    pr = np.ascontiguousarray(params, dtype=np.float32)
    if lib().rgref_car_force_phase(c.shape[0], _ptr(c), _ptr(ctl), None if cfg is None else _ptr(cfg),
                                   None if q is None else _ptr(q), phase, flags, _ptr(pr)) != 0:
        raise RuntimeError("rgref_car_force_phase failed")
    return c, q


def car_arena_run(n_env: int, cars_per_env: int, ticks: int, seed: int, p_jump: float = 0.25) -> dict:
    """Car arenas as the reference steps them: n_env arenas of cars_per_env cars (hitbox presets cycling through the
    six bodies, alternating teams) from a kickoff, with random controls held for 1-30 ticks (axes uniform in
    [-1, 1], jump / boost / handbrake with probability p_jump / 0.3 / 0.05).  Returns the records before each tick:
    cars [ticks, cap, 880], wheels [ticks, cap, 292], controls [ticks, cap, 24], envs [ticks, n_env, 44], and the
    batch's configs [cap, 120] and teams [cap]."""
    rng = np.random.default_rng(seed)
    s = RefSim(n_env, cars_per_env)
    cap = n_env * cars_per_env
    teams = (np.arange(cap) % 2).astype(np.uint8)
    s.set_env_car_count(np.full(n_env, cars_per_env, np.int32))
    s.set_teams(teams)
    s.set_car_presets(np.arange(cap) % 6)
    s.reset_to_kickoff(seed)
    configs = s.download_car_configs(cap)
    ctl = np.zeros(cap, CONTROLS_DTYPE)
    left = np.zeros(cap, np.int64)
    out = {"cars": [], "wheels": [], "controls": [], "envs": []}
    for _ in range(ticks):
        for i in range(cap):
            if left[i] <= 0:
                left[i] = rng.integers(1, 31)
                for k in ("throttle", "steer", "pitch", "yaw", "roll"):
                    ctl[i][k] = rng.uniform(-1.0, 1.0)
                ctl[i]["jump"] = rng.random() < p_jump
                ctl[i]["boost"] = rng.random() < 0.3
                ctl[i]["handbrake"] = rng.random() < 0.05
            left[i] -= 1
        cb = ctl.view(np.uint8).reshape(cap, 24).copy()
        out["cars"].append(s.download(CAR, cap))
        out["wheels"].append(s.download_wheels(cap))
        out["controls"].append(cb)
        out["envs"].append(s.download(ENV, n_env))
        s.step_with_controls(cb, 1, cap)
    s.close()
    res = {k: np.stack(v) for k, v in out.items()}
    res["configs"] = configs
    res["teams"] = teams
    return res
