"""rocketgpu_amd: the batched arena step of RocketGPU on MI355X (gfx950).

The native library (rocketgpu_amd/_build/librocketgpu_amd.so) holds the HIP kernels, BatchSim and the C-ABI of
include/rocketgpu_amd/rg_capi.h.  This module is the Python face of that C-ABI, mirroring the reference's Python
surface (python/rocketgpu_bindings.cpp: BatchSim, state_layouts, download_state_array, device arrays):

    sim = BatchSim(env_count=4096, max_cars_per_env=6)        # procedural arena PA(1, 256)
    sim.upload("ball_states", balls)                            # numpy uint8 [n, 288] or a structured array
    sim.step(8)
    balls = sim.download_state_array("ball_states")

There is no CPU fallback: importing works anywhere, but constructing a BatchSim loads the HIP library and needs a
GPU; a missing library raises.
"""
from __future__ import annotations

import ctypes
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# The built library; RG_LIBRARY_PATH points the package at another build of it.
LIB_PATH = os.environ.get("RG_LIBRARY_PATH") or os.path.join(HERE, "_build", "librocketgpu_amd.so")

RG_ARRAY = {"car_states": 0, "ball_states": 1, "env_states": 2, "boost_pad_states": 3, "controls": 4,
            "env_car_count": 5, "wheel_states": 6, "unsupported": 7}
ITEMSIZE = {"car_states": 880, "ball_states": 288, "env_states": 44, "boost_pad_states": 16, "controls": 24,
            "env_car_count": 4, "car_configs": 120, "teams": 1, "env_commands": 16, "wheel_states": 292,
            "unsupported": 4}

# EnvCommand (include/rocketgpu/rocketgpu.h:683-707 of the reference), 16 bytes.
ENV_COMMAND_DTYPE = np.dtype([("reset_to_kickoff", "u1"), ("reset_scoreboard", "u1"), ("reset_tick_count", "u1"),
                              ("kickoff_seed_mode", "u1"), ("new_round_bodies", "u1"), ("keep_boost_pads", "u1"),
                              ("kickoff_countdown_ticks", "<u2"), ("kickoff_seed", "<u8")])
# Physics event stream records (PhysicsEvent / PhysicsEventWindow, include/rocketgpu/rocketgpu.h:45-72 of the
# reference), 104 and 16 bytes; padding bytes are unspecified.
PHYSICS_EVENT_TYPES = ("BallTouchStart", "BallWorldContact", "CarWorldContact", "BoostPickup", "CarBump", "Goal")
PHYSICS_EVENT_DTYPE = np.dtype({
    "names": ["tick", "type", "car_id", "other_car_id", "team", "pad_index", "car_on_ground", "is_big_pad", "is_demo",
              "position", "normal", "car_velocity", "ball_position", "ball_velocity", "boost_before", "boost_after"],
    "formats": ["<u8", "u1", "<u4", "<u4", "<i4", "<u4", "u1", "u1", "u1", ("<f4", 3), ("<f4", 3), ("<f4", 3),
                ("<f4", 3), ("<f4", 3), "<f4", "<f4"],
    "offsets": [0, 8, 12, 16, 20, 24, 28, 29, 30, 32, 44, 56, 68, 80, 92, 96],
    "itemsize": 104})
PHYSICS_EVENT_WINDOW_DTYPE = np.dtype([("tick", "<u8"), ("count", "<u4"), ("overflow", "<u4")])


def physics_event_layout() -> dict:
    """rocketgpu_py.physics_event_layout() (python/rocketgpu_bindings.cpp:77-86 of the reference)."""
    return {"record_size": PHYSICS_EVENT_DTYPE.itemsize, "window_size": PHYSICS_EVENT_WINDOW_DTYPE.itemsize,
            "type_offset": PHYSICS_EVENT_DTYPE.fields["type"][1],
            "velocity_offset": PHYSICS_EVENT_DTYPE.fields["ball_velocity"][1],
            "count_offset": PHYSICS_EVENT_WINDOW_DTYPE.fields["count"][1],
            "overflow_offset": PHYSICS_EVENT_WINDOW_DTYPE.fields["overflow"][1],
            "goal_type": PHYSICS_EVENT_TYPES.index("Goal")}


KICKOFF_COUNTDOWN_GAME_TIMERS = 0xFFFF  # kKickoffCountdownGameTimers
KICKOFF_COUNTDOWN_TICKS = 480           # kKickoffCountdownTicks

_lib = None


class RocketGpuError(RuntimeError):
    pass


class _SimConfig(ctypes.Structure):
    _fields_ = [("env_count", ctypes.c_int32), ("max_cars_per_env", ctypes.c_int32), ("tick_dt", ctypes.c_float),
                ("enable_boost_pads", ctypes.c_uint8), ("enable_car_car_collisions", ctypes.c_uint8),
                ("enable_demos", ctypes.c_uint8), ("enable_scoring", ctypes.c_uint8),
                ("respawn_rng_seed", ctypes.c_uint64)]


class _DeviceView(ctypes.Structure):
    _fields_ = [(name, ctypes.c_void_p) for name in
                ("car_states", "controls", "car_configs", "teams", "ball_states", "env_states", "boost_pad_states",
                 "env_car_count")]


class ServeConfig(ctypes.Structure):
    """rg_serve_config: simulated clients, Poisson arrivals, horizons, launch budget (include/rocketgpu_amd/rg_capi.h)."""
    _fields_ = [("clients", ctypes.c_int32), ("min_arenas", ctypes.c_int32), ("max_arenas", ctypes.c_int32),
                ("horizon_count", ctypes.c_int32), ("horizons", ctypes.c_int32 * 4),
                ("max_arenas_per_launch", ctypes.c_int32), ("max_in_flight", ctypes.c_int32),
                ("request_rate_hz", ctypes.c_double), ("duration_s", ctypes.c_double), ("seed", ctypes.c_uint64),
                ("launch_budget_bytes", ctypes.c_int64), ("reset_every", ctypes.c_int32),
                ("fresh_controls", ctypes.c_int32), ("p_boost", ctypes.c_float), ("reserved", ctypes.c_int32)]


class ServeReport(ctypes.Structure):
    _fields_ = [("requests", ctypes.c_int64), ("arena_ticks", ctypes.c_int64), ("launches", ctypes.c_int64),
                ("arenas_resident", ctypes.c_int64), ("resident_bytes", ctypes.c_int64), ("seconds", ctypes.c_double),
                ("arena_ticks_per_s", ctypes.c_double), ("p50_ms", ctypes.c_double), ("p99_ms", ctypes.c_double),
                ("mean_ms", ctypes.c_double), ("max_ms", ctypes.c_double)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


def library_path() -> str:
    return LIB_PATH


def lib():
    """Load the HIP library (raises when it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RocketGpuError(f"librocketgpu_amd.so is not built ({LIB_PATH}); run __graft_entry__.build()")
        # One HIP runtime per process: torch ships its own libamdhip64.so.7.  Loaded first, it is the one this
        # library binds to (same soname).  Loaded after this library, the process would hold two HIP/HSA runtimes
        # and the second to open the GPU sees no device.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        L.rg_last_error.restype = ctypes.c_char_p
        L.rg_version.restype = ctypes.c_char_p
        L.rg_batch_create_procedural.argtypes = [ctypes.POINTER(_SimConfig), ctypes.c_double, ctypes.c_double,
                                                 ctypes.POINTER(vp)]
        L.rg_batch_destroy.argtypes = [vp]
        L.rg_batch_destroy.restype = None
        L.rg_batch_env_count.argtypes = [vp]
        L.rg_batch_env_count.restype = ctypes.c_int64
        L.rg_batch_car_capacity.argtypes = [vp]
        L.rg_batch_car_capacity.restype = ctypes.c_int64
        L.rg_batch_set_env_car_count.argtypes = [vp, vp, ctypes.c_size_t, vp]
        L.rg_batch_upload.argtypes = [vp, ctypes.c_int, vp, ctypes.c_size_t, vp]
        L.rg_batch_upload_mode.argtypes = [vp, ctypes.c_int, vp, ctypes.c_size_t, ctypes.c_int32, vp]
        L.rg_batch_download.argtypes = [vp, ctypes.c_int, vp, ctypes.c_size_t, vp]
        L.rg_batch_step.argtypes = [vp, ctypes.c_int32, vp]
        L.rg_batch_step_split_no_alloc.argtypes = [vp, ctypes.c_int32, vp]
        L.rg_batch_step_with_device_controls.argtypes = [vp, vp, ctypes.c_int32, ctypes.c_int32, vp]
        L.rg_batch_synchronize.argtypes = [vp, vp]
        L.rg_batch_device_view.argtypes = [vp, ctypes.POINTER(_DeviceView)]
        L.rg_batch_env_digests.argtypes = [vp, vp, vp]
        L.rg_workload_ball_states.argtypes = [ctypes.c_uint64, ctypes.c_int64, ctypes.c_int64, ctypes.c_double, vp]
        L.rg_workload_controls.argtypes = [ctypes.c_uint64, ctypes.c_int64, ctypes.c_int64, ctypes.c_int32,
                                           ctypes.c_int32, ctypes.c_float, vp]
        L.rg_batch_step_arenas.argtypes = [vp, vp, vp, ctypes.c_int32, ctypes.c_int32, vp]
        L.rg_batch_step_graph.argtypes = [vp, ctypes.c_int32, ctypes.c_int32, vp]
        L.rg_batch_ball_path.argtypes = [vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_float, vp, vp, vp]
        L.rg_batch_set_launch_shape.argtypes = [vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]
        L.rg_batch_set_split.argtypes = [vp, ctypes.c_int32, ctypes.c_int32]
        L.rg_serve_run.argtypes = [vp, ctypes.POINTER(ServeConfig), ctypes.POINTER(ServeReport), vp]
        L.rg_serve_client_arenas.argtypes = [ctypes.POINTER(ServeConfig), ctypes.c_int32, vp, vp]
        L.rg_serve_client_requests.argtypes = [ctypes.POINTER(ServeConfig), ctypes.c_int32, vp, ctypes.c_int64]
        L.rg_serve_client_requests.restype = ctypes.c_int64
        L.rg_workload_request_controls.argtypes = [ctypes.c_uint64, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64,
                                                   ctypes.c_int32, ctypes.c_float, vp]
        L.rg_batch_arena_bytes.argtypes = [vp, ctypes.c_int32]
        L.rg_batch_arena_bytes.restype = ctypes.c_int64
        L.rg_fold_digests.argtypes = [vp, ctypes.c_int64]
        L.rg_car_preset_config.argtypes = [ctypes.c_int32, vp]
        L.rg_batch_download_async.argtypes = [vp, ctypes.c_int, vp, ctypes.c_size_t, vp]
        L.rg_batch_copy_state_from.argtypes = [vp, vp, vp]
        L.rg_mutator_config_defaults.argtypes = [vp]
        L.rg_mutator_config_defaults.restype = None
        L.rg_batch_set_mutator_config.argtypes = [vp, vp]
        L.rg_batch_get_mutator_config.argtypes = [vp, vp]
        L.rg_fold_digests.restype = ctypes.c_uint64
        L.rg_batch_boost_pad_capacity.argtypes = [vp]
        L.rg_batch_boost_pad_capacity.restype = ctypes.c_int64
        L.rg_batch_set_teams.argtypes = [vp, vp, ctypes.c_size_t, vp]
        L.rg_batch_set_car_configs.argtypes = [vp, vp, ctypes.c_size_t, vp]
        L.rg_batch_set_kickoff_settled.argtypes = [vp, ctypes.c_int32]
        L.rg_batch_set_physics_event_stream.argtypes = [vp, vp, vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                                        ctypes.c_int32]
        L.rg_batch_state_snapshot_size.argtypes = [vp, ctypes.POINTER(ctypes.c_size_t)]
        L.rg_batch_save_state.argtypes = [vp, vp, ctypes.c_size_t, vp]
        L.rg_batch_load_state.argtypes = [vp, vp, ctypes.c_size_t, vp]
        L.rg_batch_reset_to_kickoff.argtypes = [vp, ctypes.c_uint64, vp]
        L.rg_batch_apply_env_commands.argtypes = [vp, vp, ctypes.c_size_t, vp]
        L.rg_batch_apply_env_commands_device.argtypes = [vp, vp, ctypes.c_size_t, vp]
        L.rg_batch_step_with_device_commands.argtypes = [vp, vp, ctypes.c_int32, vp, ctypes.c_int32, ctypes.c_int32,
                                                          vp]
        L.rg_state_layouts_json.argtypes = [ctypes.c_char_p, ctypes.c_size_t]
        L.rg_state_layouts_json.restype = ctypes.c_int64
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise RocketGpuError(lib().rg_last_error().decode())


def state_layouts() -> dict:
    """Field layouts of the state records (the reference's state_layouts())."""
    L = lib()
    n = L.rg_state_layouts_json(None, 0)
    buf = ctypes.create_string_buffer(int(n) + 1)
    L.rg_state_layouts_json(buf, n + 1)
    return json.loads(buf.value.decode())


def workload_ball_states(seed: int, first_arena: int, count: int, scale: float = 1.0) -> np.ndarray:
    """Seeded ball-launch BallState records ([count, 288] uint8) for the ball-arena workload."""
    out = np.zeros((count, ITEMSIZE["ball_states"]), dtype=np.uint8)
    _check(lib().rg_workload_ball_states(seed, first_arena, count, scale, out.ctypes.data))
    return out


def workload_controls(seed: int, first_arena: int, arena_count: int, cars_per_arena: int, ticks: int,
                      p_boost: float = 0.3) -> np.ndarray:
    """Seeded held controls ([ticks, arena_count * cars_per_arena, 24] uint8 CarControls records)."""
    out = np.zeros((ticks, arena_count * cars_per_arena, ITEMSIZE["controls"]), dtype=np.uint8)
    _check(lib().rg_workload_controls(seed, first_arena, arena_count, cars_per_arena, ticks, p_boost,
                                      out.ctypes.data))
    return out


CAR_PRESETS = ("octane", "dominus", "plank", "breakout", "hybrid", "merc", "psyclops", "batmobile")


def car_preset_config(preset) -> np.ndarray:
    """The CarConfig record (120 bytes) of a stock body, by index or name (OctaneConfig() ... BatmobileConfig())."""
    i = CAR_PRESETS.index(preset.lower()) if isinstance(preset, str) else int(preset)
    out = np.zeros(ITEMSIZE["car_configs"], dtype=np.uint8)
    _check(lib().rg_car_preset_config(i, out.ctypes.data))
    return out


def mutator_config(**fields) -> np.ndarray:
    """A MutatorConfig record (rocketgpu.h:708-840) with the reference's defaults and `fields` replaced: a one-element
    array of state_dtype("mutator_config")."""
    from .views import state_dtype

    rec = np.zeros(1, state_dtype("mutator_config"))
    lib().rg_mutator_config_defaults(rec.ctypes.data)
    for k, v in fields.items():
        if k not in rec.dtype.names:
            raise KeyError(f"MutatorConfig has no field {k}")
        rec[0][k] = v
    return rec


def serve_client_plan(cfg: ServeConfig, client: int):
    """A client's arena range and the ticks of its requests in the order the serving loop serves them."""
    first, count = ctypes.c_int32(), ctypes.c_int32()
    _check(lib().rg_serve_client_arenas(ctypes.byref(cfg), client, ctypes.byref(first), ctypes.byref(count)))
    n = lib().rg_serve_client_requests(ctypes.byref(cfg), client, None, 0)
    if n < 0:
        raise RocketGpuError(lib().rg_last_error().decode())
    ticks = np.zeros(max(n, 1), np.int32)
    lib().rg_serve_client_requests(ctypes.byref(cfg), client, ticks.ctypes.data, n)
    return int(first.value), int(count.value), ticks[:n]


def request_controls(seed: int, request: int, first_arena: int, arena_count: int, cars_per_arena: int,
                     p_boost: float = 0.3) -> np.ndarray:
    """The held controls of a serving request ([arena_count * cars_per_arena, 24] uint8 CarControls records)."""
    out = np.zeros((arena_count * cars_per_arena, ITEMSIZE["controls"]), np.uint8)
    _check(lib().rg_workload_request_controls(seed, request, first_arena, arena_count, cars_per_arena, p_boost,
                                              out.ctypes.data))
    return out


def _stream_ptr(stream):
    if stream is None:
        return None
    if isinstance(stream, int):
        return ctypes.c_void_p(stream)
    return ctypes.c_void_p(stream.cuda_stream)


class BatchSim:
    """BatchSim over the procedural arena PA(scale, fillet_radius) (include/rocketgpu/rocketgpu.h:988-1089)."""

    def __init__(self, env_count: int, max_cars_per_env: int = 6, scale: float = 1.0, fillet_radius: float = 256.0,
                 tick_dt: float = 1.0 / 120.0, enable_scoring: bool = True):
        L = lib()
        cfg = _SimConfig(env_count, max_cars_per_env, tick_dt, 1, 1, 1, 1 if enable_scoring else 0, 0)
        h = ctypes.c_void_p()
        _check(L.rg_batch_create_procedural(ctypes.byref(cfg), scale, fillet_radius, ctypes.byref(h)))
        self._h = h
        self.env_count = int(L.rg_batch_env_count(h))
        self.car_capacity = int(L.rg_batch_car_capacity(h))
        self.boost_pad_capacity = int(L.rg_batch_boost_pad_capacity(h))
        self.max_cars_per_env = max_cars_per_env

    def close(self):
        if getattr(self, "_h", None):
            lib().rg_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _count(self, name):
        return {"car_states": self.car_capacity, "controls": self.car_capacity, "car_configs": self.car_capacity,
                "teams": self.car_capacity, "wheel_states": self.car_capacity,
                "boost_pad_states": self.boost_pad_capacity}.get(name, self.env_count)

    def set_env_car_count(self, counts, stream=None):
        a = np.ascontiguousarray(counts, dtype=np.int32)
        _check(lib().rg_batch_set_env_car_count(self._h, a.ctypes.data, a.size, _stream_ptr(stream)))

    def upload(self, name: str, rows, stream=None, preserve_internal: bool = False):
        """upload_*_states; preserve_internal selects StateUploadMode::PreserveInternal (records as given)."""
        a = np.ascontiguousarray(rows)
        n = a.nbytes // ITEMSIZE[name]
        if n * ITEMSIZE[name] != a.nbytes:
            raise ValueError(f"{name}: byte size is not a multiple of {ITEMSIZE[name]}")
        _check(lib().rg_batch_upload_mode(self._h, RG_ARRAY[name], a.ctypes.data, n, 1 if preserve_internal else 0,
                                          _stream_ptr(stream)))

    def download_state_array(self, name: str, host_ptr: int = 0, count: int | None = None, stream=None):
        """download_state_array (python/rocketgpu_bindings.cpp:984-1000): `count` records (all by default) of
        `name` into host memory at host_ptr; without a pointer, returns them as a [count, itemsize] uint8 array."""
        n = self._count(name) if count is None else int(count)
        if host_ptr:
            _check(lib().rg_batch_download(self._h, RG_ARRAY[name], ctypes.c_void_p(host_ptr), n,
                                           _stream_ptr(stream)))
            return None
        out = np.zeros((n, ITEMSIZE[name]), dtype=np.uint8)
        _check(lib().rg_batch_download(self._h, RG_ARRAY[name], out.ctypes.data, n, _stream_ptr(stream)))
        return out

    def download_state_array_async(self, name: str, out: np.ndarray, stream=None):
        """download_*_async (rocketgpu.h:1050-1053): queues the copy of all records of `name` into `out` (a
        contiguous [count, itemsize] uint8 array, pinned for a true asynchronous copy) on `stream` without waiting;
        synchronize the stream before reading `out`."""
        n = self._count(name)
        if not out.flags["C_CONTIGUOUS"] or out.nbytes < n * ITEMSIZE[name]:
            raise ValueError("out must be a contiguous array of at least count * itemsize bytes")
        _check(lib().rg_batch_download_async(self._h, RG_ARRAY[name], out.ctypes.data, n, _stream_ptr(stream)))

    def copy_state_from(self, source: "BatchSim", stream=None):
        """copy_state_from (rocketgpu.h:1045): every state array of `source` (same shape, arena and physics)."""
        _check(lib().rg_batch_copy_state_from(self._h, source._h, _stream_ptr(stream)))

    def upload_state_array(self, name: str, host_ptr: int, count: int, stream=None):
        """upload_state_array (python/rocketgpu_bindings.cpp:955-983), public-state mode: `count` raw records."""
        _check(lib().rg_batch_upload(self._h, RG_ARRAY[name], ctypes.c_void_p(host_ptr), count, _stream_ptr(stream)))

    def set_teams(self, teams, stream=None):
        a = np.ascontiguousarray(teams, dtype=np.uint8)
        _check(lib().rg_batch_set_teams(self._h, a.ctypes.data, a.size, _stream_ptr(stream)))

    def set_car_configs(self, configs, stream=None):
        a = np.ascontiguousarray(configs).view(np.uint8).reshape(-1, ITEMSIZE["car_configs"])
        _check(lib().rg_batch_set_car_configs(self._h, a.ctypes.data, a.shape[0], _stream_ptr(stream)))

    def set_kickoff_settled(self, on: bool = True):
        """MutatorConfig::kickoff_settled: a kickoff reset without a countdown puts the ball where the countdown
        leaves it (True, the default) or asleep at its spawn (False)."""
        _check(lib().rg_batch_set_kickoff_settled(self._h, 1 if on else 0))

    def set_physics_event_stream(self, records: int, windows: int, capacity: int, goal_only: bool = False):
        """BatchSim.set_physics_event_stream(records, windows, capacity, goal_only) (python/rocketgpu_bindings.cpp:
        832-845 of the reference): device addresses of env_count * capacity PHYSICS_EVENT_DTYPE records and env_count
        PHYSICS_EVENT_WINDOW_DTYPE windows.  Pass zeros to disable; synchronize before freeing the buffers."""
        envs = cars = 0
        if records or windows or capacity:
            envs = self.env_count
            cars = self.car_capacity // max(self.env_count, 1)
        _check(lib().rg_batch_set_physics_event_stream(self._h, ctypes.c_void_p(records or None),
                                                      ctypes.c_void_p(windows or None), capacity, envs, cars,
                                                      1 if goal_only else 0))

    def save_state(self, stream=None) -> bytes:
        """The batch's full state as one snapshot blob (rg_batch_save_state)."""
        n = ctypes.c_size_t()
        _check(lib().rg_batch_state_snapshot_size(self._h, ctypes.byref(n)))
        buf = ctypes.create_string_buffer(n.value)
        _check(lib().rg_batch_save_state(self._h, buf, n.value, _stream_ptr(stream)))
        return buf.raw

    def load_state(self, blob: bytes, stream=None):
        """Restore a snapshot taken from a batch of the same shape and arena (rg_batch_load_state)."""
        buf = ctypes.create_string_buffer(bytes(blob), len(blob))
        _check(lib().rg_batch_load_state(self._h, buf, len(blob), _stream_ptr(stream)))

    def get_full_state(self, arrays=None) -> dict:
        """Arena::GetFullState(arrays) (src/rocketsim_compat_dev.cpp:369-388 of the reference) over the public
        arrays: {"array.field": bytes}, each field's bytes for every element in order, fields as state_layouts()
        lists them."""
        layouts = state_layouts()
        names = [a for a in ("car_states", "ball_states", "env_states", "boost_pad_states", "controls")]
        for a in arrays or ():
            if a not in names:
                raise ValueError("GetFullState: unknown array " + a)
        out = {}
        for a in names:
            if arrays and a not in arrays:
                continue
            raw = self.download_state_array(a).view(np.uint8).reshape(-1, layouts[a]["itemsize"])
            for name, fmt, off, shape in layouts[a]["fields"]:
                size = np.dtype(fmt).itemsize * int(np.prod(shape, dtype=np.int64))
                out[a + "." + name] = np.ascontiguousarray(raw[:, off:off + size]).tobytes()
        return out

    def reset_to_kickoff(self, seed: int = 0, stream=None):
        _check(lib().rg_batch_reset_to_kickoff(self._h, seed, _stream_ptr(stream)))

    def apply_env_commands(self, commands, stream=None):
        """Command i (ENV_COMMAND_DTYPE records, or [n, 16] uint8) to env i."""
        a = np.ascontiguousarray(commands).view(np.uint8).reshape(-1, ITEMSIZE["env_commands"])
        _check(lib().rg_batch_apply_env_commands(self._h, a.ctypes.data, a.shape[0], _stream_ptr(stream)))

    def apply_env_commands_device(self, device_commands_ptr: int, count: int, stream=None):
        _check(lib().rg_batch_apply_env_commands_device(self._h, ctypes.c_void_p(device_commands_ptr), count,
                                                        _stream_ptr(stream)))

    def step_with_device_commands(self, device_controls_ptr: int, control_stride: int, device_commands_ptr: int,
                                  command_stride: int, ticks: int, stream=None):
        """Per tick t, env e's command device_commands[t * command_stride + e] (EnvCommand records), then the tick;
        the last tick's controls become the batch's controls."""
        _check(lib().rg_batch_step_with_device_commands(self._h, ctypes.c_void_p(device_controls_ptr), control_stride,
                                                        ctypes.c_void_p(device_commands_ptr), command_stride, ticks,
                                                        _stream_ptr(stream)))

    def step(self, ticks: int = 1, stream=None):
        _check(lib().rg_batch_step(self._h, ticks, _stream_ptr(stream)))

    def step_with_device_controls(self, device_controls_ptr: int, ticks: int, control_stride: int, stream=None):
        _check(lib().rg_batch_step_with_device_controls(self._h, ctypes.c_void_p(device_controls_ptr), ticks,
                                                        control_stride, _stream_ptr(stream)))

    def step_arenas(self, device_ids_ptr: int, count: int, ticks: int = 0, device_ticks_ptr: int = 0, stream=None):
        """Step `count` distinct arenas (device int32 ids), each for device_ticks[i] ticks or `ticks` ticks."""
        _check(lib().rg_batch_step_arenas(self._h, ctypes.c_void_p(device_ids_ptr),
                                          ctypes.c_void_p(device_ticks_ptr or None), count, ticks,
                                          _stream_ptr(stream)))

    def set_mutator_config(self, config=None, **fields):
        """set_mutator_config (rocketgpu.h:1003): a record from mutator_config(), or the defaults with `fields`
        replaced.  Fields this build keeps fixed must keep their defaults (RocketGpuError names the field)."""
        rec = mutator_config(**fields) if config is None else np.ascontiguousarray(config)
        _check(lib().rg_batch_set_mutator_config(self._h, rec.ctypes.data))

    def get_mutator_config(self) -> np.ndarray:
        from .views import state_dtype

        rec = np.zeros(1, state_dtype("mutator_config"))
        _check(lib().rg_batch_get_mutator_config(self._h, rec.ctypes.data))
        return rec

    def step_split_no_alloc(self, ticks: int = 1, stream=None):
        """step(ticks) through the split schedule with its arena lists in the library's fixed device array: allocates
        no device memory (ball-only batches of at most 262,144 arenas; overlapping calls on one GPU race)."""
        _check(lib().rg_batch_step_split_no_alloc(self._h, ticks, _stream_ptr(stream)))

    def step_graph(self, ticks_per_launch: int, launches: int = 1, stream=None):
        """`launches` x step(ticks_per_launch), captured once into a hipGraph and replayed."""
        _check(lib().rg_batch_step_graph(self._h, ticks_per_launch, launches, _stream_ptr(stream)))

    def ball_path(self, samples: int, ticks: int, goal_y: float, out_ptr: int = 0, result_ptr: int = 0, stream=None):
        """BatchSim::ball_path: every arena's ball stepped in place for samples x ticks ticks.  With device pointers,
        writes out [env, samples, 6] f32 and result [env, 2] i32 there; without, returns them as torch tensors."""
        ret = None
        if not out_ptr or not result_ptr:
            import torch

            out = torch.zeros((self.env_count, samples, 6), dtype=torch.float32, device="cuda")
            res = torch.zeros((self.env_count, 2), dtype=torch.int32, device="cuda")
            out_ptr, result_ptr, ret = out.data_ptr(), res.data_ptr(), (out, res)
        _check(lib().rg_batch_ball_path(self._h, samples, ticks, goal_y, ctypes.c_void_p(out_ptr),
                                        ctypes.c_void_p(result_ptr), _stream_ptr(stream)))
        return ret

    def set_split(self, rounds: int = 26, cap_divisor: int = 16):
        """Launch stages 6 and 7: the heavy-arena threshold (contact-walk windows) and list capacity divisor."""
        _check(lib().rg_batch_set_split(self._h, rounds, cap_divisor))

    def set_launch_shape(self, block: int = 64, stage: int = 0, arenas_per_wave: int = 64):
        """Lanes per workgroup, state staging (0 in place, 1 lane-private copy), arenas per 64-lane wavefront."""
        _check(lib().rg_batch_set_launch_shape(self._h, block, stage, arenas_per_wave))

    @staticmethod
    def serve_config(clients=256, min_arenas=16, max_arenas=256, horizons=(1, 4, 8), max_arenas_per_launch=65536,
                     max_in_flight=2, request_rate_hz=2000.0, duration_s=2.0, seed=0, launch_budget_bytes=0,
                     reset_every=0, fresh_controls=False, p_boost=0.3) -> ServeConfig:
        h = (ctypes.c_int32 * 4)(*(list(horizons) + [0] * (4 - len(horizons))))
        return ServeConfig(clients, min_arenas, max_arenas, len(horizons), h, max_arenas_per_launch, max_in_flight,
                           request_rate_hz, duration_s, seed, launch_budget_bytes, reset_every,
                           1 if fresh_controls else 0, p_boost, 0)

    def serve(self, clients=256, min_arenas=16, max_arenas=256, horizons=(1, 4, 8), max_arenas_per_launch=65536,
              max_in_flight=2, request_rate_hz=2000.0, duration_s=2.0, seed=0, ticks_per_arena=False,
              launch_budget_bytes=0, reset_every=0, fresh_controls=False, p_boost=0.3):
        """Run the serving loop (continuous batching over this batch's arenas); returns the report as a dict and,
        with ticks_per_arena, the ticks each arena was stepped.  launch_budget_bytes admits requests by the HBM
        bytes of the arena records a launch touches; reset_every resets a client's arenas before every n-th of its
        requests; fresh_controls gives each request its own held controls."""
        cfg = self.serve_config(clients, min_arenas, max_arenas, horizons, max_arenas_per_launch, max_in_flight,
                                request_rate_hz, duration_s, seed, launch_budget_bytes, reset_every, fresh_controls,
                                p_boost)
        rep = ServeReport()
        per_arena = np.zeros(self.env_count, np.int32) if ticks_per_arena else None
        _check(lib().rg_serve_run(self._h, ctypes.byref(cfg), ctypes.byref(rep),
                                  per_arena.ctypes.data if per_arena is not None else None))
        return (rep.as_dict(), per_arena) if ticks_per_arena else rep.as_dict()

    def synchronize(self, stream=None):
        _check(lib().rg_batch_synchronize(self._h, _stream_ptr(stream)))

    def device_view(self) -> dict:
        v = _DeviceView()
        _check(lib().rg_batch_device_view(self._h, ctypes.byref(v)))
        return {name: getattr(v, name) for name, _ in _DeviceView._fields_}

    def env_digests_into(self, device_ptr: int, stream=None):
        """Per-env body digests (uint64) into a device buffer of env_count entries."""
        _check(lib().rg_batch_env_digests(self._h, ctypes.c_void_p(device_ptr), _stream_ptr(stream)))

    def device_arrays(self) -> dict:
        """device_arrays() (python/rocketgpu_bindings.cpp:846-898): per state array, a dict with a zero-copy
        `__cuda_array_interface__` (v3, one opaque record per element), itemsize, count, per_env and what one
        element is.  Records are the byte-compatible ones of state_layouts()."""
        view = self.device_view()
        cars = self.max_cars_per_env
        pads = self.boost_pad_capacity // self.env_count if self.env_count else 0
        spec = [("car_states", cars, "car slot"), ("controls", cars, "car slot"), ("car_configs", cars, "car slot"),
                ("teams", cars, "car slot"), ("ball_states", 1, "env"), ("env_states", 1, "env"),
                ("boost_pad_states", pads, "boost pad"), ("env_car_count", 1, "env")]
        out = {}
        for name, per_env, element in spec:
            count = self._count(name)
            out[name] = {
                "__cuda_array_interface__": {"data": (int(view[name] or 0), False), "shape": (count,),
                                             "typestr": f"|V{ITEMSIZE[name]}", "strides": None, "version": 3},
                "itemsize": ITEMSIZE[name], "count": count, "per_env": per_env, "element": element}
        return out

    def torch_arrays(self) -> dict:
        """Zero-copy torch views of the state arrays as raw bytes ([count, itemsize] uint8 tensors on cuda:0)."""
        import torch

        return {name: _torch_from_ptr(e["__cuda_array_interface__"]["data"][0], (e["count"], e["itemsize"]),
                                      torch.uint8)
                for name, e in self.device_arrays().items() if e["count"] and e["__cuda_array_interface__"]["data"][0]}


def _torch_from_ptr(ptr: int, shape, dtype):
    import torch

    class _Holder:
        def __init__(self):
            self.__cuda_array_interface__ = {"shape": tuple(shape), "typestr": "|u1", "data": (ptr, False),
                                             "version": 3, "strides": None}

    return torch.as_tensor(_Holder(), device="cuda")


from .views import (DeviceArray, DeviceFieldArray, device_arrays, field_arrays, host_struct_array,  # noqa: E402
                    state_dtype, state_layout)
