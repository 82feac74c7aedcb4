"""Car-arena cases shared by the twin and GPU parity tests: the reference batch at a kickoff, and byte comparison of
its records with this build's, arena by arena.

The reference (oracle/_ref) is the master: it is reset to a kickoff, its records are copied into the batch under test
(PreserveInternal uploads: records as they are), and both then step the same seeded held controls.  An arena is
compared while this build's unsupported word for it is 0 (every tick stepped in full); once a bit is set, the arena
met physics this build does not compute and it is no longer compared with the reference.
"""
import numpy as np

WHEEL_BYTES = 292
PADS_PER_ARENA = 34


def reference_at_kickoff(R, n_env, cars, seed, presets=None):
    """The reference batch after reset_to_kickoff(seed): n_env arenas of `cars` cars, Blue then Orange halves,
    hitbox presets (Octane when None)."""
    ref = R.RefSim(n_env, cars, respawn_seed=seed)
    cap = n_env * cars
    counts = np.full(n_env, cars, np.int32)
    teams = (np.arange(cap) % cars >= cars // 2).astype(np.uint8)
    ref.set_env_car_count(counts)
    ref.set_teams(teams)
    ref.set_car_presets(np.zeros(cap, np.uint8) if presets is None else np.asarray(presets, np.uint8))
    ref.reset_to_kickoff(seed)
    return ref, counts, teams, ref.download_car_configs(cap)


def reference_records(R, ref, n_env, cars):
    cap = n_env * cars
    return {"cars": ref.download(R.CAR, cap), "wheels": ref.download_wheels(cap), "balls": ref.download(R.BALL, n_env),
            "envs": ref.download(R.ENV, n_env), "pads": ref.download(R.PAD, n_env * PADS_PER_ARENA)}


def arena_mismatches(a, b, n_env, cars):
    """Arenas whose records differ between record sets a and b."""
    bad = np.zeros(n_env, bool)
    bad |= (a["cars"].reshape(n_env, -1) != b["cars"].reshape(n_env, -1)).any(axis=1)
    bad |= (a["wheels"].reshape(n_env, -1) != b["wheels"].reshape(n_env, -1)).any(axis=1)
    bad |= (a["balls"] != b["balls"]).any(axis=1)
    bad |= (a["envs"] != b["envs"]).any(axis=1)
    bad |= (a["pads"].reshape(n_env, -1) != b["pads"].reshape(n_env, -1)).any(axis=1)
    return np.nonzero(bad)[0]


def batch_records(sim):
    """The records of a rocketgpu_amd.BatchSim in the layout of reference_records, and its unsupported words."""
    get = sim.download_state_array
    return ({"cars": get("car_states"), "wheels": get("wheel_states"), "balls": get("ball_states"),
             "envs": get("env_states"), "pads": get("boost_pad_states")},
            get("unsupported").view(np.uint32).ravel())


_CONTROLS = np.dtype([("throttle", "<f4"), ("steer", "<f4"), ("pitch", "<f4"), ("yaw", "<f4"), ("roll", "<f4"),
                      ("jump", "u1"), ("boost", "u1"), ("handbrake", "u1"), ("pad", "u1")])


def teammate_chase_controls(car_records, n_env, cars):
    """Controls [n_env * cars, 24] u8 that steer every car at full throttle toward the next car of its own team
    (slots 0..cars/2-1 Blue, the rest Orange), so that teammates drive into each other: car-car contacts and bumps
    without the ball (which sits between the teams) and below supersonic speed (no boost)."""
    from rocketgpu_amd import views

    c = np.ascontiguousarray(car_records).view(views.state_dtype("car_states")).reshape(n_env, cars)
    half = cars // 2
    target = np.array([(k + 1) % half + (k // half) * half for k in range(cars)])
    d = c["pos"][:, target] - c["pos"]
    fwd = c["basis_forward"]
    angle = np.arctan2(fwd[..., 0] * d[..., 1] - fwd[..., 1] * d[..., 0], (fwd[..., :2] * d[..., :2]).sum(-1))
    out = np.zeros(n_env * cars, _CONTROLS)
    out["throttle"] = 1.0
    out["steer"] = np.clip(3.0 * angle, -1.0, 1.0).ravel()
    return np.ascontiguousarray(out.view(np.uint8).reshape(n_env * cars, _CONTROLS.itemsize))
