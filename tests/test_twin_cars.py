"""The car-arena tick of the host twin against the reference CPU build, tick by tick.

Arenas of two Octanes (the 1v1 of the headline workload) and of mixed bodies start from the reference's kickoff and
step seeded held controls (workload.h ControlStream).  Every record of every arena must match the reference byte for
byte on every tick while the arena's unsupported word is 0, and the word must rise no later than the first
difference.  The test also requires most arena-ticks to be compared, so it cannot pass by flagging everything.
"""
import numpy as np
import pytest

from car_cases import arena_mismatches, reference_at_kickoff, reference_records, teammate_chase_controls
from conftest import oracle_or_skip, twin_cars
from rocketgpu_amd import views


def _twin_records(T, h, n_env, cars):
    cap = n_env * cars
    def get(which, shape, dtype=np.uint8):
        a = np.zeros(shape, dtype)
        assert T.rgt_cars_download(h, which, a.ctypes.data, shape[0]) == 0
        return a
    return {"cars": get(0, (cap, 880)), "wheels": get(8, (cap, 292)), "balls": get(1, (n_env, 288)),
            "envs": get(2, (n_env, 44)), "pads": get(3, (n_env * 34, 16))}, get(9, (n_env,), np.uint32)


def _run(n_env, cars, ticks, seed, presets=None, p_boost=0.3):
    R = oracle_or_skip()
    T = twin_cars()
    import rocketgpu_amd as rga

    ref, counts, teams, cfgs = reference_at_kickoff(R, n_env, cars, seed, presets)
    h = T.rgt_cars_create(n_env, cars, 1.0, 256.0, seed)
    assert h
    rec = reference_records(R, ref, n_env, cars)
    for which, arr in ((6, counts), (7, teams), (5, cfgs), (0, rec["cars"]), (8, rec["wheels"]), (1, rec["balls"]),
                       (2, rec["envs"]), (3, rec["pads"])):
        a = np.ascontiguousarray(arr)
        assert T.rgt_cars_upload(h, which, a.ctypes.data, a.shape[0]) == 0
    controls = rga.workload_controls(seed, 0, n_env, cars, ticks, p_boost)
    compared = 0
    body_contacts = 0
    has_contact = views.state_dtype("car_states")
    flagged_at = np.full(n_env, -1)
    for t in range(ticks):
        ctl = np.ascontiguousarray(controls[t])
        assert T.rgt_cars_upload(h, 4, ctl.ctypes.data, ctl.shape[0]) == 0
        ref.step_with_controls(ctl, 1, ctl.shape[0])
        assert T.rgt_cars_step(h, 1) == 0
        mine, uns = _twin_records(T, h, n_env, cars)
        flagged_at[(uns != 0) & (flagged_at < 0)] = t
        bad = arena_mismatches(mine, reference_records(R, ref, n_env, cars), n_env, cars)
        live = uns == 0
        assert not live[bad].any(), f"tick {t}: arenas {bad[live[bad]]} differ from the reference while unflagged"
        compared += int(live.sum())
        wc = mine["cars"].view(has_contact)["world_contact.has_contact"].reshape(n_env, cars)
        body_contacts += int(np.count_nonzero(wc[live]))
    T.rgt_cars_destroy(h)
    ref.close()
    return compared, flagged_at, body_contacts


def test_twin_cars_1v1_octane_match_reference():
    n_env, ticks = 12, 400
    compared, flagged_at, body_contacts = _run(n_env, 2, ticks, seed=7)
    assert compared >= 0.95 * n_env * ticks, f"only {compared} of {n_env * ticks} arena-ticks compared"
    # The compared ticks include hitbox-on-mesh contacts (landings on the roof or side, wall hits), not only wheels.
    assert body_contacts > 20, f"only {body_contacts} car-ticks with a hitbox-mesh contact"


@pytest.mark.parametrize("cars", [4, 6])
def test_twin_cars_mixed_bodies_match_reference(cars):
    n_env, ticks = 6, 200
    compared, _, _ = _run(n_env, cars, ticks, seed=11 + cars, presets=np.arange(n_env * cars) % 6)
    assert compared >= 0.9 * n_env * ticks, f"only {compared} of {n_env * ticks} arena-ticks compared"


def test_twin_cars_teammates_collide_match_reference():
    """Teammates steered into each other: hitbox-hitbox contacts, the pair manifolds, bumps, wheels resting on other
    cars and the joint solve of the touching cars, compared with the reference on every tick while unflagged."""
    R = oracle_or_skip()
    T = twin_cars()
    n_env, cars, ticks, seed = 16, 6, 1000, 9
    presets = np.arange(n_env * cars) % 6
    ref, counts, teams, cfgs = reference_at_kickoff(R, n_env, cars, seed, presets)
    h = T.rgt_cars_create(n_env, cars, 1.0, 256.0, seed)
    theirs = reference_records(R, ref, n_env, cars)
    for which, arr in ((6, counts), (7, teams), (5, cfgs), (0, theirs["cars"]), (8, theirs["wheels"]),
                       (1, theirs["balls"]), (2, theirs["envs"]), (3, theirs["pads"])):
        a = np.ascontiguousarray(arr)
        assert T.rgt_cars_upload(h, which, a.ctypes.data, a.shape[0]) == 0
    ever_flagged = np.zeros(n_env, bool)
    compared = 0
    for t in range(ticks):
        ctl = teammate_chase_controls(theirs["cars"], n_env, cars)
        assert T.rgt_cars_upload(h, 4, ctl.ctypes.data, ctl.shape[0]) == 0
        ref.step_with_controls(ctl, 1, ctl.shape[0])
        assert T.rgt_cars_step(h, 1) == 0
        mine, uns = _twin_records(T, h, n_env, cars)
        theirs = reference_records(R, ref, n_env, cars)
        bad = arena_mismatches(mine, theirs, n_env, cars)
        live = uns == 0
        assert not live[bad].any(), f"tick {t}: arenas {bad[live[bad]]} differ from the reference while unflagged"
        ever_flagged |= ~live
        compared += int(live.sum())
    bumps = mine["cars"].view(views.state_dtype("car_states"))["bump_event_bumps"].reshape(n_env, -1).sum(axis=1)
    T.rgt_cars_destroy(h)
    ref.close()
    # Arenas that were never flagged went through many bumps (each bump is one car-car contact's outcome).
    assert bumps[~ever_flagged].sum() >= 100, f"bumps in unflagged arenas: {bumps[~ever_flagged].sum()}"
    assert compared >= 0.9 * n_env * ticks


def test_twin_cars_mixed_long_run_prefix_matches_oracle_chains():
    """The first 64 arenas of the 4,096-arena, 10,000-tick oracle fixture, stepped by the host twin from the
    reference's kickoff for 2,000 ticks: every arena unflagged up to a checkpoint has the oracle's running body-digest
    chain there (arenas are independent, and their kickoff and car ids depend only on the seed and the arena index)."""
    import os

    from digests import FNV_OFFSET, arena_body_digests, chain_step
    import rocketgpu_amd as rga

    R = oracle_or_skip()
    T = twin_cars()
    fx = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cars_mixed_4096_t10000.npz"))
    n, seed, period, every = 64, int(fx["seed"]), int(fx["period"]), int(fx["every"])
    counts, presets, teams = fx["counts"][:n], fx["presets"][:n * 6], fx["teams"][:n * 6]
    ref = R.RefSim(n, 6, respawn_seed=seed)
    ref.set_env_car_count(counts)
    ref.set_teams(teams)
    ref.set_car_presets(presets)
    ref.reset_to_kickoff(seed)
    rec = reference_records(R, ref, n, 6)
    h = T.rgt_cars_create(n, 6, 1.0, 256.0, seed)
    for which, arr in ((6, counts), (7, teams), (5, ref.download_car_configs(n * 6)), (0, rec["cars"]),
                       (8, rec["wheels"]), (1, rec["balls"]), (2, rec["envs"]), (3, rec["pads"])):
        a = np.ascontiguousarray(arr)
        assert T.rgt_cars_upload(h, which, a.ctypes.data, a.shape[0]) == 0
    ref.close()
    controls = rga.workload_controls(seed, 0, n, 6, period)
    chain = chain_step(np.full(n, FNV_OFFSET, dtype=np.uint64),
                       arena_body_digests(rec["balls"], rec["cars"], counts, 6))
    assert np.array_equal(chain, fx["chain"][0][:n])
    flagged = np.zeros(n, bool)
    for t in range(2 * every):
        ctl = np.ascontiguousarray(controls[t % period])
        assert T.rgt_cars_upload(h, 4, ctl.ctypes.data, ctl.shape[0]) == 0
        assert T.rgt_cars_step(h, 1) == 0
        mine, uns = _twin_records(T, h, n, 6)
        flagged |= uns != 0
        chain = chain_step(chain, arena_body_digests(mine["balls"], mine["cars"], counts, 6))
        if (t + 1) % every == 0:
            ref_chain = fx["chain"][(t + 1) // every][:n]
            bad = np.nonzero(~flagged & (chain != ref_chain))[0]
            assert bad.size == 0, f"tick {t + 1}: unflagged arenas {bad} differ from the oracle"
    T.rgt_cars_destroy(h)
    assert (~flagged).sum() >= n // 2
