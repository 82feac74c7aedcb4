"""The golden fixtures: self-consistency, and (where the reference sources were built) reproduction by the oracle."""
import numpy as np
import pytest

from conftest import fixture_names, load_fixture, oracle_or_skip
from digests import ball_env_digests, fold

NAMES = fixture_names()


def test_fixtures_present():
    assert {"ball_arena_96_t1000", "ball_arena_1024_t3000", "ball_arena_small_scale_64_t1500"} <= set(NAMES)


@pytest.mark.parametrize("name", NAMES)
def test_fixture_self_consistent(name):
    f = load_fixture(name)
    n, ticks, every = f["meta"]["arenas"], f["meta"]["ticks"], int(f["checkpoint_every"])
    assert f["init_ball"].shape == (n, 288) and f["final_ball"].shape == (n, 288)
    assert f["init_env"].shape == (n, 44) and f["final_env"].shape == (n, 44)
    assert f["batch_digest"].shape == (ticks,)
    assert f["env_digest"].shape == (ticks // every, n)
    # Every checkpoint row folds to that tick's batch digest.
    ticks_at = np.arange(1, ticks // every + 1) * every - 1
    assert np.array_equal(fold(f["env_digest"]), f["batch_digest"][ticks_at])
    # The final states digest to the last checkpoint (the last tick is a checkpoint in every fixture).
    assert ticks % every == 0
    assert np.array_equal(ball_env_digests(f["final_ball"]), f["env_digest"][-1])
    # EnvState.tick counts the ticks.
    assert np.all(f["final_env"][:, 0:4].copy().view(np.uint32)[:, 0] == ticks)
    # The workload moves: most digests change from tick to tick early on, some arenas score.
    assert len(np.unique(f["batch_digest"])) == ticks


def test_oracle_reproduces_fixture():
    R = oracle_or_skip()
    f = load_fixture("ball_arena_96_t1000")
    m = f["meta"]
    n = m["arenas"]
    s = R.RefSim(n, 2, m["scale"], m["fillet_radius"])
    s.set_env_car_count(np.zeros(n, np.int32))
    s.upload(R.ENV, f["init_env"])
    s.upload(R.BALL, f["init_ball"])
    batch, env = s.run(np.zeros(n, np.int32), 0, m["ticks"], int(f["checkpoint_every"]))
    assert np.array_equal(batch, f["batch_digest"])
    assert np.array_equal(env, f["env_digest"])
    assert np.array_equal(s.download(R.BALL, n), f["final_ball"])


def test_oracle_workload_matches_fixture_inputs():
    R = oracle_or_skip()
    for name in NAMES:
        f = load_fixture(name)
        m = f["meta"]
        assert np.array_equal(R.workload_ball_states(m["seed"], 0, m["arenas"], m["scale"]), f["init_ball"])
        assert np.array_equal(np.tile(R.default_state(R.ENV), (m["arenas"], 1)), f["init_env"])
