"""rg_batch_step_split_no_alloc against rg_batch_step on the GPU, through the C-ABI.

The no-allocation entry point runs the split schedule with its arena lists in the library's fixed device array.
Two calls in a row, the second on a smaller batch (so the array still holds the first call's lists): each equals the
plain step under the split launch shape byte for byte (ball, env and the whole snapshot with the internal state),
and the reference's final records from the committed fixtures.
"""
import sys

import numpy as np
import pytest

from conftest import load_fixture

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _torch():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _sim(f):
    import rocketgpu_amd as rga

    m = f["meta"]
    n = m["arenas"]
    sim = rga.BatchSim(n, 2, m["scale"], m["fillet_radius"])
    sim.set_env_car_count(np.zeros(n, np.int32))
    sim.upload("env_states", f["init_env"])
    sim.upload("ball_states", f["init_ball"])
    return sim


def _state(sim):
    return (sim.download_state_array("ball_states"), sim.download_state_array("env_states"), sim.save_state())


@pytest.mark.parametrize("fixture,stage,launches", [("ball_arena_1024_t3000", 6, 375),
                                                    ("ball_arena_96_t1000", 7, 125)])
def test_no_alloc_step_matches_step(fixture, stage, launches):
    _torch()
    f = load_fixture(fixture)
    say(f"case: {fixture}, split stage {stage}, {launches} launches of 8 ticks")
    plain = _sim(f)
    plain.set_launch_shape(64, stage, 8)
    free = _sim(f)
    if stage == 7:  # 16 lanes per heavy arena; otherwise the batch keeps its default shape and owns no lists
        free.set_launch_shape(64, 7, 8)
    for _ in range(launches):
        plain.step(8)
        free.step_split_no_alloc(8)
    want, got = _state(plain), _state(free)
    plain.close()
    free.close()
    for w, g in zip(want, got):
        assert np.array_equal(np.frombuffer(w, np.uint8) if isinstance(w, bytes) else w,
                              np.frombuffer(g, np.uint8) if isinstance(g, bytes) else g)
    assert np.array_equal(got[0], f["final_ball"])
    assert np.array_equal(got[1], f["final_env"])


def test_no_alloc_step_rejects_car_batches():
    _torch()
    import rocketgpu_amd as rga

    sim = rga.BatchSim(4, 2)
    sim.set_env_car_count(np.array([1, 0, 0, 0], np.int32))
    with pytest.raises(rga.RocketGpuError, match="ball-only"):
        sim.step_split_no_alloc(1)
    sim.close()
