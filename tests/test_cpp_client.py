"""The C++ BatchSim header (include/rocketgpu_amd/batch_sim.hpp) driven by a compiled C++ caller
(tests/cpp/batch_sim_client.cpp, built by __graft_entry__.build()): a mixed 1v1 / 2v2 batch with a non-default
mutator, a kickoff, seeded device controls, copy_state_from and asynchronous downloads.  Its digest of the final car
and ball records must equal the same run through the Python binding, and a fixed mutator field must be refused
with std::invalid_argument."""
import os
import subprocess

import numpy as np
import pytest

from conftest import REPO

pytestmark = pytest.mark.gpu
CLIENT = os.path.join(REPO, "rocketgpu_amd", "_build", "batch_sim_client")


def _fnv(data: bytes, h: int = 0xCBF29CE484222325) -> int:
    a = np.frombuffer(data, np.uint8)
    for b in a.tolist():
        h = ((h ^ b) * 0x100000001B3) & 0xFFFFFFFFFFFFFFFF
    return h


def test_cpp_client_matches_python_binding():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    if not os.path.exists(CLIENT):
        pytest.fail("batch_sim_client is not built")
    import rocketgpu_amd as rga

    n, ticks, seed = 16, 40, 5
    run = subprocess.run([CLIENT, str(n), str(ticks), str(seed)], capture_output=True, text=True, timeout=120)
    print(run.stdout, run.stderr, flush=True)
    assert run.returncode == 0
    lines = run.stdout.strip().splitlines()
    assert lines[0] == "mutator gravity_z -600.0"
    assert lines[1].startswith("refused: mutator field sticky_force")
    assert "copy_state_from equal" in lines
    digest = lines[-1].split()[1]

    counts = np.array([2 + 2 * (e % 2) for e in range(n)], np.int32)
    cap = n * 4
    teams = np.array([1 if (i % 4) >= counts[i // 4] // 2 else 0 for i in range(cap)], np.uint8)
    cfgs = np.stack([rga.car_preset_config(i % 6) for i in range(cap)])
    sim = rga.BatchSim(n, 4)
    sim.set_env_car_count(counts)
    sim.set_teams(teams)
    sim.set_car_configs(cfgs)
    sim.set_mutator_config(gravity_z=-600.0)
    sim.reset_to_kickoff(seed)

    def stage():
        cars = sim.download_state_array("car_states").tobytes()
        balls = sim.download_state_array("ball_states").tobytes()
        return f"{_fnv(balls, _fnv(cars)):016x}"

    stages = [stage()]
    ctrl = torch.from_numpy(rga.workload_controls(seed, 0, n, 4, ticks).reshape(ticks, -1)).cuda()
    for t in range(0, ticks, 8):
        sim.step_with_device_controls(ctrl.data_ptr() + t * ctrl.shape[1], min(8, ticks - t), cap)
        stages.append(stage())
    sim.close()
    cpp_stages = [line.split()[2] for line in lines if line.startswith("stage ")]
    assert cpp_stages == stages
    assert stages[-1] == digest
