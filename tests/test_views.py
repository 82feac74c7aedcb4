"""Typed zero-copy views (the reference's python/rocketgpu_py_cuda.py helpers and BatchSim.device_arrays,
python/rocketgpu_bindings.cpp:846-898) over this build's device arrays."""
import numpy as np
import pytest

import rocketgpu_amd as rga

ARRAYS = ("controls", "car_configs", "teams", "env_car_count", "car_states", "ball_states", "boost_pad_states",
          "env_states")


def test_layouts_name_the_reference_arrays():
    lay = rga.state_layouts()
    for name in ARRAYS:
        assert name in lay, name
        assert lay[name]["itemsize"] == rga.ITEMSIZE[name]


@pytest.mark.parametrize("name", ARRAYS)
def test_state_dtype_places_every_field(name):
    dt = rga.state_dtype(name)
    lay = rga.state_layout(name)
    assert dt.itemsize == lay["itemsize"]
    for field, typestr, off, shape in lay["fields"]:
        assert dt.fields[field][1] == off
        assert dt.fields[field][0].base == np.dtype(typestr)
        assert dt.fields[field][0].shape == tuple(shape)


def test_state_dtype_decodes_default_records():
    """A default BallState decodes to the reference's defaults (rocketgpu.h BallState initializers)."""
    from conftest import oracle_or_skip

    R = oracle_or_skip()
    b = R.default_state(R.BALL).view(rga.state_dtype("ball_states"))[0]
    assert tuple(b["pos"]) == (0.0, 0.0, np.float32(93.15))
    assert tuple(b["rot"]) == (1.0, 0.0, 0.0, 0.0)
    assert b["hs_info.cur_target_speed"] == 2900.0 and b["ds_info.charge_level"] == 1
    pad = np.zeros(1, rga.state_dtype("boost_pad_states"))
    assert pad.itemsize == 16


def test_host_struct_array_needs_a_cpu_build():
    with pytest.raises(RuntimeError, match="CPU build"):
        rga.host_struct_array(None, "ball_states")


def test_field_array_shapes_and_strides():
    entry = {"__cuda_array_interface__": {"data": (0x10000, False), "shape": (12,), "typestr": "|V16",
                                          "strides": None, "version": 3},
             "itemsize": 16, "count": 12, "per_env": 4, "element": "boost pad"}

    class Fake:
        def device_arrays(self):
            return {"boost_pad_states": entry}

    f = rga.field_arrays(Fake(), "boost_pad_states")
    cai = f["cooldown"].__cuda_array_interface__
    assert cai["shape"] == (12,) and cai["strides"] == (16,) and cai["data"][0] == 0x10000
    g = rga.field_arrays(Fake(), "boost_pad_states", by_env=True)
    cai = g["pickup_frac_q"].__cuda_array_interface__
    assert cai["shape"] == (3, 4) and cai["strides"] == (64, 16) and cai["data"][0] == 0x10006
    assert cai["typestr"] == "<u2"


@pytest.mark.gpu
def test_gpu_field_arrays_alias_the_batch():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    n = 64
    sim = rga.BatchSim(n, 2)
    sim.upload("ball_states", rga.workload_ball_states(3, 0, n))
    arrays = rga.device_arrays(sim)
    assert arrays["ball_states"].count == n and arrays["boost_pad_states"].count == n * 34
    f = rga.field_arrays(sim, "ball_states")
    pos = torch.as_tensor(f["pos"], device="cuda")
    rot = torch.as_tensor(f["rot"], device="cuda")
    for _ in range(2):
        torch.cuda.synchronize()
        host = sim.download_state_array("ball_states").view(rga.state_dtype("ball_states"))[:, 0]
        assert np.array_equal(pos.cpu().numpy(), host["pos"])
        assert np.array_equal(rot.cpu().numpy(), host["rot"])
        sim.step(8)  # the views see the step: same memory
    pads = torch.as_tensor(rga.field_arrays(sim, "boost_pad_states", by_env=True)["is_big"], device="cuda")
    assert tuple(pads.shape) == (n, 34)
    assert (pads.cpu().numpy() == (np.arange(34) >= 28)).all()
    t = sim.torch_arrays()["env_states"]
    assert tuple(t.shape) == (n, 44)
    torch.cuda.synchronize()
    sim.close()
