"""Typed, zero-copy views of a BatchSim's device arrays: the reference's Python helpers (python/rocketgpu_py_cuda.py:
DeviceArray, device_arrays, state_layout, state_dtype, DeviceFieldArray, field_arrays, host_struct_array) over this
build's C-ABI.

    import rocketgpu_amd as rg
    sim = rg.BatchSim(4096)
    f = rg.field_arrays(sim, "ball_states")         # {"pos": DeviceFieldArray [4096, 3] f4, ...}
    pos = torch.as_tensor(f["pos"], device="cuda")  # zero-copy strided view through __cuda_array_interface__

Records are stored array-of-structs (one record per element, the byte layout of state_layouts()); a field view is
the field's column with the record size as its leading stride.  Views alias the batch's memory: they stay valid
while the batch lives and see every step.
"""
from __future__ import annotations

import numpy as np


class DeviceArray:
    """One state array of a batch (python/rocketgpu_py_cuda.py:4-21)."""

    def __init__(self, name, entry, owner=None):
        self._owner = owner
        self.name = name
        self.__cuda_array_interface__ = entry["__cuda_array_interface__"]
        self.itemsize = entry["itemsize"]
        self.count = entry["count"]

    @property
    def ptr(self):
        return self.__cuda_array_interface__["data"][0]

    @property
    def nbytes(self):
        return self.itemsize * self.count

    def __repr__(self):
        return f"DeviceArray({self.name}, count={self.count}, itemsize={self.itemsize}, ptr=0x{self.ptr:x})"


def device_arrays(sim) -> dict:
    """DeviceArray per state array of `sim` (python/rocketgpu_py_cuda.py:24-25)."""
    return {name: DeviceArray(name, entry, sim) for name, entry in sim.device_arrays().items()}


def state_layout(name: str) -> dict:
    from . import state_layouts

    return state_layouts()[name]


def state_dtype(name: str) -> np.dtype:
    """numpy structured dtype of one record (python/rocketgpu_py_cuda.py:34-45): the listed fields at their offsets,
    itemsize the record's size."""
    layout = state_layout(name)
    names, formats, offsets = [], [], []
    for field, typestr, offset, shape in layout["fields"]:
        names.append(field)
        formats.append((np.dtype(typestr), tuple(shape)) if shape else np.dtype(typestr))
        offsets.append(offset)
    return np.dtype({"names": names, "formats": formats, "offsets": offsets, "itemsize": layout["itemsize"]})


class DeviceFieldArray:
    """A strided device view of one field of every record (python/rocketgpu_py_cuda.py:48-82): shape
    (count, *field_shape), or (envs, per_env, *field_shape) when grouped by env."""

    def __init__(self, array_name, field, typestr, base_ptr, count, itemsize, offset, shape, owner=None,
                 per_env=None, read_only=False):
        elem = np.dtype(typestr)
        inner = []
        stride = elem.itemsize
        for dim in reversed(tuple(shape)):
            inner.insert(0, stride)
            stride *= dim
        lead, lead_strides = (count,), (itemsize,)
        if per_env:
            lead, lead_strides = (count // per_env, per_env), (itemsize * per_env, itemsize)
        self._owner = owner
        self.name = f"{array_name}.{field}"
        self.dtype = elem
        self.shape = lead + tuple(shape)
        self.strides = lead_strides + tuple(inner)
        self.__cuda_array_interface__ = {
            "data": (base_ptr + offset, read_only),
            "shape": self.shape,
            "typestr": typestr,
            "strides": self.strides,
            "version": 3,
        }

    def __repr__(self):
        return f"DeviceFieldArray({self.name}, shape={self.shape}, dtype={self.dtype}, strides={self.strides})"


def field_arrays(sim, array_name: str, by_env: bool = False) -> dict:
    """DeviceFieldArray per field of `array_name` (python/rocketgpu_py_cuda.py:86-109).  With by_env, the leading
    dimension is split into (env, element of the env)."""
    entry = sim.device_arrays()[array_name]
    base_ptr, read_only = entry["__cuda_array_interface__"]["data"]
    count = entry["count"]
    layout = state_layout(array_name)
    itemsize = layout["itemsize"]
    if itemsize != entry["itemsize"]:
        raise ValueError(f"{array_name}: layout itemsize {itemsize} != device itemsize {entry['itemsize']}")
    per_env = entry["per_env"] if by_env else None
    return {
        field: DeviceFieldArray(array_name, field, typestr, base_ptr, count, itemsize, offset, shape, sim,
                                per_env=per_env, read_only=read_only)
        for field, typestr, offset, shape in layout["fields"]
    }


def host_struct_array(sim, array_name: str):
    """The reference maps a CPU build's arrays as numpy records (python/rocketgpu_py_cuda.py:121-134) and raises
    on a GPU build.  This build is GPU-only, so it raises the same way; download_state_array(name).view(
    state_dtype(name)) gives a host copy instead."""
    raise RuntimeError("host_struct_array requires a CPU build")
