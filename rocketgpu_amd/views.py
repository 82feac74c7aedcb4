"""Typed, zero-copy views of a BatchSim's device arrays, under the names of the reference's Python helpers
(python/rocketgpu_py_cuda.py: DeviceArray, device_arrays, state_layout, state_dtype, DeviceFieldArray,
field_arrays, host_struct_array).

    import rocketgpu_amd as rg
    sim = rg.BatchSim(4096)
    f = rg.field_arrays(sim, "ball_states")         # {"pos": DeviceFieldArray [4096, 3] f4, ...}
    pos = torch.as_tensor(f["pos"], device="cuda")  # zero-copy strided view through __cuda_array_interface__

How the views are built here: every state array is a run of fixed-size records in HBM (the byte layout that
state_layouts() lists).  A record layout becomes one numpy structured dtype, and a field view is described by a
numpy array of zero length over that dtype: numpy computes the field's element shape and its strides inside the
record, and the view only adds the leading record dimension (or the env x element dimensions).  The views alias
the batch's memory: they stay valid while the batch lives and see every step.
"""
from __future__ import annotations

import numpy as np


def state_layout(name: str) -> dict:
    """The record layout of one state array: {"itemsize": bytes, "fields": [(name, typestr, offset, shape)]}."""
    from . import state_layouts

    return state_layouts()[name]


def _record_dtype(layout: dict) -> np.dtype:
    spec = {"names": [], "formats": [], "offsets": [], "itemsize": layout["itemsize"]}
    for field, typestr, offset, shape in layout["fields"]:
        spec["names"].append(field)
        spec["formats"].append(np.dtype((np.dtype(typestr), tuple(shape))))
        spec["offsets"].append(offset)
    return np.dtype(spec)


def state_dtype(name: str) -> np.dtype:
    """numpy structured dtype of one record of `name`: every listed field at its byte offset, itemsize the record
    size (padding bytes unnamed), so download_state_array(name).view(state_dtype(name)) decodes a host copy."""
    return _record_dtype(state_layout(name))


class DeviceArray:
    """One whole state array of a batch: device pointer, record size and count, and the array's
    __cuda_array_interface__ (a flat run of records)."""

    __slots__ = ("name", "itemsize", "count", "_entry", "_owner")

    def __init__(self, name, entry, owner=None):
        self.name = name
        self.itemsize = int(entry["itemsize"])
        self.count = int(entry["count"])
        self._entry = entry
        self._owner = owner  # keeps the batch (and so the memory) alive while the view exists

    @property
    def __cuda_array_interface__(self):
        return self._entry["__cuda_array_interface__"]

    @property
    def ptr(self) -> int:
        return int(self._entry["__cuda_array_interface__"]["data"][0])

    @property
    def nbytes(self) -> int:
        return self.count * self.itemsize

    def __repr__(self):
        return f"<DeviceArray {self.name}: {self.count} x {self.itemsize} B at {self.ptr:#x}>"


def device_arrays(sim) -> dict:
    """Every state array of `sim` as a DeviceArray, by name."""
    out = {}
    for name, entry in sim.device_arrays().items():
        out[name] = DeviceArray(name, entry, sim)
    return out


class DeviceFieldArray:
    """One field of every record of a state array, as a strided device array.

    `lead` is the leading shape over the records, (count,) or (envs, per_env), with `lead_strides` its byte
    strides; the field's own shape and strides come from the record dtype."""

    def __init__(self, array_name, field, record, base_ptr, lead, lead_strides, owner=None, read_only=False):
        sub = np.zeros(0, record)[field]  # numpy derives the field's element shape and in-record strides
        self._owner = owner
        self.name = f"{array_name}.{field}"
        self.dtype = sub.dtype
        self.shape = tuple(lead) + sub.shape[1:]
        self.strides = tuple(lead_strides) + sub.strides[1:]
        self._ptr = base_ptr + record.fields[field][1]
        self._read_only = read_only
        self._typestr = sub.dtype.str

    @property
    def __cuda_array_interface__(self):
        return {"shape": self.shape, "strides": self.strides, "typestr": self._typestr,
                "data": (self._ptr, self._read_only), "version": 3}

    def __repr__(self):
        return f"<DeviceFieldArray {self.name}: {self.shape} {self.dtype} strides {self.strides}>"


def field_arrays(sim, array_name: str, by_env: bool = False) -> dict:
    """A DeviceFieldArray per field of `array_name`.  With by_env the records are viewed as (env, element of the
    env), e.g. boost pads as [envs, pads per env]."""
    entry = sim.device_arrays()[array_name]
    record = state_dtype(array_name)
    if record.itemsize != entry["itemsize"]:
        raise ValueError(f"{array_name}: layout record of {record.itemsize} B, device record of {entry['itemsize']} B")
    base_ptr, read_only = entry["__cuda_array_interface__"]["data"]
    count = entry["count"]
    size = record.itemsize
    if by_env:
        k = entry["per_env"]
        lead, lead_strides = (count // k, k), (size * k, size)
    else:
        lead, lead_strides = (count,), (size,)
    return {field: DeviceFieldArray(array_name, field, record, base_ptr, lead, lead_strides, sim, read_only)
            for field in record.names}


def host_struct_array(sim, array_name: str):
    """The reference maps a CPU build's state arrays as host record arrays and raises on a GPU build; this library
    is GPU-only, so this raises.  download_state_array(name).view(state_dtype(name)) is a host copy."""
    raise RuntimeError("host_struct_array requires a CPU build")
