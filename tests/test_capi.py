"""The C-ABI library: it loads without a GPU, exports every declared entry point, and its host-side utilities
(state layouts, workload generator, digest fold) agree with the reference build and with the tests' own code.
No GPU computation is started here."""
import ctypes
import glob
import os
import re

import numpy as np
import pytest

from conftest import REPO, load_fixture, oracle_or_skip
from digests import BALL_ANG_VEL, BALL_POS, BALL_ROT, BALL_VEL, fold

import rocketgpu_amd as rga


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(REPO, "include", "rocketgpu_amd", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*[A-Za-z_][\w \*]*?\b(rg_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_header_declares_the_boundary():
    names = declared_functions()
    for must in ("rg_batch_create", "rg_batch_step", "rg_batch_step_with_device_controls", "rg_batch_upload",
                 "rg_batch_download", "rg_batch_device_view", "rg_state_layouts_json", "rg_last_error"):
        assert must in names


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(rga.library_path())
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_version_string():
    assert b"gfx950" in rga.lib().rg_version()


def test_layouts_match_reference_offsets():
    R = oracle_or_skip()
    lay = rga.state_layouts()
    assert lay["ball_states"]["itemsize"] == R.sizeof(R.BALL) == 288
    assert lay["env_states"]["itemsize"] == R.sizeof(R.ENV) == 44
    assert lay["car_states"]["itemsize"] == R.sizeof(R.CAR) == 880
    assert lay["controls"]["itemsize"] == R.sizeof(R.CONTROLS) == 24
    assert lay["boost_pad_states"]["itemsize"] == R.sizeof(R.PAD) == 16
    assert lay["car_configs"]["itemsize"] == R.sizeof(5) == 120
    which = {"ball_states": R.BALL, "env_states": R.ENV, "car_states": R.CAR, "boost_pad_states": R.PAD,
             "controls": 4, "car_configs": 5}
    checked = 0
    for arr, w in which.items():
        for name, _typestr, off, _shape in lay[arr]["fields"]:
            ref = R.offsetof(w, name)
            # Every field of the records this build steps is checked against the reference's offsetof.
            assert ref >= 0, (arr, name)
            if ref >= 0:
                assert off == ref, (arr, name, off, ref)
                checked += 1
    assert checked >= 70
    assert (BALL_POS, BALL_ROT, BALL_VEL, BALL_ANG_VEL) == tuple(
        R.offsetof(R.BALL, f) for f in ("pos", "rot", "vel", "ang_vel"))


def test_layout_fields_fit_itemsize():
    for arr, lay in rga.state_layouts().items():
        for name, typestr, off, shape in lay["fields"]:
            assert isinstance(shape, list)
            size = np.dtype(typestr).itemsize * int(np.prod(shape))
            assert 0 <= off and off + size <= lay["itemsize"], (arr, name)


@pytest.mark.parametrize("seed,first,count,scale", [(7, 0, 96, 1.0), (11, 1000, 24, 1.0), (13, 5, 64, 0.8)])
def test_workload_matches_reference_generator(seed, first, count, scale):
    R = oracle_or_skip()
    assert np.array_equal(rga.workload_ball_states(seed, first, count, scale),
                          R.workload_ball_states(seed, first, count, scale))


def test_workload_matches_fixture_without_oracle():
    f = load_fixture("ball_arena_96_t1000")
    m = f["meta"]
    assert np.array_equal(rga.workload_ball_states(m["seed"], 0, m["arenas"], m["scale"]), f["init_ball"])


def test_fold_matches_numpy():
    L = rga.lib()
    L.rg_fold_digests.restype = ctypes.c_uint64
    L.rg_fold_digests.argtypes = [ctypes.c_void_p, ctypes.c_int64]
    f = load_fixture("ball_arena_96_t1000")
    row = np.ascontiguousarray(f["env_digest"][3])
    assert L.rg_fold_digests(row.ctypes.data, row.size) == int(fold(row)) == int(f["batch_digest"][399])
    assert L.rg_fold_digests(None, 0) == 0xCBF29CE484222325
