import ctypes
import json
import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(REPO, "tests", "golden")
if REPO not in sys.path:
    sys.path.insert(0, REPO)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library on cuda:0)")


def load_fixture(name):
    d = dict(np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False))
    d["meta"] = json.loads(bytes(d["meta"]).decode())
    return d


def fixture_names(prefix="ball_arena_"):
    return sorted(f[:-4] for f in os.listdir(GOLDEN) if f.endswith(".npz") and f.startswith(prefix))


TWIN_PATH = os.path.join(REPO, "rocketgpu_amd", "_build", "librg_twin.so")


def twin():
    """The host-compiled twin of the step core (test-only library)."""
    if not os.path.exists(TWIN_PATH):
        pytest.skip("host twin not built (run __graft_entry__.build())")
    T = ctypes.CDLL(TWIN_PATH)
    vp = ctypes.c_void_p
    T.rgt_create.restype = vp
    T.rgt_create.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double]
    T.rgt_destroy.argtypes = [vp]
    T.rgt_upload.argtypes = [vp, ctypes.c_int, vp, ctypes.c_long]
    T.rgt_download.argtypes = [vp, ctypes.c_int, vp, ctypes.c_long]
    T.rgt_step.argtypes = [vp, ctypes.c_int]
    T.rgt_env_digests.argtypes = [vp, vp, vp]
    T.rgt_math.argtypes = [ctypes.c_int, vp, vp, vp, ctypes.c_long]
    T.rgt_bvh_containment_violations.argtypes = [vp]
    T.rgt_bvh_containment_violations.restype = ctypes.c_long
    T.rgt_ball_path.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_float, vp, vp]
    T.rgt_mesh_tables.argtypes = [vp, vp, vp, vp, vp, vp]
    T.rgt_set_kickoff_settled.argtypes = [vp, ctypes.c_int]
    T.rgt_set_kickoff_settled.restype = None
    T.rgt_settled_ball.argtypes = [vp, vp]
    T.rgt_reset_to_kickoff.argtypes = [vp, ctypes.c_uint64]
    T.rgt_apply_env_commands.argtypes = [vp, vp, ctypes.c_long]
    T.rgt_step_with_commands.argtypes = [vp, vp, ctypes.c_int, ctypes.c_int]
    T.rgt_set_event_stream.argtypes = [vp, vp, vp, ctypes.c_uint32, ctypes.c_int]
    T.rgt_set_event_stream.restype = None
    T.rgt_car_pretick.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_uint64, vp, vp,
                                  vp, vp, vp, vp]
    T.rgt_car_pretick.restype = ctypes.c_long
    T.rgt_car_params.argtypes = [vp]
    T.rgt_quantize_axis.argtypes = [vp, vp, ctypes.c_long]
    T.rgt_quantize_axis.restype = None
    T.rgt_car_posttick.argtypes = [ctypes.c_long, vp, vp, vp, ctypes.c_int]
    T.rgt_car_posttick.restype = None
    T.rgt_car_force_phase.argtypes = [ctypes.c_long, vp, vp, vp, vp, ctypes.c_int, ctypes.c_int, vp]
    return T


def oracle_or_skip():
    from oracle import refsim

    if not refsim.available():
        pytest.skip("oracle/_ref not built (reference sources absent)")
    return refsim


def twin_cars():
    """The host twin's car-arena API (rgt_cars_*)."""
    T = twin()
    vp = ctypes.c_void_p
    T.rgt_cars_create.restype = vp
    T.rgt_cars_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_uint64]
    T.rgt_cars_destroy.argtypes = [vp]
    T.rgt_cars_upload.argtypes = [vp, ctypes.c_int, vp, ctypes.c_long]
    T.rgt_cars_download.argtypes = [vp, ctypes.c_int, vp, ctypes.c_long]
    T.rgt_cars_step.argtypes = [vp, ctypes.c_int]
    return T
