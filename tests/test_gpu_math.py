"""The step core's math on gfx950 (the test-only phase library's kernels) against the reference build and the
host twin.

- sinf, cosf (with the large-argument path) and acosf over all 2^32 float inputs: an order-free checksum of every
  result bit pattern must equal the reference build's (tests/golden/math_checksums.json, written by
  tests/golden/gen_math_checksums.py; the host twin reproduces the same three checksums).
- atan2f, powf and the double SafeAtan2 on 4 Mi seeded argument pairs: every result equals the host twin's bit for
  bit, except that SafeAtan2 may differ by one float ulp where the device's and the host C library's double atan2
  (each within an ulp of double) round to different floats.
"""
import ctypes
import json
import os

import numpy as np
import pytest

from conftest import REPO, twin

pytestmark = pytest.mark.gpu
PHASES_PATH = os.path.join(REPO, "rocketgpu_amd", "_build", "librg_phases.so")


def _phases():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    if not os.path.exists(PHASES_PATH):
        pytest.fail("librg_phases.so is not built")
    P = ctypes.CDLL(PHASES_PATH)
    P.rgp_math_checksum.argtypes = [ctypes.c_int32, ctypes.c_uint64, ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64)]
    P.rgp_math_eval.argtypes = [ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                ctypes.c_void_p]
    return torch, P


@pytest.mark.parametrize("name,fn", [("sin", 0), ("cos", 1), ("acos", 4)])
def test_gpu_math_exhaustive_checksum(name, fn):
    _, P = _phases()
    with open(os.path.join(REPO, "tests", "golden", "math_checksums.json")) as fh:
        want = json.load(fh)["checksums"][name]
    got = ctypes.c_uint64()
    assert P.rgp_math_checksum(fn, 0, 1 << 32, ctypes.byref(got)) == 0
    print(f"{name}: {got.value:016x} (reference {want})", flush=True)
    assert f"{got.value:016x}" == want


def _pairs(fn, n=1 << 22, seed=7):
    rng = np.random.default_rng(seed + fn)
    if fn == 3:  # powf over the step's range: x in [2^-8, 2^8], y in [-8, 8]
        x = np.exp2(rng.uniform(-8, 8, n)).astype(np.float32)
        y = rng.uniform(-8, 8, n).astype(np.float32)
        return x, y
    bits = rng.integers(0, 1 << 32, (2, n), dtype=np.uint64).astype(np.uint32)
    return bits[0].view(np.float32), bits[1].view(np.float32)


@pytest.mark.parametrize("fn", [0, 1, 2, 3, 4, 5])
def test_gpu_math_matches_twin(fn):
    torch, P = _phases()
    T = twin()
    x, y = _pairs(fn)
    dx, dy = torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda()
    dout = torch.empty_like(dx)
    assert P.rgp_math_eval(fn, dx.data_ptr(), dy.data_ptr(), dout.data_ptr(), x.size, None) == 0
    torch.cuda.synchronize()
    got = dout.cpu().numpy()
    want = np.empty_like(x)
    assert T.rgt_math(fn, x.ctypes.data, y.ctypes.data, want.ctypes.data, x.size) == 0
    same = (got.view(np.uint32) == want.view(np.uint32)) | (np.isnan(got) & np.isnan(want))
    bad = np.flatnonzero(~same)
    if fn == 5:
        ulps = np.abs(got[bad].view(np.int32).astype(np.int64) - want[bad].view(np.int32).astype(np.int64))
        print(f"SafeAtan2: {bad.size} of {x.size} results one ulp apart", flush=True)
        assert (ulps <= 1).all() and bad.size <= x.size // 100000
        return
    assert bad.size == 0, f"{bad.size} differ, e.g. x={x[bad[:4]]} y={y[bad[:4]]}"
