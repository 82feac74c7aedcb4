"""Writes tests/golden/math_checksums.json: order-free checksums of the reference build's sinf, cosf and acosf
(oracle/_ref, RocketGpuSinf / RocketGpuCosf / RocketGpuUcrtAcosf) over all 2^32 float bit patterns.

checksum(fn) = sum over i in [0, 2^32) of splitmix64(i << 32 | bits(fn(float with bits i))) mod 2^64, a NaN result
counted as 0x7fc00000.  tests/test_gpu_math.py computes the same sum for the gfx950 build of the step core's math.

    python tests/golden/gen_math_checksums.py        (about a minute on 8 cores)
"""
import json
import os
import sys
from multiprocessing import Pool

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, REPO)
CHUNK = 1 << 22
FNS = {"sin": 0, "cos": 1, "acos": 4}


def mix64(z):
    z = z + np.uint64(0x9E3779B97F4A7C15)
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def chunk_sum(args):
    fn, start = args
    from oracle import refsim as R

    i = np.arange(start, start + CHUNK, dtype=np.uint64)
    x = i.astype(np.uint32).view(np.float32)
    r = R.math(fn, x)
    b = r.view(np.uint32).copy()
    b[np.isnan(r)] = 0x7FC00000
    with np.errstate(over="ignore"):
        return int(mix64((i << np.uint64(32)) | b.astype(np.uint64)).sum(dtype=np.uint64))


def main():
    out = {}
    with Pool(min(8, os.cpu_count() or 1)) as pool:
        for name, fn in FNS.items():
            parts = pool.map(chunk_sum, [(fn, s) for s in range(0, 1 << 32, CHUNK)])
            out[name] = f"{sum(parts) % (1 << 64):016x}"
            print(name, out[name], flush=True)
    with open(os.path.join(REPO, "tests", "golden", "math_checksums.json"), "w") as fh:
        json.dump({"inputs": "all 2^32 float bit patterns", "mix": "splitmix64(i << 32 | result bits), NaN as "
                   "0x7fc00000, summed mod 2^64", "source": "oracle/_ref", "checksums": out}, fh, indent=1)


if __name__ == "__main__":
    main()
