"""Digest helpers shared by the tests, written independently of the C++ code they check.

Body digest (rocketgpu_amd/workload.h BodyDigest, following tools/bench_cpp.cpp:505-525 of the reference):
FNV-1a 64 over the raw bits of pos (3 f32), vel (3 f32), ang_vel (3 f32), rot (w, x, y, z f32) of the ball, then of
each car in slot order.  Batch digest of one tick: FNV-1a 64 over the little-endian bytes of the per-env digests.
"""
from __future__ import annotations

import numpy as np

FNV_OFFSET = np.uint64(0xCBF29CE484222325)
FNV_PRIME = np.uint64(0x100000001B3)

# BallState field offsets (bytes) of the fields the digest reads; checked against the reference build's
# offsetof() in test_capi.py.
BALL_POS, BALL_ROT, BALL_VEL, BALL_ANG_VEL = 40, 52, 68, 80


def fnv_rows(rows: np.ndarray, h0=None) -> np.ndarray:
    """FNV-1a 64 of each row of a [R, B] uint8 array (vectorized over rows)."""
    rows = np.ascontiguousarray(rows, dtype=np.uint8)
    h = np.full(rows.shape[0], FNV_OFFSET if h0 is None else h0, dtype=np.uint64)
    with np.errstate(over="ignore"):
        for j in range(rows.shape[1]):
            h ^= rows[:, j].astype(np.uint64)
            h *= FNV_PRIME
    return h


def fold(env_digests: np.ndarray) -> np.ndarray:
    """Batch digests of [T, N] (or [N]) per-env digests."""
    d = np.atleast_2d(np.ascontiguousarray(env_digests, dtype=np.uint64))
    out = fnv_rows(d.view(np.uint8).reshape(d.shape[0], -1))
    return out if np.ndim(env_digests) == 2 else out[0]


def ball_env_digests(balls: np.ndarray) -> np.ndarray:
    """Per-env digests of ball-only arenas from [N, 288] BallState bytes."""
    b = np.ascontiguousarray(balls, dtype=np.uint8)
    rot = b[:, BALL_ROT:BALL_ROT + 16]  # stored w, x, y, z
    cols = np.concatenate([b[:, BALL_POS:BALL_POS + 12], b[:, BALL_VEL:BALL_VEL + 12],
                           b[:, BALL_ANG_VEL:BALL_ANG_VEL + 12], rot], axis=1)
    return fnv_rows(cols)


# CarState field offsets (bytes) of the fields the digest reads (rocketgpu_amd.views car_states).
CAR_POS, CAR_ROT, CAR_VEL, CAR_ANG_VEL = 0, 12, 28, 40


def _fnv_rows_from(rows: np.ndarray, h: np.ndarray) -> np.ndarray:
    """FNV-1a 64 of each row of a [R, B] uint8 array, continuing from per-row states h."""
    h = h.copy()
    with np.errstate(over="ignore"):
        for j in range(rows.shape[1]):
            h ^= rows[:, j].astype(np.uint64)
            h *= FNV_PRIME
    return h


def arena_body_digests(balls: np.ndarray, cars: np.ndarray, counts: np.ndarray, max_cars: int) -> np.ndarray:
    """Per-arena body digests of car arenas: the ball's pos, vel, ang_vel, rot, then each of the arena's cars' in
    slot order, as one FNV-1a 64 stream.  balls [N, 288] u8, cars [N * max_cars, 880] u8, counts [N]."""
    b = np.ascontiguousarray(balls, dtype=np.uint8)
    c = np.ascontiguousarray(cars, dtype=np.uint8).reshape(b.shape[0], max_cars, -1)
    ball_cols = np.concatenate([b[:, BALL_POS:BALL_POS + 12], b[:, BALL_VEL:BALL_VEL + 12],
                                b[:, BALL_ANG_VEL:BALL_ANG_VEL + 12], b[:, BALL_ROT:BALL_ROT + 16]], axis=1)
    h = _fnv_rows_from(ball_cols, np.full(b.shape[0], FNV_OFFSET, dtype=np.uint64))
    for k in range(max_cars):
        ck = c[:, k]
        cols = np.concatenate([ck[:, CAR_POS:CAR_POS + 12], ck[:, CAR_VEL:CAR_VEL + 12],
                               ck[:, CAR_ANG_VEL:CAR_ANG_VEL + 12], ck[:, CAR_ROT:CAR_ROT + 16]], axis=1)
        h = np.where(k < np.asarray(counts), _fnv_rows_from(cols, h), h)
    return h


def chain_step(chain: np.ndarray, digests: np.ndarray) -> np.ndarray:
    """One step of the per-arena running chain: FNV-1a 64 over the 8 little-endian bytes of the new digest."""
    d = np.ascontiguousarray(digests, dtype=np.uint64)
    return _fnv_rows_from(d.view(np.uint8).reshape(-1, 8), np.asarray(chain, dtype=np.uint64))
