"""Seeded synthetic data: SplitMix64 streams, vectorised with numpy.

`mix` and `record` restate the reference's synthetic records (src/service.rs:43-62):
a SplitMix64 generator whose state advances by 0x9e3779b97f4a7c15 per draw, and records
made of little-endian u16 words `mix % 65535`.  `stream(seed, count)` is the same
generator started at state `seed`: draw i (0-based) is the finalised value of
seed + (i + 1) * 0x9e3779b97f4a7c15, which lets numpy produce any slice of the stream at
once.  All benchmark and test inputs of this package come from such streams.
"""
from __future__ import annotations

import numpy as np

GAMMA = np.uint64(0x9E3779B97F4A7C15)
M1 = np.uint64(0xBF58476D1CE4E5B9)
M2 = np.uint64(0x94D049BB133111EB)
SEED = 0x416C746F504952  # src/service.rs:44
ROW_MUL = 0xD1B54A32D192ED03  # src/service.rs:56
Q = 4294955009

_CHUNK = 1 << 23


def _finalise(z: np.ndarray) -> np.ndarray:
    z = (z ^ (z >> np.uint64(30))) * M1
    z = (z ^ (z >> np.uint64(27))) * M2
    return z ^ (z >> np.uint64(31))


def mix(state: int) -> tuple[int, int]:
    """One draw of the reference generator (src/service.rs:46-52): returns (value, new_state)."""
    m = (1 << 64) - 1
    state = (state + 0x9E3779B97F4A7C15) & m
    z = state
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & m
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & m
    return z ^ (z >> 31), state


def stream(seed: int, count: int, start: int = 0) -> np.ndarray:
    """Draws start .. start+count-1 of the SplitMix64 stream started at state `seed` (u64)."""
    out = np.empty(count, dtype=np.uint64)
    base = np.uint64(seed & ((1 << 64) - 1))
    with np.errstate(over="ignore"):
        for lo in range(0, count, _CHUNK):
            hi = min(count, lo + _CHUNK)
            idx = np.arange(start + lo + 1, start + hi + 1, dtype=np.uint64)
            out[lo:hi] = _finalise(base + idx * GAMMA)
    return out


def uniform_bytes(seed: int, nbytes: int) -> np.ndarray:
    """nbytes i.i.d. uniform bytes: the little-endian bytes of the stream, in order."""
    return stream(seed, (nbytes + 7) // 8).view(np.uint8)[:nbytes]


def uniform_bytes_device(seed: int, nbytes: int, device="cuda", chunk: int = 1 << 27):
    """uniform_bytes(seed, nbytes) generated on the GPU with torch (int64 arithmetic wraps
    like u64; the logical right shifts are arithmetic shifts masked to their width)."""
    import torch

    words = (nbytes + 7) // 8
    out = torch.empty(words * 8, dtype=torch.uint8, device=device)
    w64 = out.view(torch.int64)
    s64 = lambda v: v - (1 << 64) if v >= (1 << 63) else v  # noqa: E731
    seed64 = s64(seed & ((1 << 64) - 1))
    g, m1, m2 = s64(0x9E3779B97F4A7C15), s64(0xBF58476D1CE4E5B9), s64(0x94D049BB133111EB)
    for lo in range(0, words, chunk):
        hi = min(words, lo + chunk)
        z = torch.arange(lo + 1, hi + 1, dtype=torch.int64, device=device)
        z.mul_(g).add_(seed64)
        z.bitwise_xor_((z >> 30) & ((1 << 34) - 1))
        z.mul_(m1)
        z.bitwise_xor_((z >> 27) & ((1 << 37) - 1))
        z.mul_(m2)
        z.bitwise_xor_((z >> 31) & ((1 << 33) - 1))
        w64[lo:hi] = z
        del z
    return out[:nbytes]


def uniform_u32(seed: int, count: int) -> np.ndarray:
    """count i.i.d. uniform u32 values (low halves of the stream)."""
    return (stream(seed, count) & np.uint64(0xFFFFFFFF)).astype(np.uint32)


def uniform_below_q(seed: int, count: int) -> np.ndarray:
    """count values in [0, Q) (stream mod Q)."""
    return (stream(seed, count) % np.uint64(Q)).astype(np.uint32)


def record(row: int, nbytes: int) -> np.ndarray:
    """The reference's synthetic record `row` (src/service.rs:54-62)."""
    assert nbytes % 2 == 0
    state = (SEED ^ ((ROW_MUL * (row + 1)) & ((1 << 64) - 1))) & ((1 << 64) - 1)
    words = (stream(state, nbytes // 2) % np.uint64(65535)).astype("<u2")
    return words.view(np.uint8).copy()
