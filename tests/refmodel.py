"""Request-level test models restated from the reference's integration tests.

- decoder_accepts: the accept/reject rule of decode_request (src/service.rs:200-225),
  the rejection oracle of tests/engine.rs:149-151,188-190.
- hand_frames: the eight hand-made frames and their statuses (tests/engine.rs:130-148).
- mutate: the frame-corruption model of tests/engine.rs:29-68, driven by the SplitMix64
  draw of src/service.rs:46-52 from seed 0x5eed_f00d (tests/engine.rs:170).
"""
from __future__ import annotations

import struct

import numpy as np

from altopir_public_amd import synthetic

Q = synthetic.Q
FUZZ_SEED = 0x5EED_F00D
SPECIAL = [0, 1, Q - 1, Q, Q + 1, 0xFFFFFFFF, 1 << 31]  # tests/engine.rs:31 (as u32)


def decoder_accepts(f, rows: int, kp: int) -> bool:
    if f is None:
        return False
    f = bytes(f)
    if len(f) != 9 + 4 * (2 * kp + rows) or f[0] != 0:
        return False
    if struct.unpack_from("<I", f, 1)[0] != kp or struct.unpack_from("<I", f, 5)[0] != kp:
        return False
    keys = np.frombuffer(f, dtype="<u4", count=2 * kp, offset=9)
    return bool((keys < Q).all())


def hand_frames(r0: np.ndarray, r1: np.ndarray, kp: int):
    """Frames and expected statuses of tests/engine.rs:130-148."""
    r0 = bytes(r0)
    frames = [r0, r0[:-1], r0 + b"\x00"]
    v = bytearray(r0); v[0] = 1; frames.append(bytes(v))
    v = bytearray(r0); v[1:5] = struct.pack("<I", 0xFFFFFFFF); frames.append(bytes(v))
    v = bytearray(r0); v[9:13] = struct.pack("<I", Q & 0xFFFFFFFF); frames.append(bytes(v))
    v = bytearray(r0); off = 9 + 8 * kp - 4; v[off:off + 4] = struct.pack("<I", 0xFFFFFFFF); frames.append(bytes(v))
    frames.append(bytes(r1))
    return frames, [0, 1, 1, 2, 2, 3, 3, 0]


class Rnd:
    def __init__(self, seed: int = FUZZ_SEED):
        self.state = seed

    def __call__(self) -> int:
        v, self.state = synthetic.mix(self.state)
        return v


def mutate(valid: bytes, rnd: Rnd, count: int) -> bytes:
    """tests/engine.rs:29-68 (u64 draws, usize/u8 truncations as in Rust)."""
    v = bytearray(valid)
    sel = rnd() % 8
    if sel == 0:
        n = rnd() % len(v)
        del v[n:]
    elif sel == 1:
        for _ in range(1 + rnd() % 64):
            v.append(rnd() & 0xFF)
    elif sel == 2:
        i = rnd() % 13
        v[i] = rnd() & 0xFF
    elif sel == 3:
        i = rnd() % 13
        v[i] ^= 1 << (rnd() % 8)
    elif sel == 4:
        w = 9 + 4 * (rnd() % (2 * count))
        x = SPECIAL[rnd() % len(SPECIAL)] & 0xFFFFFFFF
        v[w:w + 4] = struct.pack("<I", x)
    elif sel == 5:
        nq = (len(v) - 9 - 8 * count) // 4
        w = 9 + 8 * count + 4 * (rnd() % nq)
        x = SPECIAL[rnd() % len(SPECIAL)] & 0xFFFFFFFF
        v[w:w + 4] = struct.pack("<I", x)
    elif sel == 6:
        i = rnd() % len(v)
        v[i] ^= 1 << (rnd() % 8)
    return bytes(v)


def fuzz_batches(base: list, count: int, batches: int = 8, per: int = 64):
    """The 8 x 64 mutated frames of tests/engine.rs:170-192 (base frames chosen by rnd() % 4)."""
    rnd = Rnd()
    out = []
    for _ in range(batches):
        frames = []
        for _ in range(per):
            k = rnd() % len(base)
            frames.append(mutate(bytes(base[k]), rnd, count))
        out.append(frames)
    return out
