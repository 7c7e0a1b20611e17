"""Seeded engine setup inputs and request frames for tests and benchmarks.

The answer engine treats its setup artefacts -- the database, the packing matrix U, the
tail keys `that`, the inverse-NTT table and the row-0 prefixes -- as opaque inputs
(src/amx_kernels.cpp:546-560), so any seeded values exercise the same arithmetic as the
offline precomputation of a real server.  Shapes and the tail layout follow the
reference's parameters: q1 = 2^10, q2 = 2^18, row 1 starts at bit 18 * n, a part is
ceil((18 + 10) * n / 64) * 8 bytes (src/service.rs:164-167, src/modulus_switch.rs:136-147).

Requests follow the wire format of src/service.rs:178-198: flag byte 0, the two key
lengths (u32 LE, = tg * n), y and z (u32 LE, each < Q), then one u32 query word per record.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import synthetic

Q = synthetic.Q


def part_len_for(n: int, q1_bits: int = 10, q2_bits: int = 18) -> int:
    """Bytes of one response part (src/modulus_switch.rs:139-140)."""
    return ((q2_bits * n + q1_bits * n + 63) // 64) * 64 // 8


def request_len(rows: int, kp: int) -> int:
    return 9 + 4 * (2 * kp + rows)


def bitrev(x: int, bits: int) -> int:
    r = 0
    for _ in range(bits):
        r = (r << 1) | (x & 1)
        x >>= 1
    return r


def primitive_root_2n(n: int) -> int:
    """Smallest primitive 2n-th root of unity mod Q (2n must divide Q - 1 = 2^12 * 1048573)."""
    assert (Q - 1) % (2 * n) == 0
    for g in range(2, 1 << 16):
        psi = pow(g, (Q - 1) // (2 * n), Q)
        if pow(psi, n, Q) == Q - 1:
            return psi
    raise ValueError("no root")


def inverse_table(n: int, psi: int | None = None) -> np.ndarray:
    """Twiddles for the Gentleman-Sande inverse transform of the engine:
    table[k] = psi^-bitrev(k) / 2 mod Q, so that log2(n) halvings give the 1/n factor."""
    psi = primitive_root_2n(n) if psi is None else psi
    logn = n.bit_length() - 1
    psi_inv = pow(psi, Q - 2, Q)
    inv2 = (Q + 1) // 2
    t = np.zeros(n, dtype=np.uint64)
    for k in range(1, n):
        t[k] = pow(psi_inv, bitrev(k, logn), Q) * inv2 % Q
    return t


@dataclass
class EngineInputs:
    rows: int
    cols: int
    n: int
    tg: int
    ld: int
    db_col: np.ndarray           # uint8, cols * ld (column-major, ld >= rows)
    ub: np.ndarray               # int8, 4 * tg * n * cols (reference tile layout)
    prefixes: list = field(default_factory=list)  # cols / n uint8 arrays of part_len
    that: list = field(default_factory=list)      # cols / n uint64 arrays of tg * n
    inv_table: np.ndarray = None  # uint64, n
    part_len: int = 0
    row1_off: int = 0
    q1: int = 1024
    q1_bits: int = 10

    @property
    def kp(self) -> int:
        return self.tg * self.n

    @property
    def outputs(self) -> int:
        return self.cols // self.n

    @property
    def request_len(self) -> int:
        return request_len(self.rows, self.kp)

    @property
    def response_len(self) -> int:
        return self.outputs * self.part_len

    def db_record_major(self) -> np.ndarray:
        """The database as records x bytes (row k = record k)."""
        return self.db_col.reshape(self.cols, self.ld)[:, : self.rows].T


def make_inputs(rows: int, cols: int, n: int = 2048, tg: int = 4, seed: int = 0, ld: int | None = None,
                table: str = "ntt", prefix: str = "row0", db=None) -> EngineInputs:
    """Seeded setup artefacts.  db_col byte (col * ld + k) is byte (col * rows + k) of the
    stream `seed` (i.e. byte `col` of record `k`); U's tile bytes come from stream seed+1,
    `that` from seed+2 (full 64-bit values), random inverse tables from seed+3, prefixes
    from seed+4.  table="ntt" gives the real twiddles of a 2n-th root of unity,
    table="random" arbitrary u64 values (only their low 32 bits matter).  prefix="row0"
    fills the row-0 bits and leaves row 1 zero as a server does; prefix="random" fills
    every byte, which exercises the OR of row-1 fields onto the prefix."""
    ld = rows if ld is None else ld
    kp = tg * n
    if db is False:
        db_col = None  # the caller generates the database elsewhere (e.g. on the device)
    elif db is None:
        body = synthetic.uniform_bytes(seed, rows * cols).reshape(cols, rows)
        if ld == rows:
            db_col = body.reshape(-1)
        else:
            db_col = np.zeros((cols, ld), dtype=np.uint8)
            db_col[:, :rows] = body
            db_col = db_col.reshape(-1)
    else:
        db_col = db
    ub = synthetic.uniform_bytes(seed + 1, 4 * kp * cols).view(np.int8)
    outputs = cols // n
    th = synthetic.stream(seed + 2, outputs * kp).reshape(outputs, kp)
    if table == "ntt" and (Q - 1) % (2 * n) == 0:
        inv = inverse_table(n)
    else:
        inv = synthetic.stream(seed + 3, n)
    plen = part_len_for(n)
    row1_off = 18 * n
    pre = synthetic.uniform_bytes(seed + 4, outputs * plen).reshape(outputs, plen).copy()
    if prefix == "row0":
        pre[:, row1_off // 8:] = 0
        if row1_off % 8:
            pre[:, row1_off // 8] &= (1 << (row1_off % 8)) - 1
    return EngineInputs(rows=rows, cols=cols, n=n, tg=tg, ld=ld, db_col=db_col, ub=ub,
                        prefixes=[pre[o] for o in range(outputs)], that=[th[o] for o in range(outputs)],
                        inv_table=inv, part_len=plen, row1_off=row1_off)


def encode_request(y: np.ndarray, z: np.ndarray, q: np.ndarray) -> np.ndarray:
    """Wire frame of src/service.rs:178-198 (as a uint8 array)."""
    kp = len(y)
    assert len(z) == kp
    out = np.empty(9 + 4 * (2 * kp + len(q)), dtype=np.uint8)
    out[0] = 0
    out[1:5] = np.frombuffer(np.uint32(kp).tobytes(), dtype=np.uint8)
    out[5:9] = out[1:5]
    body = np.concatenate([np.asarray(y, dtype="<u4"), np.asarray(z, dtype="<u4"), np.asarray(q, dtype="<u4")])
    out[9:] = body.view(np.uint8)
    return out


def make_requests(B: int, rows: int, kp: int, seed: int) -> list[np.ndarray]:
    """B valid frames: y, z uniform in [0, Q) and q uniform u32, from stream seed*1000003 + c."""
    reqs = []
    for c in range(B):
        s = seed * 1000003 + c
        y = synthetic.uniform_below_q(s * 3 + 1, kp)
        z = synthetic.uniform_below_q(s * 3 + 2, kp)
        q = synthetic.uniform_u32(s * 3 + 3, rows)
        reqs.append(encode_request(y, z, q))
    return reqs


def fill_requests(arena: np.ndarray, B: int, rows: int, kp: int, seed: int) -> None:
    """Writes make_requests(B, ...) frames into rows of a (B, >= request_len) uint8 arena."""
    L = request_len(rows, kp)
    for c, f in enumerate(make_requests(B, rows, kp, seed)):
        arena[c, :L] = f


def frame_status(frame, rows: int, kp: int) -> int:
    """The status the engine gives a frame (src/amx_kernels.cpp:571-581): 1 length, 2 flag or
    key-length field, 3 a key word >= Q, 0 valid."""
    if frame is None:
        return 1
    f = bytes(frame)
    if len(f) != request_len(rows, kp):
        return 1
    if f[0] != 0 or int.from_bytes(f[1:5], "little") != kp or int.from_bytes(f[5:9], "little") != kp:
        return 2
    keys = np.frombuffer(f, dtype="<u4", count=2 * kp, offset=9)
    return 3 if (keys >= Q).any() else 0


def corrupt(frame, rng: np.random.Generator, kp: int) -> bytes:
    """One corruption of a valid frame, as a client bug or an attacker would make it: a
    truncated or extended frame, a flipped header byte, or a key word set to Q or above."""
    f = bytearray(bytes(frame))
    kind = int(rng.integers(4))
    if kind == 0:
        del f[int(rng.integers(1, len(f))):]
    elif kind == 1:
        f += bytes(int(rng.integers(1, 64)))
    elif kind == 2:
        f[int(rng.integers(9))] ^= 1 << int(rng.integers(8))
    else:
        w = 9 + 4 * int(rng.integers(2 * kp))
        f[w:w + 4] = int(Q + int(rng.integers(12287))).to_bytes(4, "little")
    return bytes(f)


def _stream_device(seeds, count: int, device):
    """synthetic.stream(seed, count) for each seed of `seeds` (a list), as a (len(seeds), count)
    torch int64 tensor of the u64 bit patterns."""
    import torch

    s64 = lambda v: v - (1 << 64) if v >= (1 << 63) else v  # noqa: E731
    sd = torch.tensor([s64(x & ((1 << 64) - 1)) for x in seeds], dtype=torch.int64, device=device)
    z = torch.arange(1, count + 1, dtype=torch.int64, device=device).unsqueeze(0).repeat(len(seeds), 1)
    z.mul_(s64(0x9E3779B97F4A7C15)).add_(sd.unsqueeze(1))
    z.bitwise_xor_((z >> 30) & ((1 << 34) - 1))
    z.mul_(s64(0xBF58476D1CE4E5B9))
    z.bitwise_xor_((z >> 27) & ((1 << 37) - 1))
    z.mul_(s64(0x94D049BB133111EB))
    z.bitwise_xor_((z >> 31) & ((1 << 33) - 1))
    return z


def make_requests_device(B: int, rows: int, kp: int, seed: int, stride: int, offset: int = 7, device="cuda",
                         group: int = 64):
    """make_requests(B, rows, kp, seed) generated on the GPU straight into a (B, stride) uint8
    arena, frame c at byte `offset` of row c (the layout of the device entry point); `group`
    requests are generated per pass."""
    import torch

    L = request_len(rows, kp)
    assert offset + L <= stride
    arena = torch.zeros((B, stride), dtype=torch.uint8, device=device)
    two64_mod_q = (1 << 64) % Q
    hdr = torch.tensor(list(bytes([0]) + kp.to_bytes(4, "little") * 2), dtype=torch.uint8, device=device)
    arena[:, offset:offset + 9] = hdr
    for c0 in range(0, B, group):
        cs = range(c0, min(B, c0 + group))
        ss = [seed * 1000003 + c for c in cs]
        words = torch.empty((len(cs), 2 * kp + rows), dtype=torch.int64, device=device)
        for lo, k in ((0, 1), (kp, 2)):
            z = _stream_device([s * 3 + k for s in ss], kp, device)
            r = torch.remainder(z, Q)  # z as a signed value; fix up the u64 meaning below
            words[:, lo:lo + kp] = torch.where(z < 0, torch.remainder(r + two64_mod_q, Q), r)
            del z, r
        words[:, 2 * kp:] = _stream_device([s * 3 + 3 for s in ss], rows, device) & 0xFFFFFFFF
        arena[c0:c0 + len(cs), offset + 9:offset + L] = words.to(torch.int32).view(torch.uint8)
        del words
    return arena
