"""Host mirror of the reference's engine wrapper (src/engine.rs) over the HIP library.

Same names, argument meaning and error behaviour as the Rust module:
`Method`, `amx_available`/`vnni_available`/`init_amx`, `MappedBytes`, `TiledDb`,
`build_packing_matrix` and `Engine` with `new` (the constructor), `request_len`,
`response_len` and `answer`.  Where the Rust code asserts, this module raises
AssertionError; a negative engine return raises RuntimeError("engine error N")
(src/engine.rs:186).  `Method.HIP` and `Engine.answer_device` are additions.
"""
from __future__ import annotations

import ctypes as C
import enum
from typing import Sequence

import numpy as np

from . import _lib

Q = 4294955009


class Method(enum.IntEnum):
    """Method codes of src/engine.rs:22-50; HIP (4) is this engine's own code.  All three
    run the HIP engine."""
    AMX = 0
    VNNI = 2
    HIP = 4

    @staticmethod
    def parse(s: str) -> "Method | None":
        return {"amx": Method.AMX, "vnni": Method.VNNI, "hip": Method.HIP}.get(s)

    def name_str(self) -> str:
        return {Method.AMX: "amx", Method.VNNI: "vnni", Method.HIP: "hip"}[self]


def cpu_flag(flag: str) -> bool:
    """src/engine.rs:52-56."""
    try:
        with open("/proc/cpuinfo") as f:
            return any(flag in line.split() for line in f if line.startswith("flags"))
    except OSError:
        return False


def amx_available() -> bool:
    return cpu_flag("amx_int8") and cpu_flag("amx_tile")


def vnni_available() -> bool:
    return cpu_flag("avx512_vnni")


def hip_available() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:  # pragma: no cover - torch is part of the image
        return False


def init_amx() -> None:
    """src/engine.rs:66-69 (the HIP library grants it unconditionally)."""
    r = _lib.load().sw_amx_init()
    assert r == 0, "the kernel did not grant AMX tile state (arch_prctl ARCH_REQ_XCOMP_PERM)"


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


class MappedBytes:
    """src/engine.rs:71-100: a zeroed, 2 MiB-aligned host mapping."""

    def __init__(self, ptr: int, length: int):
        self.ptr, self.len = ptr, length

    @classmethod
    def zeroed(cls, length: int) -> "MappedBytes":
        assert length > 0
        p = _lib.load().sw_alloc_zeroed(length)
        assert p, "out of memory"
        return cls(p, length)

    def as_i8(self) -> np.ndarray:
        buf = (C.c_int8 * self.len).from_address(self.ptr)
        return np.frombuffer(buf, dtype=np.int8)

    def __del__(self):
        if getattr(self, "ptr", None):
            _lib.load().sw_free(self.ptr)
            self.ptr = None


class TiledDb:
    """src/engine.rs:102-125: the reference's tile-order copy of the database."""

    def __init__(self, db: np.ndarray, ld: int, rows: int, cols: int, threads: int):
        assert rows % 64 == 0 and cols % 32 == 0 and len(db) >= (cols - 1) * ld + rows
        lib = _lib.load()
        self.ptr = lib.sw_alloc(rows * cols)
        assert self.ptr, "out of memory"
        self.rows, self.cols = rows, cols
        self._db = db  # keeps the source alive for as long as the engine may borrow it
        lib.sw_tile_db(self.ptr, _ptr(db), ld, rows, cols, threads)

    @classmethod
    def new(cls, db, ld, rows, cols, threads) -> "TiledDb":
        return cls(db, ld, rows, cols, threads)

    def as_u8(self) -> np.ndarray:
        return np.frombuffer((C.c_uint8 * (self.rows * self.cols)).from_address(self.ptr), dtype=np.uint8)

    def __del__(self):
        if getattr(self, "ptr", None):
            _lib.load().sw_free(self.ptr)
            self.ptr = None


def hint0(db_col, ld: int, rows: int, cols: int, n: int, a: np.ndarray, on_device: bool = False) -> np.ndarray:
    """The offline hint (src/server.rs:656-709, 738-768) of a column-major database without an
    engine (alto_hint0): db_col is a host uint8 array (column col at col * ld) or, with
    on_device, a device address; a is (rows / n, n) uint64.  Returns (n, cols) uint64 mod Q."""
    lib = _lib.load()
    a = np.ascontiguousarray(a, dtype=np.uint64)
    out = np.zeros((n, cols), np.uint64)
    ptr = int(db_col) if on_device else _ptr(db_col)
    if lib.alto_hint0(ptr, int(on_device), ld, rows, cols, n, a.ctypes.data, a.size // n, out.ctypes.data) != 0:
        raise RuntimeError(f"alto_hint0 failed ({_lib.last_error()})")
    return out


def build_packing_matrix(ub: np.ndarray, bold_t: Sequence[np.ndarray], bold_t_bar: Sequence[np.ndarray],
                         tables: np.ndarray, gen_pows: np.ndarray, num_iter: int, tg: int, n: int, qmod: int,
                         threads: int, device: bool = False) -> None:
    """src/engine.rs:127-138.  device=True builds it on the GPU (alto_pack_build, the same bytes)."""
    rho = len(bold_t)
    assert len(ub) == 4 * tg * n * rho * n and len(bold_t_bar) == rho
    assert len(tables) == n * n and len(gen_pows) >= num_iter
    for s in list(bold_t) + list(bold_t_bar):
        assert len(s) >= num_iter * tg * n and s.dtype == np.uint64
    p = (C.c_void_p * rho)(*[_ptr(s) for s in bold_t])
    pb = (C.c_void_p * rho)(*[_ptr(s) for s in bold_t_bar])
    tables = np.ascontiguousarray(tables, dtype=np.uint32)
    gen_pows = np.ascontiguousarray(gen_pows, dtype=np.uint32)
    lib = _lib.load()
    if device:
        r = lib.alto_pack_build(_ptr(ub), C.addressof(p), C.addressof(pb), _ptr(tables), _ptr(gen_pows), num_iter, tg,
                                n, rho, qmod, threads)
        if r != 0:
            raise RuntimeError(f"alto_pack_build failed ({_lib.last_error()})")
        return
    lib.sw_pack_build(_ptr(ub), C.addressof(p), C.addressof(pb), _ptr(tables), _ptr(gen_pows), num_iter, tg, n, rho,
                      qmod, threads)


class Engine:
    """src/engine.rs:140-195.  Unlike the reference, the engine copies its setup buffers
    into HBM, so the host arrays may be dropped after construction."""

    def __init__(self, db_col: np.ndarray, tiled: TiledDb, ld: int, ub: np.ndarray, n: int, tg: int,
                 prefixes: Sequence[np.ndarray], row1_off: int, q1: int, q1_bits: int,
                 that: Sequence[np.ndarray], inv_table: np.ndarray, q: int, threads: int):
        rows, cols = tiled.rows, tiled.cols
        assert len(db_col) >= (cols - 1) * ld + rows and len(ub) == 4 * tg * n * cols and len(prefixes) == cols // n
        assert len(that) == cols // n and all(len(t) >= tg * n for t in that) and len(inv_table) >= n
        part_len = len(prefixes[0])
        assert all(len(p) == part_len for p in prefixes)
        self._lib = _lib.load()
        prefixes = [np.ascontiguousarray(p, dtype=np.uint8) for p in prefixes]
        that = [np.ascontiguousarray(t, dtype=np.uint64) for t in that]
        inv_table = np.ascontiguousarray(inv_table, dtype=np.uint64)
        pp = (C.c_void_p * len(prefixes))(*[_ptr(p) for p in prefixes])
        tp = (C.c_void_p * len(that))(*[_ptr(t) for t in that])
        self.ptr = self._lib.alto_engine_new(_ptr(db_col), tiled.ptr, ld, rows, cols, _ptr(ub), n, tg,
                                             C.addressof(pp), part_len, row1_off, q1, q1_bits, C.addressof(tp),
                                             _ptr(inv_table), q, threads)
        assert self.ptr, f"unsupported engine configuration ({_lib.last_error()})"
        self.rows, self.cols = rows, cols

    @classmethod
    def new(cls, *args, **kw) -> "Engine":
        return cls(*args, **kw)

    @classmethod
    def from_inputs(cls, inp, threads: int = 8) -> "Engine":
        """Convenience constructor from workload.EngineInputs."""
        tiled = TiledDb(inp.db_col, inp.ld, inp.rows, inp.cols, threads)
        return cls(inp.db_col, tiled, inp.ld, inp.ub, inp.n, inp.tg, inp.prefixes, inp.row1_off, inp.q1,
                   inp.q1_bits, inp.that, inp.inv_table, Q, threads)

    @classmethod
    def from_records(cls, records_ptr: int, on_device: bool, ld: int, inp, threads: int = 8) -> "Engine":
        """Engine over a database given as records (row-major, record k at records_ptr + k * ld,
        host or device memory), transposed into the engine's layout on the device; the other
        setup inputs come from workload.EngineInputs."""
        self = cls.__new__(cls)
        self._lib = _lib.load()
        prefixes = [np.ascontiguousarray(p, dtype=np.uint8) for p in inp.prefixes]
        that = [np.ascontiguousarray(t, dtype=np.uint64) for t in inp.that]
        inv_table = np.ascontiguousarray(inp.inv_table, dtype=np.uint64)
        pp = (C.c_void_p * len(prefixes))(*[_ptr(p) for p in prefixes])
        tp = (C.c_void_p * len(that))(*[_ptr(t) for t in that])
        self.ptr = self._lib.alto_engine_new_records(records_ptr, int(on_device), ld, inp.rows, inp.cols, _ptr(inp.ub),
                                                     inp.n, inp.tg, C.addressof(pp), inp.part_len, inp.row1_off,
                                                     inp.q1, inp.q1_bits, C.addressof(tp), _ptr(inv_table), Q, threads)
        assert self.ptr, f"unsupported engine configuration ({_lib.last_error()})"
        self.rows, self.cols = inp.rows, inp.cols
        return self

    @classmethod
    def from_device_db(cls, db_ptr: int, inp, threads: int = 8) -> "Engine":
        """Engine over a column-major database already in device memory (db_ptr, ld =
        inp.ld); the other setup inputs come from workload.EngineInputs."""
        self = cls.__new__(cls)
        self._lib = _lib.load()
        prefixes = [np.ascontiguousarray(p, dtype=np.uint8) for p in inp.prefixes]
        that = [np.ascontiguousarray(t, dtype=np.uint64) for t in inp.that]
        inv_table = np.ascontiguousarray(inp.inv_table, dtype=np.uint64)
        pp = (C.c_void_p * len(prefixes))(*[_ptr(p) for p in prefixes])
        tp = (C.c_void_p * len(that))(*[_ptr(t) for t in that])
        self.ptr = self._lib.alto_engine_new_device(db_ptr, inp.ld, inp.rows, inp.cols, _ptr(inp.ub), inp.n, inp.tg,
                                                    C.addressof(pp), inp.part_len, inp.row1_off, inp.q1, inp.q1_bits,
                                                    C.addressof(tp), _ptr(inv_table), Q, threads)
        assert self.ptr, f"unsupported engine configuration ({_lib.last_error()})"
        self.rows, self.cols = inp.rows, inp.cols
        return self

    def request_len(self) -> int:
        return self._lib.alto_request_len(self.ptr)

    def response_len(self) -> int:
        return self._lib.alto_response_len(self.ptr)

    def device_bytes(self) -> int:
        return self._lib.alto_engine_device_bytes(self.ptr)

    def answer(self, req: Sequence, out, status: np.ndarray, method: Method = Method.HIP):
        """src/engine.rs:174-189.  req: frames (bytes or uint8 arrays); out: a list of
        writable uint8 arrays of response_len bytes, or a (B, response_len) uint8 array;
        status: int32 array of B.  Returns (answered, stage_ms[4])."""
        b = len(req)
        rl = self.response_len()
        if isinstance(out, np.ndarray) and out.ndim == 2:
            assert out.shape == (b, rl) and out.dtype == np.uint8 and out.flags.c_contiguous
            optrs = [out.ctypes.data + i * rl for i in range(b)]
        else:
            assert len(out) == b and all(len(o) == rl and o.dtype == np.uint8 for o in out)
            optrs = [_ptr(o) for o in out]
        assert len(status) == b and status.dtype == np.int32 and status.flags.c_contiguous
        keep = []
        rp = (C.c_void_p * max(b, 1))()
        lens = (C.c_size_t * max(b, 1))()
        for i, r in enumerate(req):
            if r is None:
                rp[i] = None
                lens[i] = 0
                continue
            a = np.frombuffer(r, dtype=np.uint8) if isinstance(r, (bytes, bytearray)) else np.ascontiguousarray(r)
            keep.append(a)
            rp[i] = a.ctypes.data if len(a) else None
            lens[i] = len(a)
        op = (C.c_void_p * max(b, 1))(*optrs)
        stage = (C.c_double * 4)()
        r = self._lib.alto_answer(self.ptr, C.addressof(rp), C.addressof(lens), b, C.addressof(op),
                                  status.ctypes.data, int(method), C.addressof(stage))
        if r < 0:
            raise RuntimeError(f"engine error {r} ({_lib.last_error()})")
        return r, list(stage)

    def hint0(self, a: np.ndarray) -> np.ndarray:
        """The offline hint from the resident database (alto_engine_hint0; src/server.rs:656-709,
        738-768): a is (rows / n, n) uint64, the query polynomials in coefficient form; returns
        (n, cols) uint64 values mod Q."""
        a = np.ascontiguousarray(a, dtype=np.uint64)
        n = a.shape[1]
        out = np.zeros((n, self.cols), np.uint64)
        if self._lib.alto_engine_hint0(self.ptr, a.ctypes.data, a.shape[0], out.ctypes.data) != 0:
            raise RuntimeError(f"alto_engine_hint0 failed ({_lib.last_error()})")
        return out

    def answer_device(self, slots_ptr: int, slot_stride: int, b: int, out_ptr: int, out_stride: int,
                      status_ptr: int, stream: int = 0) -> int:
        """Device-resident batch (alto_answer_device): all pointers are device addresses."""
        return self._lib.alto_answer_device(self.ptr, slots_ptr, slot_stride, b, out_ptr, out_stride, status_ptr,
                                            stream)

    def set_graphs(self, on: bool) -> None:
        """Graph capture/replay of device-resident batches (alto_set_graphs)."""
        self._lib.alto_set_graphs(self.ptr, 1 if on else 0)
        self._graphs = bool(on)

    @property
    def graphs(self) -> bool:
        """Whether device-resident batches are captured and replayed (on for a new engine)."""
        return getattr(self, "_graphs", True)

    def set_graph_timing(self, on: bool) -> None:
        """Graph replays record the stage events too (one captured graph per event set)."""
        self._lib.alto_set_graph_timing(self.ptr, int(on))

    def set_fusion(self, on) -> None:
        """One launch for scan and packing where the chosen shape pair allows: True / 1 for
        every such pair, False / 0 never, -1 the built-in choice (batches of 1-4 queries on
        databases of at most 4 KP records); 2: scan, packing and tail in one launch (the
        step kernel) for every batch size whose shapes have a step pair, else as 1."""
        if int(on) not in (-1, 0, 1, 2):
            raise ValueError(on)
        self._lib.alto_set_fusion(self.ptr, int(on))

    def fused(self, b: int) -> bool:
        """Whether a batch of b queries runs its two contractions in one launch."""
        return self._lib.alto_fused(self.ptr, b) >= 1

    def step(self, b: int) -> bool:
        """Whether a batch of b queries runs scan, packing and tail in one launch."""
        return self._lib.alto_fused(self.ptr, b) == 2

    def last_timings(self) -> list:
        ms = (C.c_double * 6)()
        k = self._lib.alto_last_timings(self.ptr, C.addressof(ms), 6)
        return list(ms)[: max(k, 0)]

    def stage_totals(self, reset: bool = False) -> tuple[int, list]:
        """(batches, summed stage ms) over the timed batches since the last reset."""
        ms = (C.c_double * 6)()
        k = self._lib.alto_stage_totals(self.ptr, C.addressof(ms), 6, int(reset))
        if k < 0:
            raise RuntimeError(f"alto_stage_totals failed ({_lib.last_error()})")
        return k, list(ms)

    def close(self):
        if getattr(self, "ptr", None):
            self._lib.alto_engine_free(self.ptr)
            self.ptr = None

    def __del__(self):
        self.close()
