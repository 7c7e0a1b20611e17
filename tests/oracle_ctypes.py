"""Test-side bindings of the checkers under oracle/.

oracle/liboracle.so is this project's C restatement of the answer engine;
oracle/_ref/libaltopir_ref.so is the reference's own src/amx_kernels.cpp compiled
unchanged (oracle/Makefile, target `ref`), present wherever it was built.  Both are test
infrastructure: only tests, the smoke check and the benchmark's CPU baseline load them.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent.parent
ORACLE_DIR = ROOT / "oracle"
ORACLE_LIB = ORACLE_DIR / "liboracle.so"
REF_LIB = ORACLE_DIR / "_ref" / "libaltopir_ref.so"
Q = 4294955009

P = C.c_void_p
SZ = C.c_size_t


class OrcEngine(C.Structure):
    _fields_ = [
        ("db_col", P), ("ld", SZ), ("rows", SZ), ("cols", SZ), ("ub", P), ("n", SZ), ("tg", SZ),
        ("prefix", P), ("part_len", SZ), ("row1_off", SZ), ("q1", C.c_uint64), ("q1_bits", C.c_int),
        ("that", P), ("inv_table", P),
    ]


_oracle = None
_ref = None


def oracle() -> C.CDLL:
    global _oracle
    if _oracle is None:
        if not ORACLE_LIB.exists():
            subprocess.run(["make", "-s", "-C", str(ORACLE_DIR), "liboracle.so"], check=True)
        lib = C.CDLL(str(ORACLE_LIB))
        lib.orc_request_len.restype = SZ
        lib.orc_request_len.argtypes = [SZ, SZ]
        lib.orc_validate.restype = C.c_int
        lib.orc_validate.argtypes = [P, SZ, SZ, SZ]
        lib.orc_unpack_u.restype = C.c_uint32
        lib.orc_unpack_u.argtypes = [P, SZ, SZ, SZ, SZ]
        lib.orc_scan.argtypes = [P, P, SZ, SZ, SZ, P, SZ]
        lib.orc_scan_acc.argtypes = [P, P, SZ, SZ, SZ, P, SZ]
        lib.orc_pack.argtypes = [P, P, P, SZ, SZ, SZ]
        lib.orc_mulmod.restype = C.c_uint64
        lib.orc_mulmod.argtypes = [C.c_uint64, C.c_uint64]
        lib.orc_intt.restype = C.c_int
        lib.orc_intt.argtypes = [P, P, SZ]
        lib.orc_tail_pointwise.argtypes = [P, P, P, P, SZ, SZ]
        lib.orc_rescale_q1.restype = C.c_uint64
        lib.orc_rescale_q1.argtypes = [C.c_uint64, C.c_uint64]
        lib.orc_rescale_i64.restype = C.c_uint64
        lib.orc_rescale_i64.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64]
        lib.orc_write_bits.argtypes = [P, SZ, C.c_uint64, SZ, SZ]
        lib.orc_switch_row1.restype = C.c_int
        lib.orc_switch_row1.argtypes = [P, SZ, SZ, P, P, SZ, C.c_uint64, C.c_int]
        lib.orc_answer.restype = C.c_int
        lib.orc_answer.argtypes = [C.POINTER(OrcEngine), P, P, SZ, P, P]
        lib.orc_pack_build.argtypes = [P, P, P, P, P, SZ, SZ, SZ, SZ]
        lib.orc_tile_db.argtypes = [P, P, SZ, SZ, SZ]
        lib.orc_pack_online.argtypes = [P, P, P, P, P, P, SZ, SZ, SZ, SZ]
        lib.orc_hint0.argtypes = [P, SZ, P, SZ, SZ, P, SZ, SZ, SZ]
        lib.orc_ntt_naive.argtypes = [P, P, SZ, C.c_uint64]
        lib.orc_intt_naive.argtypes = [P, P, SZ, C.c_uint64]
        lib.orc_powmod.restype = C.c_uint64
        lib.orc_powmod.argtypes = [C.c_uint64, C.c_uint64]
        lib.orc_ntt.argtypes = [P, SZ, C.c_uint64]
        lib.orc_auto_slot.restype = SZ
        lib.orc_auto_slot.argtypes = [SZ, SZ, SZ]
        lib.orc_fast_batched_dot_product.argtypes = [P, SZ, P, SZ, SZ, P, C.c_int, SZ, SZ, SZ, C.c_uint64,
                                                     C.c_uint64]
        lib.orc_inspiring_offline.restype = C.c_int
        lib.orc_inspiring_offline.argtypes = [P, P, P, P, P, SZ, SZ, P, P, SZ, SZ, C.c_uint64]
        _oracle = lib
    return _oracle


def ref_available() -> bool:
    return REF_LIB.exists()


def ref() -> C.CDLL:
    """The reference engine, compiled from the reference's own source (oracle/_ref)."""
    global _ref
    if _ref is None:
        lib = C.CDLL(str(REF_LIB))
        lib.sw_amx_init.restype = C.c_int
        lib.sw_alloc.restype = P
        lib.sw_alloc.argtypes = [SZ]
        lib.sw_free.argtypes = [P]
        lib.sw_tile_db.argtypes = [P, P, SZ, SZ, SZ, C.c_int]
        lib.sw_pack_build.argtypes = [P, P, P, P, P, SZ, SZ, SZ, SZ, C.c_uint64, C.c_int]
        lib.alto_engine_new.restype = P
        lib.alto_engine_new.argtypes = [P, P, SZ, SZ, SZ, P, SZ, SZ, P, SZ, SZ, C.c_uint64, C.c_int, P, P,
                                        C.c_uint64, C.c_int]
        lib.alto_engine_free.argtypes = [P]
        lib.alto_request_len.restype = SZ
        lib.alto_request_len.argtypes = [P]
        lib.alto_response_len.restype = SZ
        lib.alto_response_len.argtypes = [P]
        lib.alto_answer.restype = C.c_int
        lib.alto_answer.argtypes = [P, P, P, SZ, P, P, C.c_int, P]
        _ref = lib
    return _ref


def cpu_has(flag: str) -> bool:
    try:
        with open("/proc/cpuinfo") as f:
            return any(flag in line.split() for line in f if line.startswith("flags"))
    except OSError:
        return False


def ref_methods() -> list[int]:
    """Methods the reference can run on this CPU (tests/engine.rs:17-27): 0 AMX, 2 VNNI."""
    m = []
    if cpu_has("amx_int8") and cpu_has("amx_tile"):
        m.append(0)
    if cpu_has("avx512_vnni"):
        m.append(2)
    return m


def as_u8(r):
    """A frame (bytes-like or uint8 array) as a contiguous uint8 array; None stays None."""
    if r is None:
        return None
    if isinstance(r, (bytes, bytearray, memoryview)):
        return np.frombuffer(r, dtype=np.uint8)
    return np.ascontiguousarray(r, dtype=np.uint8)


def _ptrs(arrs):
    return (C.c_void_p * max(len(arrs), 1))(*[a.ctypes.data if a is not None and len(a) else None for a in arrs])


class _Held:
    """Keeps ctypes arrays and numpy buffers alive next to an engine handle."""

    def __init__(self):
        self.items = []

    def add(self, *x):
        self.items.extend(x)
        return x[0] if len(x) == 1 else x


def oracle_engine(inp) -> tuple:
    """OrcEngine struct for workload.EngineInputs (plus what keeps its buffers alive)."""
    h = _Held()
    pre = [np.ascontiguousarray(p, dtype=np.uint8) for p in inp.prefixes]
    th = [np.ascontiguousarray(t, dtype=np.uint64) for t in inp.that]
    pp, tp = _ptrs(pre), _ptrs(th)
    inv = np.ascontiguousarray(inp.inv_table, dtype=np.uint64)
    h.add(pre, th, pp, tp, inv)
    e = OrcEngine(inp.db_col.ctypes.data, inp.ld, inp.rows, inp.cols, inp.ub.ctypes.data, inp.n, inp.tg,
                  C.addressof(pp), inp.part_len, inp.row1_off, inp.q1, inp.q1_bits, C.addressof(tp), inv.ctypes.data)
    return e, h


def oracle_answer(inp, reqs, out: np.ndarray, status: np.ndarray) -> int:
    """orc_answer over frames `reqs` (None = null frame) into out (B, response_len)."""
    e, h = oracle_engine(inp)
    B = len(reqs)
    frames = [as_u8(r) for r in reqs]
    rp = _ptrs(frames)
    lens = (C.c_size_t * max(B, 1))(*[0 if f is None else len(f) for f in frames])
    rl = inp.response_len
    op = (C.c_void_p * max(B, 1))(*[out.ctypes.data + i * rl for i in range(B)])
    return oracle().orc_answer(C.byref(e), C.addressof(rp), C.addressof(lens), B, C.addressof(op),
                               status.ctypes.data)


class RefEngine:
    """The compiled reference engine over workload.EngineInputs."""

    def __init__(self, inp, threads: int = 8):
        lib = ref()
        self.lib, self.inp, self.h = lib, inp, _Held()
        self.tiled = lib.sw_alloc(inp.rows * inp.cols)
        assert self.tiled
        lib.sw_tile_db(self.tiled, inp.db_col.ctypes.data, inp.ld, inp.rows, inp.cols, threads)
        if 0 in ref_methods():
            lib.sw_amx_init()
        pre = [np.ascontiguousarray(p, dtype=np.uint8) for p in inp.prefixes]
        th = [np.ascontiguousarray(t, dtype=np.uint64) for t in inp.that]
        pp, tp = _ptrs(pre), _ptrs(th)
        inv = np.ascontiguousarray(inp.inv_table, dtype=np.uint64)
        self.h.add(pre, th, pp, tp, inv)
        self.ptr = lib.alto_engine_new(inp.db_col.ctypes.data, self.tiled, inp.ld, inp.rows, inp.cols,
                                       inp.ub.ctypes.data, inp.n, inp.tg, C.addressof(pp), inp.part_len,
                                       inp.row1_off, inp.q1, inp.q1_bits, C.addressof(tp), inv.ctypes.data, Q,
                                       threads)
        assert self.ptr, "reference engine rejected the configuration"

    def answer(self, reqs, out: np.ndarray, status: np.ndarray, method: int) -> tuple[int, list]:
        B = len(reqs)
        frames = [as_u8(r) for r in reqs]
        rp = _ptrs(frames)
        lens = (C.c_size_t * max(B, 1))(*[0 if f is None else len(f) for f in frames])
        rl = self.inp.response_len
        op = (C.c_void_p * max(B, 1))(*[out.ctypes.data + i * rl for i in range(B)])
        st = (C.c_double * 4)()
        r = self.lib.alto_answer(self.ptr, C.addressof(rp), C.addressof(lens), B, C.addressof(op),
                                 status.ctypes.data, method, C.addressof(st))
        return r, list(st)

    def close(self):
        if getattr(self, "ptr", None):
            self.lib.alto_engine_free(self.ptr)
            self.lib.sw_free(self.tiled)
            self.ptr = None

    def __del__(self):
        self.close()


def oracle_hint0(a: np.ndarray, db_col: np.ndarray, ld: int, n: int, col0: int = 0, ncols: int | None = None):
    """orc_hint0: (n, ncols) uint64 hint columns col0 .. col0 + ncols of the column-major database."""
    a = np.ascontiguousarray(a, dtype=np.uint64)
    db_col = np.ascontiguousarray(db_col, dtype=np.uint8)
    npolys = a.size // n
    if ncols is None:
        ncols = db_col.size // ld - col0
    out = np.zeros((n, ncols), np.uint64)
    oracle().orc_hint0(out.ctypes.data, ncols, a.ctypes.data, npolys, n, db_col.ctypes.data, ld, col0, ncols)
    return out


def oracle_offline(hint: np.ndarray, part: int, w: np.ndarray, v: np.ndarray, n: int, tg: int, psi: int):
    """orc_inspiring_offline for output part `part` of an (n, cols) hint: (bold_t, bold_t_bar,
    bold_t_hat, a_hat) as uint64 arrays."""
    hint = np.ascontiguousarray(hint, dtype=np.uint64)
    w = np.ascontiguousarray(w, dtype=np.uint64)
    v = np.ascontiguousarray(v, dtype=np.uint64)
    m = n // 2 - 1
    bt = np.zeros(m * tg * n, np.uint64)
    bb = np.zeros(m * tg * n, np.uint64)
    bh = np.zeros(tg * n, np.uint64)
    ah = np.zeros(n, np.uint64)
    rc = oracle().orc_inspiring_offline(bt.ctypes.data, bb.ctypes.data, bh.ctypes.data, ah.ctypes.data,
                                        hint.ctypes.data, hint.shape[1], part * n, w.ctypes.data, v.ctypes.data,
                                        n, tg, psi)
    assert rc == 0
    return bt, bb, bh, ah


def oracle_pack_build(bold_t, bold_t_bar, tables, gen_pows, num_iter: int, tg: int, n: int) -> np.ndarray:
    """orc_pack_build: the packing matrix (int8, 4 tg n rho n) of rho parts."""
    rho = len(bold_t)
    ub = np.zeros(4 * tg * n * rho * n, np.int8)
    p, pb = _ptrs(list(bold_t)), _ptrs(list(bold_t_bar))
    tables = np.ascontiguousarray(tables, dtype=np.uint32)
    gen_pows = np.ascontiguousarray(gen_pows, dtype=np.uint32)
    oracle().orc_pack_build(ub.ctypes.data, C.addressof(p), C.addressof(pb), tables.ctypes.data, gen_pows.ctypes.data,
                            num_iter, tg, n, rho)
    return ub
