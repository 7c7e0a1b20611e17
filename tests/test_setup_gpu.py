"""Setup work on the device: the packing-matrix build (alto_pack_build) against the oracle's
restatement of sw_pack_build (src/amx_kernels.cpp:311-341; pinned against the reference's own
build in test_oracle.py) on small shapes, and against the library's host sw_pack_build at the
reference's full size (n = 2048, tg = 4, rho = 16 outputs)."""
from __future__ import annotations

import ctypes as C
import sys
import time

import numpy as np
import pytest

from altopir_public_amd import _lib, synthetic, workload
from altopir_public_amd.engine import build_packing_matrix
import oracle_ctypes as oc

pytestmark = pytest.mark.gpu


def pack_inputs(n, tg, rho, seed):
    num_iter = n // 2 - 1
    logn = n.bit_length() - 1
    br = np.array([workload.bitrev(j, logn) for j in range(n)], dtype=np.int64)
    # automorphism tables: row t (g = 2t + 1) maps slot j to the slot of psi^(g (2 br(j) + 1))
    g = np.arange(1, 2 * n, 2, dtype=np.int64)[:, None]
    img = (g * (2 * br[None, :] + 1)) % (2 * n)
    inv = np.empty(n, dtype=np.int64)
    inv[br] = np.arange(n)
    tables = inv[(img - 1) // 2].astype(np.uint32).reshape(-1)
    gen_pows = np.array([pow(5, i, 2 * n) for i in range(n)], dtype=np.uint32)
    bt = [synthetic.stream(seed + o, num_iter * tg * n) for o in range(rho)]
    bb = [synthetic.stream(seed + 100 + o, num_iter * tg * n) for o in range(rho)]
    return num_iter, tables, gen_pows, bt, bb


@pytest.mark.parametrize("n,tg,rho", [(16, 8, 2), (64, 2, 1), (32, 4, 4), (128, 4, 1)])
def test_device_pack_build_matches_the_oracle(hip_device, n, tg, rho):
    num_iter, tables, gen_pows, bt, bb = pack_inputs(n, tg, rho, seed=40 + n)
    # the oracle's check of the tables it is fed: the closed form of src/packing.rs:1082-1110
    # (tested against the reference in test_oracle.py) gives the same rows
    pt = (C.c_void_p * rho)(*[x.ctypes.data for x in bt])
    pb = (C.c_void_p * rho)(*[x.ctypes.data for x in bb])
    size = 4 * tg * n * rho * n
    want = np.zeros(size, np.int8)
    oc.oracle().orc_pack_build(want.ctypes.data, C.addressof(pt), C.addressof(pb), tables.ctypes.data,
                               gen_pows.ctypes.data, num_iter, tg, n, rho)
    got = np.full(size, 0x55, np.int8)
    build_packing_matrix(got, bt, bb, tables, gen_pows, num_iter, tg, n, workload.Q, 1, device=True)
    assert np.array_equal(got, want)


def test_device_pack_build_rejects_unsupported_shapes(hip_device):
    lib = _lib.load()
    num_iter, tables, gen_pows, bt, bb = pack_inputs(16, 2, 1, seed=7)  # rho n = 16: not a whole column pair
    pt = (C.c_void_p * 1)(bt[0].ctypes.data)
    pb = (C.c_void_p * 1)(bb[0].ctypes.data)
    ub = np.zeros(4 * 2 * 16 * 16, np.int8)
    assert lib.alto_pack_build(ub.ctypes.data, C.addressof(pt), C.addressof(pb), tables.ctypes.data,
                               gen_pows.ctypes.data, num_iter, 2, 16, 1, workload.Q, 1) == -1
    assert "unsupported shape" in _lib.last_error()


def test_device_pack_build_at_full_size(hip_device):
    # the reference's parameters: n = 2048, tg = 4, 16 outputs (32 KB records): a 1 GiB matrix
    # from 2 x 1.07 GB of bold_t / bold_t_bar
    n, tg, rho = 2048, 4, 16
    num_iter, tables, gen_pows, bt, bb = pack_inputs(n, tg, rho, seed=900)
    size = 4 * tg * n * rho * n
    host = np.empty(size, np.int8)
    t0 = time.perf_counter()
    build_packing_matrix(host, bt, bb, tables, gen_pows, num_iter, tg, n, workload.Q, 16)
    t_host = time.perf_counter() - t0
    dev = np.empty(size, np.int8)
    build_packing_matrix(dev, bt, bb, tables, gen_pows, num_iter, tg, n, workload.Q, 16, device=True)  # warm-up
    t0 = time.perf_counter()
    build_packing_matrix(dev, bt, bb, tables, gen_pows, num_iter, tg, n, workload.Q, 16, device=True)
    t_dev = time.perf_counter() - t0
    print(f"[setup] packing matrix n={n} tg={tg} rho={rho}: host sw_pack_build (16 threads) {t_host:.3f} s, "
          f"alto_pack_build {t_dev:.3f} s (host inputs in, host matrix out)", file=sys.stderr, flush=True)
    assert np.array_equal(dev, host)


@pytest.mark.parametrize("rows,cols,n,check", [(1024, 96, 32, None), (4096, 512, 256, None), (8192, 4096, 2048, 64),
                                               (131072, 256, 256, None)])
def test_device_hint_matches_the_oracle(hip_device, rows, cols, n, check):
    # the offline hint (src/server.rs:656-709, 738-768) from the engine's resident database:
    # 8 to 512 A tiles of query coefficients (8-row and 16-row contraction shapes); 131,072
    # records = 2,048 k-steps, so the record range is split and the exact sums are added
    from altopir_public_amd.engine import Engine

    Q = 4294955009
    inp = workload.make_inputs(rows, cols, n=n, tg=4, seed=rows + cols + n)
    eng = Engine.from_inputs(inp)
    rng = np.random.default_rng(rows + n)
    a = rng.integers(0, Q, size=rows, dtype=np.uint64)
    a[:6] = [0, Q - 1, 0x7F7F7F7F, 0x7F7F7F80, 1, Q // 2]  # the centring boundaries and extremes
    t0 = time.perf_counter()
    got = eng.hint0(a.reshape(rows // n, n))
    dt = time.perf_counter() - t0
    eng.close()
    print(f"[setup] hint {rows} records x {cols} columns, n = {n}: {dt * 1e3:.1f} ms", file=sys.stderr, flush=True)
    assert got.shape == (n, cols) and (got < Q).all()
    if check is None:
        assert np.array_equal(got, oc.oracle_hint0(a, inp.db_col, inp.ld, n))
    else:  # the first and the last `check` columns
        for c0 in (0, cols - check):
            want = oc.oracle_hint0(a, inp.db_col, inp.ld, n, c0, check)
            assert np.array_equal(got[:, c0:c0 + check], want), f"columns {c0}..{c0 + check}"


def test_device_hint_rejects_a_record_count_that_is_not_whole_polynomials(hip_device):
    from altopir_public_amd.engine import Engine

    inp = workload.make_inputs(2048 + 64, 2048, n=2048, tg=4, seed=3)
    eng = Engine.from_inputs(inp)
    with pytest.raises(RuntimeError, match="multiple of n"):
        eng.hint0(np.zeros((1, 2048), np.uint64))
    eng.close()


@pytest.mark.parametrize("on_device", [False, True])
def test_hint_without_an_engine_matches_the_oracle(hip_device, on_device):
    # alto_hint0 tiles a column-major database (host or device memory, padded columns) itself
    import torch

    from altopir_public_amd.engine import hint0

    Q = 4294955009
    rows, cols, n, ld = 4096, 512, 256, 4096 + 64
    rng = np.random.default_rng(70 + on_device)
    db_col = rng.integers(0, 256, size=ld * cols, dtype=np.uint8)
    a = rng.integers(0, Q, size=rows, dtype=np.uint64)
    if on_device:
        d = torch.from_numpy(db_col).cuda()
        torch.cuda.synchronize()
        got = hint0(d.data_ptr(), ld, rows, cols, n, a.reshape(-1, n), on_device=True)
        del d
    else:
        got = hint0(db_col, ld, rows, cols, n, a.reshape(-1, n))
    assert np.array_equal(got, oc.oracle_hint0(a, db_col, ld, n))
