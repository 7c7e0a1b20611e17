"""Parity at the benchmark's full database sizes and batch sizes (32 KB records):

    1 GiB  (32,768 records)     batch 128                  all 16 output parts
    8 GiB  (262,144 records)    batches 1, 37, 500, 2048   parts 0 and 7
    64 GiB (2,097,152 records)  batches 128, 1024          parts 0 and 15

plus the three resident databases of the serving workload (1, 16 and 32 GiB) under Poisson
load with corrupted frames, and whole columns of the offline hint of the 1 and 64 GiB databases.

The oracle cannot answer these batches whole in test time, but two properties of the path
make the check exact without it: a response depends only on its own request (the batch only
decides which requests share a launch), and response part o depends only on database columns
[o n, (o + 1) n), the same packing-matrix columns, that_o and prefix_o
(src/amx_kernels.cpp:602-607).  So each batch runs whole on the HIP engine, and a few chosen
requests of it (the first, the last, one in the middle, and one malformed frame) are compared,
part by part, with the oracle run on a one-part engine built from exactly those columns.  The
database and the request frames are generated in device memory (the same bytes as the host
generators); the checked columns and frames are copied back for the oracle.  The batches run
through the device entry point with the engine's defaults (captured graphs, built-in tile
shapes), as the benchmark does.  Mirrors the multi-block and chunked batches of
tests/engine.rs:196-245.
"""
import sys
import time

import numpy as np
import pytest

from altopir_public_amd import synthetic, workload
from altopir_public_amd.engine import Engine
import oracle_ctypes as oc

pytestmark = pytest.mark.gpu

COLS, N, TG = 32768, 2048, 4


def log(*a):
    print("[full-size]", *a, file=sys.stderr, flush=True)


def part_inputs(inp, o, db_cols_host):
    """EngineInputs of output part o alone (n columns) of a full engine's inputs."""
    n, kp, cols = inp.n, inp.kp, inp.cols
    KC = kp // 64
    pair = KC * 2048  # bytes of one 32-column pair of one byte plane
    cp0, cp1 = o * n // 32, (o + 1) * n // 32
    planes = [inp.ub[j * kp * cols + cp0 * pair: j * kp * cols + cp1 * pair] for j in range(4)]
    return workload.EngineInputs(rows=inp.rows, cols=n, n=n, tg=inp.tg, ld=inp.rows, db_col=db_cols_host,
                                 ub=np.concatenate(planes), prefixes=[inp.prefixes[o]], that=[inp.that[o]],
                                 inv_table=inp.inv_table, part_len=inp.part_len, row1_off=inp.row1_off)


class FullDb:
    """One full-size database generated in device memory (seed 0) and its engine.  Only one
    is alive at a time: the 64 GiB one needs most of the device."""

    current = None

    @classmethod
    def get(cls, rows):
        import torch

        if cls.current is not None and cls.current.rows == rows:
            return cls.current
        if cls.current is not None:
            cls.current.close()
            cls.current = None
        t0 = time.time()
        self = cls.__new__(cls)
        self.rows = rows
        self.inp = workload.make_inputs(rows, COLS, n=N, tg=TG, seed=0, db=False)
        self.d_db = synthetic.uniform_bytes_device(0, rows * COLS)
        torch.cuda.synchronize()
        self.eng = Engine.from_device_db(self.d_db.data_ptr(), self.inp)
        self.parts = {}
        log(f"{rows} records: database and engine ready in {time.time() - t0:.1f} s")
        cls.current = self
        return self

    def part(self, o):
        """Host copy of the database columns of output part o (column-major, ld = rows)."""
        if o not in self.parts:
            self.parts[o] = self.d_db[o * N * self.rows:(o + 1) * N * self.rows].cpu().numpy()
        return self.parts[o]

    def close(self):
        import torch

        self.eng.close()
        del self.d_db
        self.parts.clear()
        torch.cuda.empty_cache()


@pytest.fixture(scope="module", autouse=True)
def _free_full_db():
    yield
    if FullDb.current is not None:
        FullDb.current.close()
        FullDb.current = None


def run_batch(db, B, seed):
    """Runs a batch of B device-generated requests (frame 1 made malformed when B > 1) through
    alto_answer_device; returns (checked positions, their host frames, all statuses, their
    responses, the expected statuses)."""
    import torch

    inp, rows = db.inp, db.rows
    L, rl = inp.request_len, inp.response_len
    stride = (7 + L + 255) // 256 * 256
    slots = workload.make_requests_device(B, rows, inp.kp, seed=seed, stride=stride)
    want = [0] * B
    if B > 1:
        # a key word of frame 1 set to Q: status 3 (src/amx_kernels.cpp:578-580)
        w = 7 + 9 + 4 * 12345
        slots[1, w:w + 4] = torch.tensor(list(int(synthetic.Q).to_bytes(4, "little")), dtype=torch.uint8,
                                         device=slots.device)
        want[1] = 3
    # statuses start at 0: the device entry point keeps a nonzero status given on entry
    status = torch.zeros(B, dtype=torch.int32, device="cuda")
    out = torch.empty((B, rl), dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    log(f"{rows} records, batch {B}: alto_answer_device")
    stream = torch.cuda.current_stream().cuda_stream
    rc = db.eng.answer_device(slots.data_ptr(), stride, B, out.data_ptr(), rl, status.data_ptr(), stream)
    assert rc == 0, f"alto_answer_device returned {rc}"
    torch.cuda.synchronize()
    chk = sorted({0, min(1, B - 1), B // 2, B - 1})
    frames = [np.ascontiguousarray(slots[c, 7:7 + L].cpu().numpy()) for c in chk]
    outs = out[chk].cpu().numpy()
    st = status.cpu().numpy()
    del slots, out, status
    return chk, frames, st, outs, want


@pytest.mark.parametrize("rows,B,parts,seed", [
    (32768, 128, None, 2),
    (262144, 1, (0, 7), 3),
    (262144, 37, (0, 7), 3),
    (262144, 500, (0, 7), 3),
    (262144, 2048, (0, 7), 3),
    (2097152, 1, (0, 15), 4),
    (2097152, 128, (0, 15), 4),
    (2097152, 1024, (0, 15), 4),
])
def test_full_size_batch_matches_the_oracle(hip_device, rows, B, parts, seed):
    db = FullDb.get(rows)
    inp = db.inp
    chk, frames, st, outs, want = run_batch(db, B, seed)
    assert list(st) == want, f"statuses {[(i, s) for i, s in enumerate(st) if s != want[i]][:8]}"
    for o in (parts if parts is not None else range(inp.outputs)):
        sub = part_inputs(inp, o, db.part(o))
        out_o = np.zeros((len(chk), sub.response_len), np.uint8)
        st_o = np.full(len(chk), -1, np.int32)
        t0 = time.time()
        oc.oracle_answer(sub, frames, out_o, st_o)
        assert list(st_o) == [want[c] for c in chk]
        for k, c in enumerate(chk):
            if want[c] == 0:
                got = outs[k, o * inp.part_len:(o + 1) * inp.part_len]
                assert np.array_equal(got, out_o[k]), f"{rows} records, batch {B}: part {o} of request {c} differs"
        log(f"{rows} records, batch {B}: part {o} of requests {chk} byte-identical ({time.time() - t0:.1f} s oracle)")


def test_every_variant_on_long_record_ranges(hip_device):
    # 4 GiB (131,072 records, 2,048 k-steps): every variant, forced for both contractions,
    # must give the oracle's bytes with its own tile grid and record-range splits (split
    # results added with atomics) at a batch of 128
    from altopir_public_amd import _lib, autotune

    db = FullDb.get(131072)
    inp = db.inp
    lib = _lib.load()
    names = autotune.variant_names()
    B = 128
    checked = None
    for i, name in enumerate(names):
        for which in (0, 1):
            assert lib.alto_tune_set(db.eng.ptr, which, B, i) == 0
        chk, frames, st, outs, want = run_batch(db, B, seed=6)
        assert list(st) == want, name
        if checked is None:
            checked = {}
            for o in (0, 15):
                sub = part_inputs(inp, o, db.part(o))
                out_o = np.zeros((len(chk), sub.response_len), np.uint8)
                st_o = np.full(len(chk), -1, np.int32)
                oc.oracle_answer(sub, frames, out_o, st_o)
                checked[o] = out_o
        for o, out_o in checked.items():
            for k, c in enumerate(chk):
                if want[c] == 0:
                    got = outs[k, o * inp.part_len:(o + 1) * inp.part_len]
                    assert np.array_equal(got, out_o[k]), f"variant {name}: part {o} of request {c} differs"
    for which in (0, 1):
        lib.alto_tune_set(db.eng.ptr, which, B, -1)
    log(f"131072 records, batch 128: {len(names)} variants byte-identical on parts 0 and 15")


def test_full_size_host_entry_point(hip_device):
    # the synchronous entry point of the reference's API (host frames, one pageable copy per
    # frame) at 1 GiB, batch 128 + a truncated frame, all parts of three requests
    db = FullDb.get(32768)
    inp = db.inp
    reqs = workload.make_requests(3, inp.rows, inp.kp, seed=12)
    frames = [reqs[i % 3] for i in range(128)] + [reqs[0][:-1]]
    B, rl = len(frames), inp.response_len
    out = np.zeros((B, rl), np.uint8)
    st = np.full(B, -1, np.int32)
    log("32768 records, batch 129: alto_answer")
    answered, _ = db.eng.answer(frames, out, st)
    assert answered == 128 and list(st) == [0] * 128 + [1]
    want = np.zeros((3, rl), np.uint8)
    for o in range(inp.outputs):
        sub = part_inputs(inp, o, db.part(o))
        out_o = np.zeros((3, sub.response_len), np.uint8)
        st_o = np.full(3, -1, np.int32)
        oc.oracle_answer(sub, reqs, out_o, st_o)
        want[:, o * inp.part_len:(o + 1) * inp.part_len] = out_o
    for c in range(128):
        assert np.array_equal(out[c], want[c % 3]), f"request {c} differs"
    assert not out[128].any(), "the rejected request's buffer was written"


def test_serving_three_resident_databases(hip_device):
    # the serving workload: 1, 16 and 32 GiB databases (seeds 0-2, own U/that/prefixes),
    # Poisson arrivals at 500 req/s with uniform database and frame choice, 1% corrupted
    # frames; every status must be the decoder's and every answered response must equal the
    # oracle's for its frame (part 0 and the last part, from one-part engines)
    import torch

    from altopir_public_amd.serve import Server, run_load

    if FullDb.current is not None:
        FullDb.current.close()
        FullDb.current = None
    engines, inputs, parts = [], [], []
    for i, gib in enumerate((1, 16, 32)):
        rows = gib * 2**30 // COLS
        inp = workload.make_inputs(rows, COLS, n=N, tg=TG, seed=i, db=False)
        d_db = synthetic.uniform_bytes_device(i, rows * COLS)
        torch.cuda.synchronize()
        engines.append(Engine.from_device_db(d_db.data_ptr(), inp))
        last = inp.outputs - 1
        parts.append({o: d_db[o * N * rows:(o + 1) * N * rows].cpu().numpy() for o in (0, last)})
        del d_db
        torch.cuda.empty_cache()
        inputs.append(inp)
        log(f"serving: database {i} ({gib} GiB) ready")
    pool = 6
    pools = []
    for i, inp in enumerate(inputs):
        L = inp.request_len
        stride = (7 + L + 255) // 256 * 256
        arena = workload.make_requests_device(pool, inp.rows, inp.kp, seed=5000 + i, stride=stride).cpu().numpy()
        pools.append([np.ascontiguousarray(arena[c, 7:7 + L]) for c in range(pool)])
    # the oracle's part 0 and last part of every pool frame
    want = []
    for i, inp in enumerate(inputs):
        w = {}
        for o, cols_host in parts[i].items():
            sub = part_inputs(inp, o, cols_host)
            out_o = np.zeros((pool, sub.response_len), np.uint8)
            st_o = np.full(pool, -1, np.int32)
            oc.oracle_answer(sub, pools[i], out_o, st_o)
            assert (st_o == 0).all()
            w[o] = out_o
        want.append(w)
        log(f"serving: oracle answers of database {i} ready")
    srv = Server(engines, max_batch=128, slots=256, contexts=2)
    rng = np.random.default_rng(5)
    count = 1500
    dbs = rng.integers(3, size=count)
    pick = rng.integers(pool, size=count)
    bad = rng.random(count) < 0.01
    bad[7] = True  # at least one
    frames = []
    for i in range(count):
        f = pools[dbs[i]][pick[i]]
        frames.append(workload.corrupt(f, rng, inputs[dbs[i]].kp) if bad[i] else f)
    want_status = [workload.frame_status(frames[i], inputs[dbs[i]].rows, inputs[dbs[i]].kp) for i in range(count)]
    errors = []

    def check(i, d):
        if errors:
            return
        if d.status != want_status[i]:
            errors.append(f"request {i} (database {dbs[i]}): status {d.status}, expected {want_status[i]}")
        elif d.status == 0:
            inp = inputs[dbs[i]]
            for o, w in want[dbs[i]].items():
                if d.response[o * inp.part_len:(o + 1) * inp.part_len] != w[pick[i]].tobytes():
                    errors.append(f"request {i} (database {dbs[i]}, frame {pick[i]}): part {o} differs")
                    break

    log("serving: 1500 requests at 500 req/s")
    res = run_load(srv, frames, [int(x) for x in dbs], rate=500.0, seed=55, on_done=check)
    srv.close()
    for e in engines:
        e.close()
    assert not errors, errors[0]
    s = res.summary()
    log(f"serving: {s}")
    assert s["answered"] == sum(1 for x in want_status if x == 0)
    assert sum(s["rejected"].values()) == sum(1 for x in want_status if x != 0) >= 1
    assert s["p99_ms"] >= s["p50_ms"] > 0


@pytest.mark.parametrize("rows, cols_checked", [(32768, (0, 16383, 32767)), (2097152, (12345, 32767))])
def test_full_size_hint_columns_match_the_oracle(hip_device, rows, cols_checked):
    # the offline hint (src/server.rs:656-709, 738-768) of the 1 GiB and the 64 GiB database:
    # 2048 x records x 32,768 columns on the device; whole columns of it against the oracle
    Q = 4294955009
    db = FullDb.get(rows)
    rng = np.random.default_rng(rows)
    a = rng.integers(0, Q, size=rows, dtype=np.uint64)
    t0 = time.perf_counter()
    hint = db.eng.hint0(a.reshape(rows // N, N))
    log(f"{rows} records: hint of 2048 x {COLS} in {(time.perf_counter() - t0) * 1e3:.0f} ms (with its copies)")
    for c in cols_checked:
        col = db.d_db[c * rows:(c + 1) * rows].cpu().numpy()  # column c of the column-major database
        want = oc.oracle_hint0(a, col, rows, N, 0, 1)
        assert np.array_equal(hint[:, c], want[:, 0]), f"column {c}"
    del hint
