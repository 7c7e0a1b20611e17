"""Parity of the HIP answer engine with the oracle, through the C ABI.

Restates tests/engine.rs for the HIP engine: byte-identical responses on the reference's
shapes (ROWS = 2048 and 8192, RECORD_BYTES = 4096, batches {1, 2, 72, 136}), the
rejection statuses of hand-made and fuzzed frames with the untouched-buffer rule, plus
shapes and edge cases the reference's tests do not cover (small rings, ld > rows, batch
sizes that are not multiples of the tile width, random prefixes).  The oracle
(oracle/liboracle.so) is pinned against the reference engine in test_oracle.py.
"""
from __future__ import annotations

import numpy as np
import pytest

from altopir_public_amd import workload
from altopir_public_amd.engine import Engine, Method
import oracle_ctypes as oc
import refmodel

pytestmark = pytest.mark.gpu


def run_both(inp, reqs, method=Method.HIP, sentinel=0):
    B = len(reqs)
    rl = inp.response_len
    eng = Engine.from_inputs(inp)
    out_g = np.full((B, rl), sentinel, np.uint8)
    st_g = np.full(B, -1, np.int32)
    answered, stage = eng.answer(reqs, out_g, st_g, method)
    out_o = np.full((B, rl), sentinel, np.uint8)
    st_o = np.full(B, -1, np.int32)
    r_o = oc.oracle_answer(inp, reqs, out_o, st_o)
    eng.close()
    return (answered, st_g, out_g, stage), (r_o, st_o, out_o)


def assert_same(g, o, label=""):
    (a, sg, og, _), (r, so, oo) = g, o
    assert a == r, f"{label}: answered {a} vs {r}"
    assert np.array_equal(sg, so), f"{label}: statuses {sg} vs {so}"
    bad = [i for i in range(len(og)) if not np.array_equal(og[i], oo[i])]
    assert not bad, f"{label}: responses {bad[:8]} differ from the oracle"


@pytest.fixture(scope="module")
def small_inputs():
    # tests/engine.rs:10-11: ROWS = 2048, RECORD_BYTES = 4096 (two outputs)
    return workload.make_inputs(2048, 4096, n=2048, tg=4, seed=1)


def test_responses_match_the_oracle(hip_device, small_inputs):
    inp = small_inputs
    reqs = workload.make_requests(5, inp.rows, inp.kp, seed=21)
    g, o = run_both(inp, reqs)
    assert g[0] == 5 and (g[1] == 0).all()
    assert_same(g, o, "rows 2048")


@pytest.mark.parametrize("method", [Method.AMX, Method.VNNI, Method.HIP])
def test_every_method_code_runs_the_hip_engine(hip_device, small_inputs, method):
    inp = small_inputs
    reqs = workload.make_requests(3, inp.rows, inp.kp, seed=22)
    g, o = run_both(inp, reqs, method=method)
    assert_same(g, o, f"method {int(method)}")


def test_malformed_requests_are_rejected_like_the_decoder_rejects_them(hip_device, small_inputs):
    inp = small_inputs
    r0, r1 = workload.make_requests(2, inp.rows, inp.kp, seed=23)
    frames, want = refmodel.hand_frames(r0, r1, inp.kp)
    for f, w in zip(frames, want):
        assert refmodel.decoder_accepts(f, inp.rows, inp.kp) == (w == 0)
    g, o = run_both(inp, frames, sentinel=0xA5)
    assert g[0] == 2
    assert list(g[1]) == want
    for c in range(1, 7):
        assert (g[2][c] == 0xA5).all(), "a rejected buffer was written"
    assert_same(g, o, "hand frames")


def test_fuzzed_frames_follow_the_decoder(hip_device, small_inputs):
    inp = small_inputs
    base = workload.make_requests(4, inp.rows, inp.kp, seed=24)
    eng = Engine.from_inputs(inp)
    accepted = 0
    rl = inp.response_len
    for frames in refmodel.fuzz_batches(base, inp.kp):
        out = np.zeros((len(frames), rl), np.uint8)
        st = np.full(len(frames), -1, np.int32)
        eng.answer(frames, out, st, Method.HIP)
        for f, s in zip(frames, st):
            assert (s == 0) == refmodel.decoder_accepts(f, inp.rows, inp.kp), "engine and decoder disagree"
        accepted += int((st == 0).sum())
        out_o = np.zeros_like(out)
        st_o = np.full(len(frames), -1, np.int32)
        oc.oracle_answer(inp, frames, out_o, st_o)
        assert np.array_equal(st, st_o)
        assert np.array_equal(out[st == 0], out_o[st_o == 0])
    eng.close()
    assert 0 < accepted < 512


@pytest.mark.parametrize("B", [1, 2, 72, 136])
def test_multi_block_and_chunked_batches_match_the_oracle(hip_device, B):
    # tests/engine.rs:196-245: ROWS = 8192, batches {1, 2, 72, 136}
    inp = _block_inputs()
    reqs = workload.make_requests(B, inp.rows, inp.kp, seed=100 + B)
    g, o = run_both(inp, reqs)
    assert g[0] == B
    assert_same(g, o, f"B={B}")


@pytest.mark.parametrize("B", [3, 9, 24, 33, 48, 49, 129])
def test_built_in_shapes_at_every_bucket_edge_match_the_oracle(hip_device, B):
    # the built-in shape choice changes at 4, 8, 32 and 48 queries (streaming, 8-row ring,
    # 12-row ring, then the 16-row workgroups that stream B blocks into registers); each side
    # of those edges, with the default packing shape of the same batch size
    inp = _block_inputs()
    reqs = workload.make_requests(B, inp.rows, inp.kp, seed=300 + B)
    g, o = run_both(inp, reqs)
    assert g[0] == B
    assert_same(g, o, f"B={B}")


_BLOCK = None


def _block_inputs():
    global _BLOCK
    if _BLOCK is None:
        _BLOCK = workload.make_inputs(8192, 4096, n=2048, tg=4, seed=2)
    return _BLOCK


@pytest.mark.parametrize("rows,cols,n,tg,B,ld,prefix,table", [
    (64, 32, 16, 8, 3, 64, "random", "random"),       # smallest legal shapes
    (128, 64, 32, 4, 7, 192, "random", "ntt"),        # ld > rows
    (2048, 512, 256, 4, 13, 2048, "random", "ntt"),   # 2 outputs of a small ring
    (4096, 1024, 1024, 2, 9, 4096, "row0", "random"),  # tg = 2
    (1024, 8192, 4096, 4, 3, 1024, "random", "random"),  # n = 4096 (largest ring)
])
def test_shapes_beyond_the_reference_tests(hip_device, rows, cols, n, tg, B, ld, prefix, table):
    inp = workload.make_inputs(rows, cols, n=n, tg=tg, seed=rows + cols + n, ld=ld, prefix=prefix, table=table)
    reqs = workload.make_requests(B, rows, inp.kp, seed=B)
    g, o = run_both(inp, reqs)
    assert_same(g, o, f"{rows}x{cols} n={n}")


def test_extreme_query_and_key_values(hip_device, small_inputs):
    inp = small_inputs
    kp, rows = inp.kp, inp.rows
    reqs = []
    for yv, zv, qv in [(0, 0, 0), (workload.Q - 1, workload.Q - 1, 0xFFFFFFFF), (1, 2, 1 << 31),
                       (workload.Q - 1, 0, 0x80808080)]:
        reqs.append(workload.encode_request(np.full(kp, yv, np.uint32), np.full(kp, zv, np.uint32),
                                            np.full(rows, qv, np.uint32)))
    g, o = run_both(inp, reqs)
    assert g[0] == 4
    assert_same(g, o, "extreme values")


def test_empty_batch_and_null_frames(hip_device, small_inputs):
    inp = small_inputs
    eng = Engine.from_inputs(inp)
    out = np.zeros((0, inp.response_len), np.uint8)
    assert eng.answer([], out, np.zeros(0, np.int32))[0] == 0
    r = workload.make_requests(1, inp.rows, inp.kp, seed=5)[0]
    out = np.full((3, inp.response_len), 7, np.uint8)
    st = np.full(3, -1, np.int32)
    a, _ = eng.answer([None, r, b""], out, st)
    assert a == 1 and list(st) == [1, 0, 1]
    assert (out[0] == 7).all() and (out[2] == 7).all()
    eng.close()


def test_a_batch_of_many_row_blocks_matches_smaller_batches_and_the_oracle(hip_device):
    # 1,000 queries = 250 A tiles: 16 row blocks of the 16-row shapes, whose workgroups the tile
    # map groups as 8 row blocks x 4 column tiles per XCD (8,192 columns = 32 column tiles of
    # 256).  The same frames answered in batches of 250 (4 row blocks, one column tile's row
    # blocks together) give the same bytes, and four of them the oracle's.
    inp = workload.make_inputs(2048, 8192, n=2048, tg=4, seed=77, prefix="random")
    B, rl = 1000, inp.response_len
    reqs = workload.make_requests(B, inp.rows, inp.kp, seed=78)
    eng = Engine.from_inputs(inp)
    out = np.zeros((B, rl), np.uint8)
    st = np.full(B, -1, np.int32)
    assert eng.answer(reqs, out, st)[0] == B
    part = np.zeros_like(out)
    for c0 in range(0, B, 250):
        stp = np.full(250, -1, np.int32)
        assert eng.answer(reqs[c0:c0 + 250], part[c0:c0 + 250], stp)[0] == 250
    eng.close()
    bad = [c for c in range(B) if not np.array_equal(out[c], part[c])]
    assert not bad, f"responses {bad[:8]} differ between the batch of 1000 and batches of 250"
    pick = [0, 499, 613, 999]
    out_o = np.zeros((len(pick), rl), np.uint8)
    assert oc.oracle_answer(inp, [reqs[i] for i in pick], out_o, np.full(len(pick), -1, np.int32)) == len(pick)
    for k, i in enumerate(pick):
        assert np.array_equal(out[i], out_o[k]), f"response {i} differs from the oracle"


def test_repeated_batches_of_different_sizes_reuse_the_engine(hip_device, small_inputs):
    inp = small_inputs
    eng = Engine.from_inputs(inp)
    rl = inp.response_len
    for B, seed in [(9, 1), (2, 2), (17, 3), (1, 4)]:
        reqs = workload.make_requests(B, inp.rows, inp.kp, seed=300 + seed)
        out = np.zeros((B, rl), np.uint8)
        st = np.full(B, -1, np.int32)
        assert eng.answer(reqs, out, st)[0] == B
        out_o = np.zeros_like(out)
        oc.oracle_answer(inp, reqs, out_o, np.full(B, -1, np.int32))
        assert np.array_equal(out, out_o), f"batch {B}"
    eng.close()


@pytest.mark.parametrize("B", [1, 6, 13, 72])
def test_every_kernel_variant_gives_the_same_bytes(hip_device, B):
    from altopir_public_amd import _lib, autotune

    inp = _block_inputs()
    reqs = workload.make_requests(B, inp.rows, inp.kp, seed=500 + B)
    rl = inp.response_len
    out_o = np.zeros((B, rl), np.uint8)
    oc.oracle_answer(inp, reqs, out_o, np.full(B, -1, np.int32))
    eng = Engine.from_inputs(inp)
    lib = _lib.load()
    names = autotune.variant_names()
    for i, name in enumerate(names):
        for which in (0, 1):
            assert lib.alto_tune_set(eng.ptr, which, B, i) == 0
        out = np.zeros((B, rl), np.uint8)
        st = np.full(B, -1, np.int32)
        assert eng.answer(reqs, out, st)[0] == B
        bad = [c for c in range(B) if not np.array_equal(out[c], out_o[c])]
        assert not bad, f"variant {name}: responses {bad[:8]} differ"
    eng.close()


@pytest.mark.parametrize("steps, B", [(37, 5), (37, 72), (453, 37)])
def test_every_kernel_variant_on_an_odd_number_of_k_steps(hip_device, steps, B):
    # 2,368 records = 37 k-steps of 64 (a prime): record-range splits and the unrolled k-loops
    # of the contraction kernels end on partial groups of steps; 453 k-steps: the ring shapes
    # split the record range 4 ways, into ranges of 113 and 114 steps
    from altopir_public_amd import _lib, autotune

    inp = workload.make_inputs(64 * steps, 4096, n=2048, tg=4, seed=steps, prefix="random")
    reqs = workload.make_requests(B, inp.rows, inp.kp, seed=600 + B)
    rl = inp.response_len
    out_o = np.zeros((B, rl), np.uint8)
    oc.oracle_answer(inp, reqs, out_o, np.full(B, -1, np.int32))
    eng = Engine.from_inputs(inp)
    lib = _lib.load()
    for i, name in enumerate(autotune.variant_names()):
        for which in (0, 1):
            assert lib.alto_tune_set(eng.ptr, which, B, i) == 0
        out = np.zeros((B, rl), np.uint8)
        st = np.full(B, -1, np.int32)
        assert eng.answer(reqs, out, st)[0] == B
        bad = [c for c in range(B) if not np.array_equal(out[c], out_o[c])]
        assert not bad, f"variant {name}: responses {bad[:8]} differ"
    eng.close()


def test_one_graph_answers_every_batch_size_of_its_bucket(hip_device, small_inputs):
    # batches of 33-48 queries (12 A tiles at most) share one captured chain, sized for 48 and
    # told each batch's size at replay: every answer equals the eager one, the rows past the
    # batch are never written, and the context holds one graph per bucket
    import torch

    from altopir_public_amd import _lib

    lib = _lib.load()
    inp = small_inputs
    L, rl = inp.request_len, inp.response_len
    stride = (7 + L + 255) // 256 * 256
    eng = Engine.from_inputs(inp)
    ctx = lib.alto_engine_context(eng.ptr)
    Bmax = 48
    slots = workload.make_requests_device(Bmax, inp.rows, inp.kp, seed=811, stride=stride)
    slots[5, 7 + 9 + 4 * 77:7 + 9 + 4 * 78] = torch.tensor([255, 255, 255, 255], dtype=torch.uint8,
                                                           device=slots.device)  # key >= Q: status 3
    stream = torch.cuda.current_stream().cuda_stream

    def run(B, graphs):
        eng.set_graphs(graphs)
        out = torch.full((Bmax, rl), 0xA5, dtype=torch.uint8, device="cuda")
        st = torch.zeros(Bmax, dtype=torch.int32, device="cuda")
        assert eng.answer_device(slots.data_ptr(), stride, B, out.data_ptr(), rl, st.data_ptr(), stream) == 0
        torch.cuda.synchronize()
        return out.cpu().numpy(), st.cpu().numpy()

    for B in (37, 48, 33, 44, 40, 37):
        want_out, want_st = run(B, False)
        got_out, got_st = run(B, True)
        assert list(got_st[:B]) == list(want_st[:B]) and got_st[5] == 3 and (got_st[B:] == 0).all()
        assert np.array_equal(got_out[:B], want_out[:B]), f"batch {B}: replay differs from the eager batch"
        assert (got_out[B:] == 0xA5).all(), f"batch {B}: rows past the batch were written"
        assert lib.alto_context_graphs(ctx) == 1, "batches of one bucket captured more than one graph"
    frames = [slots[c, 7:7 + L].cpu().numpy() for c in range(37)]
    out_o = np.zeros((37, rl), np.uint8)
    st_o = np.full(37, -1, np.int32)
    oc.oracle_answer(inp, frames, out_o, st_o)
    got_out, got_st = run(37, True)
    assert list(got_st[:37]) == list(st_o)
    for c in range(37):
        if st_o[c] == 0:
            assert np.array_equal(got_out[c], out_o[c])
    run(3, True)  # another bucket: a second graph
    assert lib.alto_context_graphs(ctx) == 2
    eng.close()


def test_graph_replays_match_eager_batches(hip_device, small_inputs):
    # The device entry point captures a batch into a HIP graph and replays it for the same
    # pointers; every replay must equal an eager run of the same frames, with other eager
    # work and synchronisations between replays.
    import torch

    inp = small_inputs
    B = 6
    L, rl = inp.request_len, inp.response_len
    stride = (7 + L + 255) // 256 * 256
    eng = Engine.from_inputs(inp)
    slots = torch.zeros((B, stride), dtype=torch.uint8, device="cuda")
    status = torch.zeros(B, dtype=torch.int32, device="cuda")
    out = torch.zeros((B, rl), dtype=torch.uint8, device="cuda")
    stream = torch.cuda.Stream()
    for rep in range(4):
        reqs = workload.make_requests(B, inp.rows, inp.kp, seed=700 + rep)
        if rep == 2:
            reqs[3] = bytes(reqs[3][:9]) + b"\xff\xff\xff\xff" + bytes(reqs[3][13:])  # key word >= Q
        host = np.zeros((B, stride), np.uint8)
        for c, f in enumerate(reqs):
            host[c, 7:7 + L] = np.frombuffer(bytes(f), np.uint8)
        slots.copy_(torch.from_numpy(host))
        status.zero_()
        out.zero_()
        torch.cuda.synchronize()
        eng.set_graphs(True)
        assert eng.answer_device(slots.data_ptr(), stride, B, out.data_ptr(), rl, status.data_ptr(),
                                 stream.cuda_stream) == 0
        stream.synchronize()
        got, st = out.cpu().numpy(), status.cpu().numpy()
        _ = torch.ones(1 << 20, device="cuda").sum().item()  # unrelated eager work
        out_e = np.zeros((B, rl), np.uint8)
        st_e = np.full(B, -1, np.int32)
        eng.answer(reqs, out_e, st_e)
        assert np.array_equal(st, st_e), f"replay {rep}: statuses"
        ok = st_e == 0
        assert np.array_equal(got[ok], out_e[ok]), f"replay {rep}: responses differ from the eager batch"
    eng.close()


def test_stage_totals_count_every_timed_batch(hip_device, small_inputs):
    # Eager device batches are timed through a ring of 32 event sets; the running totals
    # must count every one of more than a ring's worth of back-to-back batches, and graph
    # replays none.
    import torch

    inp = small_inputs
    B = 3
    L, rl = inp.request_len, inp.response_len
    stride = (7 + L + 255) // 256 * 256
    eng = Engine.from_inputs(inp)
    host = np.zeros((B, stride), np.uint8)
    for c, f in enumerate(workload.make_requests(B, inp.rows, inp.kp, seed=91)):
        host[c, 7:7 + L] = f
    slots = torch.from_numpy(host).to("cuda")
    status = torch.zeros(B, dtype=torch.int32, device="cuda")
    out = torch.zeros((B, rl), dtype=torch.uint8, device="cuda")
    stream = torch.cuda.Stream()
    eng.set_graphs(False)
    eng.stage_totals(reset=True)
    for _ in range(45):
        assert eng.answer_device(slots.data_ptr(), stride, B, out.data_ptr(), rl, status.data_ptr(),
                                 stream.cuda_stream) == 0
    n, ms = eng.stage_totals()
    assert n == 45 and all(v >= 0 for v in ms) and ms[2] > 0 and ms[5] > 0
    last = eng.last_timings()
    assert len(last) == 6 and last[2] > 0
    eng.set_graphs(True)
    for _ in range(3):
        assert eng.answer_device(slots.data_ptr(), stride, B, out.data_ptr(), rl, status.data_ptr(),
                                 stream.cuda_stream) == 0
    n2, ms2 = eng.stage_totals(reset=True)
    assert n2 == 45 and ms2 == ms
    assert eng.stage_totals()[0] == 0
    assert int(status.sum()) == 0
    eng.close()


@pytest.mark.parametrize("rows,cols", [(64, 96), (50, 37), (1024, 512)])
def test_packed_product_matches_the_oracle(hip_device, rows, cols):
    # the SimplePIR packed product of src/matmul.cpp, served by the scan kernels
    import ctypes as C

    from altopir_public_amd import _lib, synthetic

    lib, orc = _lib.load(), oc.oracle()
    orc.orc_matmul_vec_packed.argtypes = [C.c_void_p] * 3 + [C.c_size_t] * 3
    for bc, fn in ((1, "matMulVecPacked"), (2, "matMulVecPacked2"), (4, "matMulVecPacked4"), (8, "matMulVecPacked8")):
        a = synthetic.uniform_u32(rows + bc, rows * cols)
        b = synthetic.uniform_u32(cols + bc, bc * 4 * cols)
        out = synthetic.uniform_u32(bc, rows * bc + 8)
        want, got = out.copy(), out.copy()
        orc.orc_matmul_vec_packed(want.ctypes.data, a.ctypes.data, b.ctypes.data, rows, cols, bc)
        getattr(lib, fn)(got.ctypes.data, a.ctypes.data, b.ctypes.data, rows, cols)
        assert np.array_equal(got, want), fn


@pytest.mark.parametrize("B", [6, 3])
def test_timed_graph_replays_feed_the_stage_totals(hip_device, small_inputs, B):
    # with graph timing on, each replay records its stage events as graph nodes (one captured
    # graph per event set); every replay is counted, its times are positive and of the same
    # size as eager launches', and the responses equal the eager ones (B = 3 runs both
    # contractions in one launch: its packing reads zero, the pair counts as the scan)
    import torch

    inp = small_inputs
    L, rl = inp.request_len, inp.response_len
    stride = (7 + L + 255) // 256 * 256
    eng = Engine.from_inputs(inp)
    one_launch = eng.fused(B)
    assert one_launch == (B <= 4)
    host = np.zeros((B, stride), np.uint8)
    for c, f in enumerate(workload.make_requests(B, inp.rows, inp.kp, seed=93)):
        host[c, 7:7 + L] = f
    slots = torch.from_numpy(host).to("cuda")
    status = torch.zeros(B, dtype=torch.int32, device="cuda")
    stream = torch.cuda.Stream()
    outs, totals = [], []
    for graphs in (False, True):
        out = torch.zeros((B, rl), dtype=torch.uint8, device="cuda")
        eng.set_graphs(graphs)
        eng.set_graph_timing(graphs)
        for _ in range(5):  # captures
            assert eng.answer_device(slots.data_ptr(), stride, B, out.data_ptr(), rl, status.data_ptr(),
                                     stream.cuda_stream) == 0
        stream.synchronize()
        eng.stage_totals(reset=True)
        for _ in range(13):
            assert eng.answer_device(slots.data_ptr(), stride, B, out.data_ptr(), rl, status.data_ptr(),
                                     stream.cuda_stream) == 0
        n, ms = eng.stage_totals(reset=True)
        assert n == 13, (graphs, n)
        assert all(v > 0 for v in (ms[0], ms[2], ms[5])), (graphs, ms)
        assert (ms[4] == 0) if one_launch else (ms[4] > 0), (graphs, ms)
        totals.append(ms)
        outs.append(out.cpu().numpy())
    eng.set_graph_timing(False)
    assert int(status.sum()) == 0
    assert np.array_equal(outs[0], outs[1])
    # the contractions' times agree within a factor of two between eager and replayed batches
    for i in ((2,) if one_launch else (2, 4)):
        assert 0.5 < totals[1][i] / totals[0][i] < 2.0, (i, totals)
    eng.close()


def test_tuner_records_a_variant_per_bucket_and_keeps_the_bytes(hip_device, small_inputs, tmp_path):
    # the tuner times every variant on device-resident batches and records the fastest per
    # batch-size bucket; answers after tuning equal the oracle's, and a saved table applies
    # to a fresh engine
    from altopir_public_amd import autotune

    inp = small_inputs
    eng = Engine.from_inputs(inp)
    # in the engine's default graph mode: the tuner times eager batches and restores the mode
    table = autotune.tune(eng, inp, [1, 6, 40], seed=3, reps=1)
    assert eng.graphs
    names = autotune.variant_names()
    assert set(table) == {autotune.bucket(b) for b in (1, 6, 40)}
    for ent in table.values():
        assert ent["scan"][0] in names and ent["pack"][0] in names
        assert eng._lib.alto_tune_get(eng.ptr, autotune.SCAN, ent["B"]) == names.index(ent["scan"][0])
    reqs = workload.make_requests(40, inp.rows, inp.kp, seed=31)
    out_g = np.zeros((40, inp.response_len), np.uint8)
    st_g = np.full(40, -1, np.int32)
    answered, stage = eng.answer(reqs, out_g, st_g)
    out_o = np.zeros_like(out_g)
    st_o = np.full(40, -1, np.int32)
    r_o = oc.oracle_answer(inp, reqs, out_o, st_o)
    assert_same((answered, st_g, out_g, stage), (r_o, st_o, out_o), "after tuning")
    path = tmp_path / "tune.json"
    autotune.save(table, path, inp.rows, inp.cols)
    fresh = Engine.from_inputs(inp)
    assert autotune.load(fresh, path, inp.rows, inp.cols) == len(table)
    for ent in table.values():
        assert fresh._lib.alto_tune_get(fresh.ptr, autotune.PACK, ent["B"]) == names.index(ent["pack"][0])
    fresh.close()
    eng.close()


@pytest.mark.parametrize("B", [1, 6, 12, 40, 60, 136])
def test_scan_and_packing_in_one_launch_give_the_same_bytes(hip_device, small_inputs, B):
    # the built-in shape pairs run both contractions in one launch (B = 1 and 6 streaming,
    # 60 and 136 the register-streaming 16 x 16 tiles; 12 and 40 scan on 8- and 12-row tiles,
    # which have no fused pair): fused, unfused and oracle answers agree.  By default only batches
    # of one A tile (1-4 queries) run fused (on this database of 2048 <= 4 KP records).
    fusable = B not in (12, 40)
    inp = small_inputs
    reqs = workload.make_requests(B, inp.rows, inp.kp, seed=1200 + B)
    eng = Engine.from_inputs(inp)
    assert eng.fused(B) == (B <= 4)
    outs = []
    for on in (True, False):
        eng.set_fusion(on)
        assert eng.fused(B) == (on and fusable)
        out = np.zeros((B, inp.response_len), np.uint8)
        st = np.full(B, -1, np.int32)
        answered, stage = eng.answer(reqs, out, st)
        assert answered == B and (st == 0).all()
        outs.append(out)
    eng.close()
    out_o = np.zeros_like(outs[0])
    st_o = np.full(B, -1, np.int32)
    assert oc.oracle_answer(inp, reqs, out_o, st_o) == B
    assert np.array_equal(outs[0], out_o), "fused responses differ from the oracle"
    assert np.array_equal(outs[1], out_o), "unfused responses differ from the oracle"


@pytest.mark.parametrize("rows,cols,n,ld_pad,on_device", [(2048, 4096, 2048, 0, False), (2048, 4096, 2048, 7, True),
                                                         (192, 352, 32, 32, False), (128, 96, 32, 0, True)])
def test_engine_from_records_matches_the_oracle(hip_device, rows, cols, n, ld_pad, on_device):
    # the database given record by record (row-major, as a file holds it) is transposed into the
    # engine's layout on the device: answers equal the column-major engine's and the oracle's
    import torch

    inp = workload.make_inputs(rows, cols, n=n, tg=4, seed=rows + cols)
    ld = cols + ld_pad
    rec = np.zeros((rows, ld), np.uint8)
    rec[:, :cols] = inp.db_record_major()
    B = 5
    reqs = workload.make_requests(B, inp.rows, inp.kp, seed=77)
    if on_device:
        d_rec = torch.from_numpy(rec).to("cuda")
        eng = Engine.from_records(d_rec.data_ptr(), True, ld, inp)
        del d_rec
    else:
        eng = Engine.from_records(rec.ctypes.data, False, ld, inp)
    out = np.zeros((B, inp.response_len), np.uint8)
    st = np.full(B, -1, np.int32)
    answered, stage = eng.answer(reqs, out, st)
    eng.close()
    out_o = np.zeros_like(out)
    st_o = np.full(B, -1, np.int32)
    r_o = oc.oracle_answer(inp, reqs, out_o, st_o)
    assert answered == r_o == B and np.array_equal(st, st_o)
    assert np.array_equal(out, out_o), "responses of the record-ingested engine differ from the oracle"


@pytest.mark.parametrize("b_bytes,K,a_elems,b_cols", [(1, 1, 4096, 300), (2, 3, 1000, 64), (4, 2, 777, 33),
                                                      (2, 11, 513, 17)])
def test_fast_batched_dot_product_matches_the_oracle(hip_device, b_bytes, K, a_elems, b_cols):
    # src/kernel.rs:14-192 with explicit moduli: the device's u64 partial sums wrap like the
    # reference's, then the CRT composition; full-range operands (sums past 2^64 included)
    from altopir_public_amd import _lib

    m0, m1 = 4294967291, 4294967279
    rng = np.random.default_rng(b_bytes * 100 + K)
    a = rng.integers(0, 1 << 63, (K, a_elems), dtype=np.uint64) * np.uint64(2) + rng.integers(0, 2, (K, a_elems),
                                                                                               dtype=np.uint64)
    b = rng.integers(0, 1 << (8 * b_bytes), (b_cols, a_elems), dtype=np.uint64).astype(f"<u{b_bytes}")
    c = rng.integers(0, m0 * m1, (K, b_cols + 3), dtype=np.uint64)
    want = c.copy()
    oc.oracle().orc_fast_batched_dot_product(want.ctypes.data, b_cols + 3, a.ctypes.data, a_elems, a_elems,
                                             b.ctypes.data, b_bytes, a_elems, b_cols, K, m0, m1)
    got = c.copy()
    assert _lib.load().alto_fast_batched_dot_product(got.ctypes.data, b_cols + 3, a.ctypes.data, a_elems, a_elems,
                                                     b.ctypes.data, b_bytes, a_elems, b_cols, K, m0, m1) == 0
    assert np.array_equal(got, want)
    assert np.array_equal(got[:, b_cols:], c[:, b_cols:])  # columns past b_cols untouched
    assert _lib.load().alto_fast_batched_dot_product(got.ctypes.data, b_cols + 3, a.ctypes.data, a_elems, a_elems,
                                                     b.ctypes.data, 3, a_elems, b_cols, K, m0, m1) == -1


@pytest.mark.parametrize("graphs", [False, True])
def test_null_stream_batch_keeps_the_null_stream_order(hip_device, small_inputs, graphs):
    # stream 0: the batch runs on the engine's own stream in the null stream's order.  The
    # frames are written on the null stream behind a spin kernel of about 10 ms, the outputs
    # copied there right after the call and the frames cleared after that, with no host wait
    # in between: the batch must see the frames and the copy the batch's outputs.
    import torch

    inp = small_inputs
    L, rl = inp.request_len, inp.response_len
    stride = (7 + L + 255) // 256 * 256
    B = 6
    good = workload.make_requests_device(B, inp.rows, inp.kp, seed=97, stride=stride)
    frames = [np.ascontiguousarray(good[c, 7:7 + L].cpu().numpy()) for c in range(B)]
    want = np.zeros((B, rl), np.uint8)
    want_st = np.full(B, -1, np.int32)
    assert oc.oracle_answer(inp, frames, want, want_st) == B
    eng = Engine.from_inputs(inp)
    eng.set_graphs(graphs)
    with torch.cuda.stream(torch.cuda.default_stream()):
        assert torch.cuda.current_stream().cuda_stream == 0
        for rep in range(3):  # the first call grows the workspace (and captures the graph)
            slots = torch.zeros_like(good)
            status = torch.zeros(B, dtype=torch.int32, device="cuda")
            out = torch.zeros((B, rl), dtype=torch.uint8, device="cuda")
            torch.cuda.synchronize()
            torch.cuda._sleep(20_000_000)
            slots.copy_(good)
            assert eng.answer_device(slots.data_ptr(), stride, B, out.data_ptr(), rl, status.data_ptr(), 0) == 0
            got, got_st = out.clone(), status.clone()
            slots.zero_()
            torch.cuda.synchronize()
            assert list(got_st.cpu().numpy()) == [0] * B, f"call {rep}: statuses {got_st.cpu().numpy()}"
            assert np.array_equal(got.cpu().numpy(), want), f"call {rep}: responses differ from the oracle"
    eng.close()


@pytest.mark.parametrize("layout", ["columns", "records"])
def test_engine_set_up_from_device_memory_keeps_the_null_stream_order(hip_device, layout):
    # the database is written into device memory on the null stream behind a spin kernel of
    # about 10 ms, and the engine is created at once, with no host wait: set-up must read it in
    # the null stream's order
    import torch

    inp = workload.make_inputs(2048, 4096, n=2048, tg=4, seed=5)
    host = np.ascontiguousarray(inp.db_col if layout == "columns" else inp.db_record_major())
    src = torch.from_numpy(host).to("cuda")
    dst = torch.zeros_like(src)
    torch.cuda.synchronize()
    with torch.cuda.stream(torch.cuda.default_stream()):
        torch.cuda._sleep(20_000_000)
        dst.copy_(src)
        if layout == "columns":
            eng = Engine.from_device_db(dst.data_ptr(), inp)
        else:
            eng = Engine.from_records(dst.data_ptr(), True, inp.cols, inp)
    B = 4
    reqs = workload.make_requests(B, inp.rows, inp.kp, seed=78)
    out = np.zeros((B, inp.response_len), np.uint8)
    st = np.full(B, -1, np.int32)
    answered, _ = eng.answer(reqs, out, st)
    eng.close()
    out_o = np.zeros_like(out)
    st_o = np.full(B, -1, np.int32)
    assert answered == oc.oracle_answer(inp, reqs, out_o, st_o) == B
    assert np.array_equal(out, out_o), f"engine set up from device {layout} differs from the oracle"
