"""Scan, packing and tail of a batch in one persistent launch (the step kernel, alto_set_fusion
mode 2) against the oracle.

The step kernel counts the workgroups that have stored their results per output part and
runs a part's tails once the count is complete, on whichever workgroups claim them; the
cases cover every step pair of the built-in policy (1, 2, 3-4, 5-8, 9-12 and more A tiles),
parts narrower and wider than a tile's columns (a tile meeting up to eight parts), record
ranges split across workgroups (atomic accumulation), rejected frames among the tails,
captured graphs sized for a bucket and replayed for smaller batches, and timed replays.
"""
from __future__ import annotations

import numpy as np
import pytest

from altopir_public_amd import workload
from altopir_public_amd.engine import Engine
import oracle_ctypes as oc

pytestmark = pytest.mark.gpu

SHAPES = {
    # two outputs of n = 2048 (tests/engine.rs:10-11)
    "two_parts": dict(rows=2048, cols=4096, n=2048, tg=4),
    # 16 parts of 256 columns: a streaming scan tile (256 columns) meets one, an 8 x 32 one two
    "sixteen_parts": dict(rows=2048, cols=4096, n=256, tg=4),
    # 16 parts of 32 columns: every scan tile meets 8 or more parts
    "parts_of_32": dict(rows=1024, cols=512, n=32, tg=4),
    # 453 k-steps: the ring shapes split the record range (atomic accumulation)
    "split_ranges": dict(rows=453 * 64, cols=4096, n=2048, tg=4),
}

CASES = ([("two_parts", b) for b in (1, 2, 3, 6, 12, 13, 40, 60, 136)]
         + [("sixteen_parts", b) for b in (1, 7, 40, 72)]
         + [("parts_of_32", b) for b in (1, 12, 40, 70)]
         + [("split_ranges", b) for b in (1, 12, 40, 52)])

_inputs = {}


def inputs(shape):
    if shape not in _inputs:
        _inputs[shape] = workload.make_inputs(**SHAPES[shape], seed=11, prefix="random")
    return _inputs[shape]


def oracle(inp, reqs):
    B = len(reqs)
    out = np.zeros((B, inp.response_len), np.uint8)
    st = np.full(B, -1, np.int32)
    r = oc.oracle_answer(inp, reqs, out, st)
    return r, st, out


@pytest.mark.parametrize("shape,B", CASES)
def test_step_launch_matches_the_oracle(hip_device, shape, B):
    inp = inputs(shape)
    reqs = workload.make_requests(B, inp.rows, inp.kp, seed=1300 + B)
    rng = np.random.default_rng(B)
    if B >= 3:
        # rejected frames among the tails: a corrupted one and a truncated one
        reqs[1] = np.frombuffer(workload.corrupt(reqs[1], rng, inp.kp), np.uint8)
        reqs[B - 1] = reqs[B - 1][:-4]
    eng = Engine.from_inputs(inp)
    eng.set_fusion(2)
    assert eng.step(B), f"no step pair for {B} queries"
    out = np.zeros((B, inp.response_len), np.uint8)
    st = np.full(B, -1, np.int32)
    answered, stage = eng.answer(reqs, out, st)
    eng.close()
    r_o, st_o, out_o = oracle(inp, reqs)
    assert answered == r_o
    assert np.array_equal(st, st_o), (st, st_o)
    bad = [i for i in range(B) if not np.array_equal(out[i], out_o[i])]
    assert not bad, f"responses {bad[:8]} differ from the oracle"
    # {validate, scan, packing, tail}: the launch is timed as the scan; packing reads zero and
    # the tail only the gap between two events recorded back to back
    assert stage[1] > 0 and stage[2] == 0 and stage[3] < 0.05, stage


def test_bucket_graphs_of_the_step_launch(hip_device):
    # one graph captured for the 9-12-tile bucket (sized for 48 queries) replayed for 33, 37,
    # 41 and 48 queries, and batches of other buckets, with rejected frames: every answer
    # equals the two-launch-contraction eager one, which equals the oracle
    import torch

    inp = inputs("two_parts")
    L, rl = inp.request_len, inp.response_len
    stride = (7 + L + 255) // 256 * 256
    eng = Engine.from_inputs(inp)
    reqs = workload.make_requests(48, inp.rows, inp.kp, seed=77)
    reqs[5] = np.frombuffer(workload.corrupt(reqs[5], np.random.default_rng(3), inp.kp), np.uint8)
    host = np.zeros((48, stride), np.uint8)
    for c, f in enumerate(reqs):
        k = min(len(f), L)
        host[c, 7:7 + k] = f[:k]
    slots = torch.from_numpy(host).to("cuda")
    stream = torch.cuda.Stream()

    def run(B, fusion, graphs):
        eng.set_fusion(fusion)
        eng.set_graphs(graphs)
        out = torch.zeros((B, rl), dtype=torch.uint8, device="cuda")
        status = torch.zeros(B, dtype=torch.int32, device="cuda")
        for c in range(B):
            if len(reqs[c]) != L:
                status[c] = 1
        assert eng.answer_device(slots.data_ptr(), stride, B, out.data_ptr(), rl, status.data_ptr(),
                                 stream.cuda_stream) == 0
        stream.synchronize()
        return status.cpu().numpy(), out.cpu().numpy()

    r_o, st_o, out_o = oracle(inp, reqs)
    for B in (33, 37, 41, 48, 5, 20, 3):
        st_e, out_e = run(B, 0, False)
        assert np.array_equal(st_e, st_o[:B]) and np.array_equal(out_e, out_o[:B]), B
        for rep in range(2):
            st_g, out_g = run(B, 2, True)
            assert eng.step(B)
            assert np.array_equal(st_g, st_e), (B, rep)
            assert np.array_equal(out_g, out_e), (B, rep)
    eng.close()


def test_timed_replays_of_the_step_launch(hip_device):
    # graph timing inserts the stage events around the three-launch chain (validate, prep,
    # step); every replay is counted, the step shows under "scan", packing and tail read zero
    import torch

    inp = inputs("two_parts")
    B = 40
    L, rl = inp.request_len, inp.response_len
    stride = (7 + L + 255) // 256 * 256
    eng = Engine.from_inputs(inp)
    eng.set_fusion(2)
    host = np.zeros((B, stride), np.uint8)
    for c, f in enumerate(workload.make_requests(B, inp.rows, inp.kp, seed=94)):
        host[c, 7:7 + L] = f
    slots = torch.from_numpy(host).to("cuda")
    status = torch.zeros(B, dtype=torch.int32, device="cuda")
    stream = torch.cuda.Stream()
    outs = []
    for graphs in (False, True):
        out = torch.zeros((B, rl), dtype=torch.uint8, device="cuda")
        eng.set_graphs(graphs)
        eng.set_graph_timing(graphs)
        for _ in range(5):
            assert eng.answer_device(slots.data_ptr(), stride, B, out.data_ptr(), rl, status.data_ptr(),
                                     stream.cuda_stream) == 0
        stream.synchronize()
        eng.stage_totals(reset=True)
        for _ in range(9):
            assert eng.answer_device(slots.data_ptr(), stride, B, out.data_ptr(), rl, status.data_ptr(),
                                     stream.cuda_stream) == 0
        n, ms = eng.stage_totals(reset=True)
        assert n == 9, (graphs, n)
        assert ms[2] > 0 and ms[4] == 0 and ms[5] < 0.05 * 9, (graphs, ms)
        outs.append(out.cpu().numpy())
    eng.set_graph_timing(False)
    assert int(status.sum()) == 0
    assert np.array_equal(outs[0], outs[1])
    eng.close()
