"""Parity at the benchmark's full database sizes (1 GiB and 64 GiB, 32 KB records).

The oracle cannot answer a 64 GiB batch in test time, but a response part o depends only on
database columns [o n, (o + 1) n), the same packing-matrix columns, that_o and prefix_o
(src/amx_kernels.cpp:602-607).  So each checked part is compared with the oracle run on a
one-part engine built from exactly those columns: the full-size engine's bytes for that part
must be identical.  The database is generated in device memory; the checked column slices are
copied back for the oracle.
"""
import numpy as np
import pytest

from altopir_public_amd import synthetic, workload
from altopir_public_amd.engine import Engine
import oracle_ctypes as oc

pytestmark = pytest.mark.gpu

KiB = 1024


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


@pytest.mark.parametrize("rows,parts,batch", [(32768, None, 2), (262144, (0, 7), 136), (2097152, (0, 15), 2)])
def test_full_size_parts_match_the_oracle(hip_device, rows, parts, batch):
    # 1 GiB, all parts; 8 GiB with an uneven batch past 128 queries (the 16 x 24-tile shape);
    # 64 GiB (32 record splits of the scan)
    import torch

    cols, n, tg = 32768, 2048, 4
    inp = workload.make_inputs(rows, cols, n=n, tg=tg, seed=0, db=False)
    d_db = synthetic.uniform_bytes_device(0, rows * cols)
    torch.cuda.synchronize()
    eng = Engine.from_device_db(d_db.data_ptr(), inp)
    reqs = workload.make_requests(batch, rows, inp.kp, seed=2)
    rng = np.random.default_rng(9)
    reqs.append(workload.corrupt(reqs[0], rng, inp.kp))  # a malformed frame at full size
    B, rl = len(reqs), inp.response_len
    out = np.zeros((B, rl), np.uint8)
    st = np.full(B, -1, np.int32)
    answered, _ = eng.answer(reqs, out, st)
    eng.close()
    want = [workload.frame_status(f, rows, inp.kp) for f in reqs]
    assert list(st) == want and answered == want.count(0)
    for o in (parts if parts is not None else range(inp.outputs)):
        sl = d_db[o * n * rows:(o + 1) * n * rows].cpu().numpy()
        sub = part_inputs(inp, o, sl)
        out_o = np.zeros((B, sub.response_len), np.uint8)
        st_o = np.full(B, -1, np.int32)
        assert oc.oracle_answer(sub, reqs, out_o, st_o) == answered
        assert list(st_o) == want
        for c in range(B):
            if want[c] == 0:
                got = out[c, o * inp.part_len:(o + 1) * inp.part_len]
                assert np.array_equal(got, out_o[c]), f"{rows} records: part {o} of request {c} differs"
    del d_db
    torch.cuda.empty_cache()
