"""The offline packing precomputation on the device (alto_inspiring_offline) and the whole
scheme end to end on the GPU.

  * alto_inspiring_offline equals the oracle's orc_inspiring_offline (itself pinned against the
    reference's direct aggregate form and by decryption, tests/test_client.py) in every output:
    bold_t, bold_t_bar, bold_t_hat and a_hat, for n = 32, 256 and 2048;
  * end to end: the server's setup on the device (alto_hint0, alto_inspiring_offline,
    alto_pack_build) feeds the HIP engine, which answers queries of the client of client.py;
    every response decrypts to its database record (src/service.rs:227-257), also at the
    benchmark's 1 GiB shape (32,768 records x 32 KB).
"""
from __future__ import annotations

import numpy as np
import pytest

from altopir_public_amd import client as cl, offline, synthetic, workload
from altopir_public_amd.engine import Engine
import oracle_ctypes as oc

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,rows,cols", [(32, 256, 64), (256, 1024, 512), (2048, 2048, 4096)])
def test_device_offline_matches_the_oracle(hip_device, n, rows, cols):
    p = cl.Params(rows, cols, n, 4)
    db = synthetic.uniform_bytes(n + 1, rows * cols)
    hint = oc.oracle_hint0(cl.hint_query_polys(p), db, rows, n)
    got = offline.device_offline(p, hint)
    w, v = cl.masks(p)
    psi = workload.primitive_root_2n(n)
    for o in range(p.outputs):
        want = oc.oracle_offline(hint, o, w, v, n, p.tg, psi)
        for name, g, x in zip(("bold_t", "bold_t_bar", "bold_t_hat", "a_hat"), got[o], want):
            assert np.array_equal(g, x), (o, name)


def _device_server(p, db_col):
    hint = offline.device_hint(p, db_col)
    parts = offline.device_offline(p, hint)
    return offline.server_inputs(p, db_col, parts, device=True)


@pytest.mark.parametrize("rows,cols,idx", [(8192, 4096, [0, 1, 2047, 2048, 5000, 8191]),
                                           (32768, 32768, [0, 12345, 32767])])
def test_device_setup_and_engine_answers_decrypt_to_the_records(hip_device, rows, cols, idx):
    p = cl.Params(rows, cols, 2048, 4)
    db = synthetic.uniform_bytes(7, rows * cols).reshape(cols, rows)
    db_col = np.ascontiguousarray(db).reshape(-1)
    inp = offline.device_setup(p, db_col) if rows >= 32768 else _device_server(p, db_col)
    eng = Engine.from_inputs(inp)
    clients, reqs = [], []
    for k, i in enumerate(idx):
        c, f = cl.new_request(p, i, seed=77 + k)
        clients.append(c)
        reqs.append(f)
    reqs.append(reqs[0][:-4])  # a short frame: status 1
    out = np.zeros((len(reqs), inp.response_len), np.uint8)
    st = np.full(len(reqs), -1, np.int32)
    answered, _ = eng.answer(reqs, out, st)
    eng.close()
    assert answered == len(idx) and list(st) == [0] * len(idx) + [1]
    for k, i in enumerate(idx):
        assert cl.decrypt_response(p, clients[k], out[k]) == db[:, i].tobytes(), i


def test_device_offline_rejects_unsupported_shapes(hip_device):
    from altopir_public_amd import _lib

    p = cl.Params(256, 64, 32, 4)
    hint = np.zeros((32, 64), np.uint64)
    lib = _lib.load()
    z = np.zeros(16 * 4 * 32, np.uint64)
    # three parts of 32 columns do not fit a hint of 64 columns
    rc = lib.alto_inspiring_offline(z.ctypes.data, z.ctypes.data, z.ctypes.data, z.ctypes.data, hint.ctypes.data, 64,
                                    3, z.ctypes.data, z.ctypes.data, 32, 4)
    assert rc == -1 and "unsupported shape" in _lib.last_error()
    assert offline.device_offline(p, hint) is not None


@pytest.mark.parametrize("n,rows,cols,on_device", [(256, 1024, 512, False), (2048, 4096, 4096, True)])
def test_one_call_setup_equals_the_staged_setup(hip_device, n, rows, cols, on_device):
    # alto_scheme_setup keeps the hint and the aggregates on the device; its packing matrix, tail
    # keys and prefixes equal those of the three separate calls
    import torch

    p = cl.Params(rows, cols, n, 4)
    db_col = synthetic.uniform_bytes(n + 3, rows * cols)
    staged = _device_server(p, db_col)
    if on_device:
        d = torch.from_numpy(db_col.copy()).to("cuda")
        one = offline.device_setup(p, d.data_ptr(), on_device=True)
    else:
        one = offline.device_setup(p, db_col)
    assert np.array_equal(one.ub, staged.ub)
    for o in range(p.outputs):
        assert np.array_equal(one.that[o], staged.that[o]), o
        assert np.array_equal(one.prefixes[o], staged.prefixes[o]), o
    assert np.array_equal(one.inv_table, staged.inv_table)
