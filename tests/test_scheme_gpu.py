"""The scheme end to end through the serving paths: clients' queries (client.py) for databases
set up on the device (offline.py) go through the serving loop and its TCP front end (the
reference server's wire protocol), and every response decrypts to its record
(src/service.rs:227-257) -- several databases, concurrent clients, corrupted frames among them."""
from __future__ import annotations

import threading

import numpy as np
import pytest

from altopir_public_amd import client as cl, net, offline, synthetic
from altopir_public_amd.engine import Engine
from altopir_public_amd.serve import Server

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def scheme_dbs():
    out = []
    for seed, (rows, cols) in enumerate([(4096, 4096), (8192, 2048)]):
        p = cl.Params(rows, cols)
        db = synthetic.uniform_bytes(40 + seed, rows * cols).reshape(cols, rows)
        db_col = np.ascontiguousarray(db).reshape(-1)
        inp = offline.server_inputs(p, db_col, offline.device_offline(p, offline.device_hint(p, db_col)))
        out.append((p, db, inp))
    return out


def _queries(p, count, seed):
    idx = [int(i) for i in np.random.default_rng(seed).integers(0, p.rows, count)]
    cs = [cl.new_request(p, i, seed=seed * 1000 + k) for k, i in enumerate(idx)]
    return idx, [c for c, _ in cs], [f for _, f in cs]


def test_served_queries_decrypt_to_the_records(hip_device, scheme_dbs):
    engines = [Engine.from_inputs(inp) for _, _, inp in scheme_dbs]
    srv = Server(engines, max_batch=16, slots=64, contexts=2)
    work = []
    for d, (p, db, _) in enumerate(scheme_dbs):
        idx, clients, frames = _queries(p, 12, 70 + d)
        work += [(d, i, c, f) for i, c, f in zip(idx, clients, frames)]
    # a corrupted frame per database: a key word >= Q (status 3)
    bad = []
    for d, (p, _, _) in enumerate(scheme_dbs):
        f = bytearray(bytes(work[d * 12][3]))
        f[9:13] = (0xFFFFFFFF).to_bytes(4, "little")
        bad.append((d, bytes(f)))

    def sender(items):
        for t, (d, _, _, f) in items:
            assert srv.submit(d, f, t) in (0, 1)

    items = list(enumerate(work))
    ts = [threading.Thread(target=sender, args=(items[k::3],)) for k in range(3)]
    for t in ts:
        t.start()
    for k, (d, f) in enumerate(bad):
        assert srv.submit(d, f, 1000 + k) in (0, 1)
    done = {}
    while len(done) < len(work) + len(bad):
        for x in srv.poll(100):
            done[x.tag] = x
    for t in ts:
        t.join()
    assert max(x.batch for x in done.values()) > 1, "no batch was shared"
    for t, (d, i, c, _) in enumerate(work):
        p, db, _ = scheme_dbs[d]
        assert done[t].status == 0, (t, srv.error())
        assert cl.decrypt_response(p, c, done[t].response) == db[:, i].tobytes(), (d, i)
    assert [done[1000 + k].status for k in range(len(bad))] == [3] * len(bad)
    srv.close()
    for e in engines:
        e.close()


def test_queries_over_tcp_decrypt_to_the_records(hip_device, scheme_dbs):
    p, db, inp = scheme_dbs[0]
    eng = Engine.from_inputs(inp)
    srv = Server([eng], max_batch=32, slots=64, contexts=2)
    ns = net.NetServer(srv)
    idx, clients, frames = _queries(p, 8, 90)
    socks = [net.connect("127.0.0.1", ns.port(0)) for _ in range(2)]
    for k, f in enumerate(frames):
        net.send_request(socks[k % 2], 5000 + k, bytes(f))
    got = {}
    for k in range(len(frames)):
        rid, status, body = net.read_reply(socks[k % 2])
        got[rid] = (status, body)
    for s in socks:
        s.close()
    for k, (i, c) in enumerate(zip(idx, clients)):
        status, body = got[5000 + k]
        assert status == 0
        assert cl.decrypt_response(p, c, body) == db[:, i].tobytes(), i
    ns.close()
    srv.close()
    eng.close()
