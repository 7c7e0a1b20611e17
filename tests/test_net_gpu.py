"""The TCP front end of the serving loop speaks the reference server's wire protocol
(src/bin/altopir.rs:282-308, 403-442): every reply carries its request's id, the status the
decoder gives the frame (1 for a frame longer than any request, discarded without losing the
stream; 9 when the slot pool is full) and, for a valid frame, the oracle's response bytes."""
from __future__ import annotations

import threading
import time

import numpy as np
import pytest

from altopir_public_amd import net, workload
from altopir_public_amd.engine import Engine
from altopir_public_amd.serve import Server
import oracle_ctypes as oc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dbs():
    a = workload.make_inputs(2048, 4096, n=2048, tg=4, seed=21)
    b = workload.make_inputs(4096, 2048, n=2048, tg=4, seed=22, prefix="random")
    return a, b


def oracle_responses(inp, frames):
    out = np.zeros((len(frames), inp.response_len), np.uint8)
    oc.oracle_answer(inp, frames, out, np.full(len(frames), -1, np.int32))
    return out


def settled_stats(ns, key, value, timeout_s=5.0):
    # the front end counts a reply once its write has returned, which can be after the client
    # has read it
    deadline = time.monotonic() + timeout_s
    st = ns.stats()
    while st[key] != value and time.monotonic() < deadline:
        time.sleep(0.005)
        st = ns.stats()
    return st


def read_all(s, count):
    got = {}
    for _ in range(count):
        rid, status, body = net.read_reply(s)
        assert rid not in got, f"two replies for id {rid}"
        got[rid] = (status, body)
    return got


def test_frames_over_tcp_get_the_oracle_answer_and_the_decoder_status(hip_device, dbs):
    inp = dbs[0]
    eng = Engine.from_inputs(inp)
    srv = Server([eng], max_batch=64, slots=64, contexts=2)
    ns = net.NetServer(srv)
    valid = workload.make_requests(4, inp.rows, inp.kp, seed=31)
    rng = np.random.default_rng(32)
    L = inp.request_len
    frames = {
        10: valid[0], 11: valid[1],
        12: np.frombuffer(workload.corrupt(valid[2], rng, inp.kp), np.uint8),
        13: valid[2][:-1],                    # one byte short
        14: np.zeros(L + 1000, np.uint8),     # longer than any request: discarded
        15: np.zeros(0, np.uint8),            # empty frame
        (1 << 63) + 5: valid[3],              # ids are opaque 64-bit values
    }
    want = dict(zip((10, 11, (1 << 63) + 5), oracle_responses(inp, [valid[0], valid[1], valid[3]])))
    s = net.connect("127.0.0.1", ns.port(0))
    for rid, f in frames.items():
        net.send_request(s, rid, f)
    got = read_all(s, len(frames))
    s.close()
    for rid, f in frames.items():
        status, body = got[rid]
        assert status == workload.frame_status(f, inp.rows, inp.kp), f"id {rid}"
        if status == 0:
            assert body == want[rid].tobytes(), f"id {rid}: response differs from the oracle"
        else:
            assert body == b""
    st = settled_stats(ns, "answered", 3)
    assert st["frames"] == len(frames) and st["discarded"] == 1 and st["answered"] == 3
    ns.close()
    srv.close()
    eng.close()


def test_a_full_slot_pool_answers_busy_and_every_frame_gets_a_reply(hip_device, dbs):
    inp = dbs[0]
    eng = Engine.from_inputs(inp)
    srv = Server([eng], max_batch=8, slots=4, contexts=1)
    ns = net.NetServer(srv)
    frames = workload.make_requests(8, inp.rows, inp.kp, seed=41)
    want = oracle_responses(inp, frames)
    s = net.connect("127.0.0.1", ns.port(0))
    # 96 frames in one burst through 4 slots: some must come back busy (status 9)
    payload = b"".join(
        np.array([len(frames[i % 8])], "<u4").tobytes() + np.array([i], "<u8").tobytes() + frames[i % 8].tobytes()
        for i in range(96))
    s.sendall(payload)
    got = read_all(s, 96)
    s.close()
    statuses = [got[i][0] for i in range(96)]
    assert set(statuses) <= {0, 9}
    assert statuses.count(9) > 0 and statuses.count(0) > 0
    for i in range(96):
        if got[i][0] == 0:
            assert got[i][1] == want[i % 8].tobytes(), f"request {i}"
    assert ns.stats()["busy"] == statuses.count(9)
    ns.close()
    srv.close()
    eng.close()


def test_two_databases_on_two_ports_and_a_stop_frame(hip_device, dbs):
    engs = [Engine.from_inputs(i) for i in dbs]
    srv = Server(engs, max_batch=32, slots=64, contexts=2)
    ns = net.NetServer(srv)
    assert ns.port(0) != ns.port(1)
    frames = [workload.make_requests(12, i.rows, i.kp, seed=50 + k) for k, i in enumerate(dbs)]
    want = [oracle_responses(i, f) for i, f in zip(dbs, frames)]
    errors = []

    def client(db, c):
        try:
            s = net.connect("127.0.0.1", ns.port(db))
            ids = list(range(c * 100, c * 100 + 12))
            for k, rid in enumerate(ids):
                net.send_request(s, rid, frames[db][k])
            got = read_all(s, 12)
            s.close()
            for k, rid in enumerate(ids):
                if got[rid] != (0, want[db][k].tobytes()):
                    errors.append((db, c, rid, got[rid][0]))
        except Exception as e:  # noqa: BLE001 - reported below
            errors.append((db, c, repr(e)))

    ts = [threading.Thread(target=client, args=(c % 2, c)) for c in range(6)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    assert not errors, errors[:4]
    assert not ns.wait(0)
    s = net.connect("127.0.0.1", ns.port(1))
    net.send_stop(s)
    assert ns.wait(5000), "the stop frame was not seen"
    s.close()
    st = settled_stats(ns, "answered", 72)
    assert st["connections"] == 7 and st["answered"] == 72 and st["stop_frames"] == 1
    ns.close()
    srv.close()
    for e in engs:
        e.close()


@pytest.mark.parametrize("client", ["python", "native"])
def test_poisson_load_over_tcp(hip_device, dbs, client):
    inp = dbs[0]
    eng = Engine.from_inputs(inp)
    srv = Server([eng], max_batch=64, slots=256, contexts=2)
    ns = net.NetServer(srv)
    frames = workload.make_requests(16, inp.rows, inp.kp, seed=61)
    rng = np.random.default_rng(62)
    frames.append(np.frombuffer(workload.corrupt(frames[0], rng, inp.kp), np.uint8))
    frames.append(np.zeros(inp.request_len + 7, np.uint8))
    statuses = [workload.frame_status(f, inp.rows, inp.kp) for f in frames]
    want = {i: r.tobytes() for i, r in enumerate(oracle_responses(inp, frames[:16]))}
    run = net.tcp_load if client == "python" else net.native_load
    r = run("127.0.0.1", ns.port(0), frames, rate=400 if client == "python" else 4000, seconds=2.0, conns=4, seed=63,
            drain_s=3, expect_status=statuses, expect=want)
    assert r["missing"] == 0 and r["wrong_status"] == 0 and r["bad_responses"] == 0, r
    assert r["answered"] > 0 and r["checked_responses"] == r["answered"]
    ns.close()
    srv.close()
    eng.close()
