"""TCP front end of the serving loop (alto_net_* in include/altopir_engine.h), the client side of
the reference's wire protocol, and an open-loop load generator over TCP after the reference's
`load` command (src/bin/altopir.rs:594-739).

Frames (src/bin/altopir.rs:282-308), little-endian:
    request  [u32 len][u64 id][len bytes]
    reply    [u32 len][u64 id][i32 status][len bytes]   (status 0: the response; else len = 0)
    stop     len = 0, id = 2^64 - 1

    python -m altopir_public_amd.net --gib 1 --rate 2000 --seconds 10 --conns 16

serves one synthetic database (R = gib * 2^30 / C records of C uniform bytes, seed 0) on
127.0.0.1, sends requests on a Poisson schedule over `--conns` connections (native client
threads, alto_net_load, unless --client python) from pre-generated
frames (--frames valid ones, one corrupted per hundred of them and one longer than any request,
chosen uniformly per request), checks every reply's status against the
decoder rule and a sample of responses against direct answers of the engine, stops the server
with a stop frame and prints one JSON line.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import socket
import struct
import sys
import threading
import time

import numpy as np

from . import _lib

STOP_ID = (1 << 64) - 1
STATUS_BAD_LENGTH = 1
STATUS_BUSY = 9


class NetStats(C.Structure):
    _fields_ = [(f, C.c_longlong) for f in ("connections", "frames", "answered", "rejected", "busy", "discarded",
                                            "stop_frames", "bytes_in", "bytes_out")]


class LoadResult(C.Structure):
    _fields_ = [(f, C.c_longlong) for f in ("requests", "answered", "busy", "rejected", "missing", "wrong_status",
                                            "checked", "bad")] + \
               [(f, C.c_double) for f in ("seconds", "answered_qps", "p50_ms", "p90_ms", "p99_ms", "max_ms")]


def _bind(lib):
    if getattr(lib, "_net_bound", False):
        return lib
    P = C.c_void_p
    lib.alto_net_load.restype = C.c_int
    lib.alto_net_load.argtypes = [C.c_char_p, C.c_int, P, P, C.c_int, P, P, C.c_size_t, C.c_double, C.c_double,
                                  C.c_int, C.c_uint64, C.c_int, C.POINTER(LoadResult)]
    lib.alto_net_new.restype = P
    lib.alto_net_new.argtypes = [P, P, C.c_int, C.c_char_p, C.POINTER(C.c_int)]
    lib.alto_net_port.restype = C.c_int
    lib.alto_net_port.argtypes = [P, C.c_int]
    lib.alto_net_wait.restype = C.c_int
    lib.alto_net_wait.argtypes = [P, C.c_int]
    lib.alto_net_get_stats.restype = C.c_int
    lib.alto_net_get_stats.argtypes = [P, C.POINTER(NetStats)]
    lib.alto_net_free.restype = None
    lib.alto_net_free.argtypes = [P]
    lib._net_bound = True
    return lib


class NetServer:
    """Listens for the reference's request frames on one port per database of `server` (a
    serve.Server) and answers them through its serving loop.  While it runs, the server's
    completions belong to it: do not poll the server or replay load on it."""

    def __init__(self, server, host: str = "127.0.0.1", ports=None):
        self.lib = _bind(_lib.load())
        self.server = server
        n = len(server.engines)
        ports = list(ports) if ports is not None else [0] * n
        if len(ports) != n:
            raise ValueError("one port per database")
        self._engines = (C.c_void_p * n)(*[e.ptr for e in server.engines])
        self._ports = (C.c_int * n)(*ports)
        self.ptr = self.lib.alto_net_new(server.ptr, C.addressof(self._engines), n, host.encode(), self._ports)
        if not self.ptr:
            raise RuntimeError("alto_net_new failed (see stderr)")
        self.host = host

    def port(self, db: int = 0) -> int:
        return self.lib.alto_net_port(self.ptr, db)

    def wait(self, timeout_ms: int = -1) -> bool:
        """True once a stop frame has arrived (waits up to timeout_ms; < 0 without limit)."""
        return self.lib.alto_net_wait(self.ptr, timeout_ms) == 1

    def stats(self) -> dict:
        s = NetStats()
        self.lib.alto_net_get_stats(self.ptr, C.byref(s))
        return {f: getattr(s, f) for f, _ in NetStats._fields_}

    def close(self):
        if getattr(self, "ptr", None):
            self.lib.alto_net_free(self.ptr)
            self.ptr = None

    def __del__(self):
        self.close()


# ------------------------------------------------------------------ client side --

def connect(host: str, port: int) -> socket.socket:
    s = socket.create_connection((host, port))
    s.setsockopt(socket.IPPROTO_TCP, socket.TCP_NODELAY, 1)
    return s


def send_request(s: socket.socket, rid: int, body) -> None:
    """src/bin/altopir.rs:291-297."""
    b = bytes(body)
    s.sendall(struct.pack("<IQ", len(b), rid) + b)


def send_stop(s: socket.socket) -> None:
    s.sendall(struct.pack("<IQ", 0, STOP_ID))


def _recv_exact(s: socket.socket, n: int) -> bytes:
    buf = bytearray(n)
    view = memoryview(buf)
    got = 0
    while got < n:
        k = s.recv_into(view[got:], n - got)
        if k == 0:
            raise ConnectionError("connection closed")
        got += k
    return bytes(buf)


def read_reply(s: socket.socket) -> tuple[int, int, bytes]:
    """src/bin/altopir.rs:299-308: (id, status, body)."""
    ln, rid, status = struct.unpack("<IQi", _recv_exact(s, 16))
    return rid, status, _recv_exact(s, ln) if ln else b""


def tcp_load(host: str, port: int, frames, rate: float, seconds: float, conns: int = 16, seed: int = 1,
             drain_s: float = 10.0, expect_status=None, expect=None) -> dict:
    """Open-loop Poisson load over `conns` connections, after src/bin/altopir.rs:594-739:
    request i (frame idx[i], uniformly chosen) is due at start + sched[i], connections take
    requests round-robin, and latency runs from the due time to the reply.  `expect_status`
    (per frame) and `expect` ({frame index: response bytes}) check the replies; the summary
    counts what disagrees."""
    rng = np.random.default_rng(seed)
    sched = []
    t = 0.0
    while True:
        t += -np.log(1.0 - rng.random()) / rate
        if t >= seconds:
            break
        sched.append(t)
    n = len(sched)
    idx = rng.integers(0, len(frames), size=n)
    bodies = [bytes(f) for f in frames]
    socks = [connect(host, port) for _ in range(conns)]
    results = [None] * n
    start = time.monotonic() + 0.2

    def sender(c):
        s = socks[c]
        for i in range(c, n, conns):
            d = start + sched[i] - time.monotonic()
            if d > 0:
                time.sleep(d)
            try:
                send_request(s, i, bodies[idx[i]])
            except OSError:
                return

    def receiver(c):
        s = socks[c]
        mine = len(range(c, n, conns))
        s.settimeout(seconds + drain_s + 5)
        for _ in range(mine):
            try:
                rid, status, body = read_reply(s)
            except (OSError, ConnectionError):
                return
            results[rid] = (status, time.monotonic(), body)

    ts = [threading.Thread(target=f, args=(c,)) for c in range(conns) for f in (sender, receiver)]
    for th in ts:
        th.start()
    for th in ts:
        th.join()
    for s in socks:
        s.close()
    got = [i for i in range(n) if results[i] is not None]
    answered = [i for i in got if results[i][0] == 0]
    busy = sum(1 for i in got if results[i][0] == STATUS_BUSY)
    lat = np.sort(np.array([(results[i][1] - start - sched[i]) * 1e3 for i in answered]))
    wrong_status = 0
    if expect_status is not None:
        wrong_status = sum(1 for i in got if results[i][0] not in (expect_status[idx[i]], STATUS_BUSY))
    bad, checked = 0, 0
    if expect:
        for i in answered:
            want = expect.get(int(idx[i]))
            if want is not None:
                checked += 1
                bad += results[i][2] != want

    def pct(p):
        return round(float(lat[min(int(len(lat) * p), len(lat) - 1)]), 3) if len(lat) else 0.0

    return {"requests": n, "answered": len(answered), "busy": busy, "rejected": len(got) - len(answered) - busy,
            "missing": n - len(got), "answered_qps": round(len(answered) / seconds, 2), "conns": conns,
            "rate": rate, "seconds": seconds,
            "latency_ms": {"p50": pct(0.5), "p90": pct(0.9), "p99": pct(0.99), "max": round(float(lat[-1]), 3)
                           if len(lat) else 0.0},
            "wrong_status": wrong_status, "checked_responses": checked, "bad_responses": bad}


def native_load(host: str, port: int, frames, rate: float, seconds: float, conns: int = 16, seed: int = 1,
                drain_s: float = 10.0, expect_status=None, expect=None) -> dict:
    """The same open-loop load from native sender/receiver threads (alto_net_load), so that the
    client does not bound the rate; `expect` maps frame index -> response bytes."""
    lib = _bind(_lib.load())
    arrs = [np.frombuffer(bytes(f), np.uint8) for f in frames]
    keep = [a if len(a) else np.zeros(1, np.uint8) for a in arrs]
    fp = (C.c_void_p * len(arrs))(*[k.ctypes.data for k in keep])
    ln = (C.c_size_t * len(arrs))(*[len(a) for a in arrs])
    es = (C.c_int * len(arrs))(*expect_status) if expect_status is not None else None
    ex, elen, hold = None, 0, []
    if expect:
        elen = len(next(iter(expect.values())))
        ex = (C.c_void_p * len(arrs))()
        for i, r in expect.items():
            b = np.frombuffer(r, np.uint8)
            assert len(b) == elen
            hold.append(b)
            ex[i] = b.ctypes.data
    r = LoadResult()
    rc = lib.alto_net_load(host.encode(), port, C.addressof(fp), C.addressof(ln), len(arrs),
                           C.addressof(es) if es is not None else None, C.addressof(ex) if ex is not None else None,
                           elen, rate, seconds, conns, seed, int(drain_s * 1000), C.byref(r))
    if rc != 0:
        raise RuntimeError("alto_net_load: bad arguments or no connection")
    d = {f: getattr(r, f) for f, _ in LoadResult._fields_}
    return {"requests": d["requests"], "answered": d["answered"], "busy": d["busy"], "rejected": d["rejected"],
            "missing": d["missing"], "answered_qps": round(d["answered_qps"], 2), "conns": conns, "rate": rate,
            "seconds": seconds, "latency_ms": {k: round(d[k + "_ms"], 3) for k in ("p50", "p90", "p99", "max")},
            "wrong_status": d["wrong_status"], "checked_responses": d["checked"], "bad_responses": d["bad"],
            "client": "native"}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gib", type=float, default=1.0)
    ap.add_argument("--record-bytes", type=int, default=32768)
    ap.add_argument("--rate", type=float, default=2000.0)
    ap.add_argument("--seconds", type=float, default=10.0)
    ap.add_argument("--conns", type=int, default=16)
    ap.add_argument("--frames", type=int, default=256, help="distinct valid frames sent")
    ap.add_argument("--max-batch", type=int, default=128)
    ap.add_argument("--slots", type=int, default=512, help="request slots (the reference's --queue)")
    ap.add_argument("--contexts", type=int, default=2)
    ap.add_argument("--check", type=int, default=16, help="frames whose responses are compared")
    ap.add_argument("--client", choices=["native", "python"], default="native",
                    help="load client: native threads (alto_net_load) or Python threads")
    a = ap.parse_args(argv)

    from . import workload
    from .loadgen import build_engines, frame_pool
    from .serve import Server

    (eng,), (inp,) = build_engines([a.gib], a.record_bytes, 2048, 4, 16)
    frames = frame_pool(inp, a.frames, seed=5)
    rng = np.random.default_rng(6)
    L = inp.request_len
    # one corrupted frame per hundred valid ones (statuses 1-3 by the decoder rule), and one
    # longer than any request
    extra = []
    for i in range(max(1, a.frames // 100)):
        extra.append(np.frombuffer(workload.corrupt(frames[i], rng, inp.kp), np.uint8))
    extra.append(np.zeros(L + 100, np.uint8))
    allf = frames + extra
    statuses = [workload.frame_status(f, inp.rows, inp.kp) for f in allf]
    expect = {}
    for i in range(min(a.check, a.frames)):
        out = np.zeros((1, inp.response_len), np.uint8)
        st = np.full(1, -1, np.int32)
        eng.answer([frames[i]], out, st)
        expect[i] = out[0].tobytes()
    srv = Server([eng], max_batch=a.max_batch, slots=a.slots, contexts=a.contexts)
    net = NetServer(srv)
    port = net.port(0)
    print(f"[net] listening on 127.0.0.1:{port}", file=sys.stderr, flush=True)
    run = native_load if a.client == "native" else tcp_load
    r = run("127.0.0.1", port, allf, a.rate, a.seconds, conns=a.conns, drain_s=3.0, expect_status=statuses,
            expect=expect)
    s = connect("127.0.0.1", port)
    send_stop(s)
    stopped = net.wait(5000)
    s.close()
    r.update(gib=a.gib, records=inp.rows, record_bytes=a.record_bytes, stopped_by_frame=stopped,
             server=net.stats())
    net.close()
    srv.close()
    eng.close()
    print(json.dumps(r), flush=True)
    return 0 if (r["wrong_status"] == 0 and r["bad_responses"] == 0 and r["missing"] == 0) else 1


if __name__ == "__main__":
    sys.exit(main())
