"""Client side of the wire protocol (src/bin/altopir.rs:282-308) on a local socket pair: the
byte layout of request, stop and reply frames."""
from __future__ import annotations

import socket
import struct

from altopir_public_amd import net


def test_request_and_stop_frames_have_the_reference_layout():
    a, b = socket.socketpair()
    net.send_request(a, 0x0102030405060708, b"\x09\x0a\x0b")
    net.send_stop(a)
    data = b.recv(64)
    assert data[:12] == struct.pack("<IQ", 3, 0x0102030405060708) and data[12:15] == b"\x09\x0a\x0b"
    assert data[15:27] == bytes(4) + b"\xff" * 8
    a.close()
    b.close()


def test_replies_are_parsed_in_pieces():
    a, b = socket.socketpair()
    body = bytes(range(200))
    frame = struct.pack("<IQi", len(body), 77, 0) + body + struct.pack("<IQi", 0, 78, 9)
    for k in range(0, len(frame), 7):  # arrives in small pieces
        a.sendall(frame[k:k + 7])
    assert net.read_reply(b) == (77, 0, body)
    assert net.read_reply(b) == (78, 9, b"")
    a.close()
    b.close()


def test_native_load_client_against_a_local_echo_server():
    # a minimal server of the protocol on the CPU: status 0 and the frame reversed for frames of
    # 64 bytes, status 1 otherwise; the native client's counts and checks follow its replies
    import threading

    import numpy as np

    srv = socket.socket()
    srv.bind(("127.0.0.1", 0))
    srv.listen(8)
    port = srv.getsockname()[1]
    stop = threading.Event()

    def serve(c):
        with c:
            while True:
                try:
                    h = net._recv_exact(c, 12)
                except (ConnectionError, OSError):
                    return
                ln, rid = struct.unpack("<IQ", h)
                body = net._recv_exact(c, ln) if ln else b""
                if ln == 64:
                    c.sendall(struct.pack("<IQi", 64, rid, 0) + body[::-1])
                else:
                    c.sendall(struct.pack("<IQi", 0, rid, 1))

    def accept():
        srv.settimeout(0.2)
        while not stop.is_set():
            try:
                c, _ = srv.accept()
            except OSError:
                continue
            threading.Thread(target=serve, args=(c,), daemon=True).start()

    acc = threading.Thread(target=accept, daemon=True)
    acc.start()
    rng = np.random.default_rng(3)
    frames = [rng.integers(0, 256, 64, dtype=np.uint8).tobytes() for _ in range(3)] + [b"short"]
    expect = {0: frames[0][::-1], 1: frames[1][::-1], 2: bytes(64)}  # frame 2's expectation is wrong
    r = net.native_load("127.0.0.1", port, frames, rate=2000, seconds=0.5, conns=3, seed=9, drain_s=0.3,
                        expect_status=[0, 0, 0, 1], expect=expect)
    stop.set()
    acc.join()
    srv.close()
    assert r["missing"] == 0 and r["wrong_status"] == 0 and r["requests"] > 500
    assert r["answered"] + r["rejected"] == r["requests"] and r["busy"] == 0
    assert r["checked_responses"] == r["answered"]
    # replies for frame 2 disagree with its (wrong) expectation, the others agree
    assert 0 < r["bad_responses"] < r["answered"]
