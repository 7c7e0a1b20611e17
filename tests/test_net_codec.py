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
