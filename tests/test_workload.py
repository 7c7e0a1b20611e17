"""The torch generators that build benchmark inputs in device memory produce the same bytes
as the host generators (run here with torch on the CPU)."""
import numpy as np
import pytest

from altopir_public_amd import synthetic, workload


@pytest.mark.parametrize("B,group", [(1, 64), (70, 64), (9, 4)])
def test_device_request_generator_matches_the_host_frames(B, group):
    rows, kp = 128, 64
    L = workload.request_len(rows, kp)
    stride = (7 + L + 255) // 256 * 256
    arena = workload.make_requests_device(B, rows, kp, seed=5, stride=stride, device="cpu", group=group).numpy()
    host = workload.make_requests(B, rows, kp, seed=5)
    for c in range(B):
        assert np.array_equal(arena[c, 7:7 + L], host[c]), c
        assert workload.frame_status(arena[c, 7:7 + L], rows, kp) == 0
    assert not arena[:, :7].any() and not arena[:, 7 + L:].any()


def test_device_database_generator_matches_the_host_bytes():
    n = (1 << 12) + 5
    dev = synthetic.uniform_bytes_device(3, n, device="cpu", chunk=100).numpy()
    assert np.array_equal(dev, synthetic.uniform_bytes(3, n))
