"""The allocation-free device entry point (alto_answer_device_noalloc) against the device
entry point it mirrors (alto_answer_device), through the C ABI: byte-identical responses
and statuses on two calls in a row, the second with a smaller batch, and a clear refusal of
a batch whose workspace exceeds the fixed scratch array."""
import ctypes as C

import numpy as np
import pytest

from altopir_public_amd import _lib, synthetic, workload
from altopir_public_amd.engine import Engine
import oracle_ctypes as oc

pytestmark = pytest.mark.gpu

SCRATCH = 64 << 20  # ALTO_NOALLOC_SCRATCH_BYTES


def workspace_bytes(B, rows, kp, cols):
    r = lambda x: (x + 255) // 256 * 256  # noqa: E731
    Bp = (B + 3) // 4 * 4
    return r(4 * Bp * rows) + r(4 * Bp * kp) + 2 * r(4 * Bp * cols) + r(4 * Bp) + r(8 * Bp)


def device_batch(inp, B, seed, bad=()):
    """B device-generated frames (slot layout of the device entry point); frames in `bad`
    get a key word equal to Q (status 3)."""
    import torch

    L = inp.request_len
    stride = (7 + L + 255) // 256 * 256
    slots = workload.make_requests_device(B, inp.rows, inp.kp, seed=seed, stride=stride)
    for c in bad:
        w = 7 + 9 + 4 * (c * 97 % (2 * inp.kp))
        slots[c, w:w + 4] = torch.tensor(list(int(synthetic.Q).to_bytes(4, "little")), dtype=torch.uint8,
                                         device=slots.device)
    return slots, stride


def run(lib, fn, eng, slots, stride, B, rl, stream):
    import torch

    out = torch.full((B, rl), 0xA5, dtype=torch.uint8, device="cuda")
    st = torch.zeros(B, dtype=torch.int32, device="cuda")
    rc = getattr(lib, fn)(eng.ptr, slots.data_ptr(), stride, B, out.data_ptr(), rl, st.data_ptr(), stream)
    assert rc == 0, (fn, rc, _lib.last_error())
    torch.cuda.synchronize()
    return out.cpu().numpy(), st.cpu().numpy()


def test_noalloc_matches_the_device_entry_point(hip_device):
    import torch

    lib = _lib.load()
    inp = workload.make_inputs(2048, 4096, n=2048, tg=4, seed=71, prefix="random")
    eng = Engine.from_inputs(inp)
    rl = inp.response_len
    stream = torch.cuda.current_stream().cuda_stream
    for B, seed, bad in ((45, 81, (3, 44)), (6, 82, (1,))):
        slots, stride = device_batch(inp, B, seed, bad)
        torch.cuda.synchronize()
        want_out, want_st = run(lib, "alto_answer_device", eng, slots, stride, B, rl, stream)
        got_out, got_st = run(lib, "alto_answer_device_noalloc", eng, slots, stride, B, rl, stream)
        exp = [3 if c in bad else 0 for c in range(B)]
        assert list(want_st) == exp and list(got_st) == exp
        assert np.array_equal(got_out, want_out), f"batch {B}: responses differ from alto_answer_device"
        assert (got_out[list(bad)] == 0xA5).all(), "a rejected request's buffer was written"
        if B == 6:
            # and against the oracle, for the smaller batch
            frames = [slots[c, 7:7 + inp.request_len].cpu().numpy() for c in range(B)]
            out_o = np.zeros((B, rl), np.uint8)
            st_o = np.full(B, -1, np.int32)
            oc.oracle_answer(inp, frames, out_o, st_o)
            assert list(st_o) == exp
            for c in range(B):
                if exp[c] == 0:
                    assert np.array_equal(got_out[c], out_o[c])
    eng.close()


def test_noalloc_refuses_a_batch_beyond_the_scratch_array(hip_device):
    import torch

    lib = _lib.load()
    inp = workload.make_inputs(2048, 4096, n=2048, tg=4, seed=72)
    eng = Engine.from_inputs(inp)
    limit = 1
    while workspace_bytes(limit + 1, inp.rows, inp.kp, inp.cols) <= SCRATCH:
        limit += 1
    B = (limit // 4 + 1) * 4 + 1  # the next multiple of 4 plus one: over the limit
    assert workspace_bytes(B, inp.rows, inp.kp, inp.cols) > SCRATCH
    L, rl = inp.request_len, inp.response_len
    stride = (7 + L + 255) // 256 * 256
    slots = torch.zeros((B, stride), dtype=torch.uint8, device="cuda")
    out = torch.zeros((B, rl), dtype=torch.uint8, device="cuda")
    st = torch.zeros(B, dtype=torch.int32, device="cuda")
    rc = lib.alto_answer_device_noalloc(eng.ptr, slots.data_ptr(), stride, B, out.data_ptr(), rl, st.data_ptr(), None)
    assert rc == -1 and "fixed scratch array" in _lib.last_error()
    torch.cuda.synchronize()
    assert not out.any() and not st.any(), "a refused batch launched work"
    # the largest batch that fits runs
    slots2, stride2 = device_batch(inp, limit, 83)
    torch.cuda.synchronize()
    got, st2 = run(lib, "alto_answer_device_noalloc", eng, slots2, stride2, limit, rl, None)
    want, st3 = run(lib, "alto_answer_device", eng, slots2, stride2, limit, rl, None)
    assert (st2 == 0).all() and (st3 == 0).all() and np.array_equal(got, want)
    eng.close()
