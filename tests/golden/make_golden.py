"""Writes the golden fixtures of cases.py with the reference engine compiled from the
reference's own source (`make -C oracle ref` -> oracle/_ref/libaltopir_ref.so).

    python tests/golden/make_golden.py

Every case is answered with each method this CPU supports (AMX, VNNI); they must agree,
and the agreed statuses and responses are stored (rejected rows stay zero)."""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path[:0] = [str(HERE.parent.parent), str(HERE.parent), str(HERE)]

import cases  # noqa: E402
import oracle_ctypes as oc  # noqa: E402


def main():
    assert oc.ref_available(), "build oracle/_ref first: make -C oracle ref"
    for name in cases.CASES:
        inp = cases.inputs(name)
        fr = cases.frames(name, inp)
        B, rl = len(fr), inp.response_len
        ref = oc.RefEngine(inp)
        got = None
        for m in oc.ref_methods():
            out = np.zeros((B, rl), np.uint8)
            st = np.full(B, -1, np.int32)
            r, _ = ref.answer(fr, out, st, m)
            if got is None:
                got = (r, st, out)
            else:
                assert r == got[0] and np.array_equal(st, got[1]) and np.array_equal(out, got[2]), name
        ref.close()
        r, st, out = got
        np.savez(HERE / f"{name}.npz", status=st, responses=out, answered=np.int64(r))
        print(f"{name}: {B} frames, {r} answered, {out.nbytes} response bytes")


if __name__ == "__main__":
    main()
