"""The committed golden fixtures (tests/golden, produced by the reference engine compiled
from its own source) against the oracle (CPU) and against the HIP engine (GPU)."""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
import pytest

sys.path.insert(0, str(Path(__file__).resolve().parent / "golden"))
import cases  # noqa: E402
import oracle_ctypes as oc  # noqa: E402

NAMES = list(cases.CASES)


@pytest.mark.parametrize("name", NAMES)
def test_oracle_reproduces_the_golden_responses(name):
    st_want, out_want, answered = cases.load(name)
    inp = cases.inputs(name)
    fr = cases.frames(name, inp)
    out = np.zeros_like(out_want)
    st = np.full(len(fr), -1, np.int32)
    assert oc.oracle_answer(inp, fr, out, st) == answered
    assert np.array_equal(st, st_want)
    assert np.array_equal(out, out_want)


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_hip_engine_reproduces_the_golden_responses(hip_device, name):
    from altopir_public_amd.engine import Engine, Method

    st_want, out_want, answered = cases.load(name)
    inp = cases.inputs(name)
    fr = cases.frames(name, inp)
    eng = Engine.from_inputs(inp)
    out = np.zeros_like(out_want)
    st = np.full(len(fr), -1, np.int32)
    assert eng.answer(fr, out, st, Method.HIP)[0] == answered
    eng.close()
    assert np.array_equal(st, st_want)
    bad = [i for i in range(len(fr)) if not np.array_equal(out[i], out_want[i])]
    assert not bad, f"responses {bad[:8]} differ from the golden bytes"
