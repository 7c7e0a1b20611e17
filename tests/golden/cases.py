"""Golden cases: seeded engine inputs and frames whose expected responses were produced by
the reference engine compiled from its own source (oracle/_ref), stored next to this file
as <name>.npz (statuses + responses).  Regenerate with make_golden.py."""
from __future__ import annotations

from pathlib import Path

import numpy as np

from altopir_public_amd import workload

HERE = Path(__file__).resolve().parent

CASES = {
    # tests/engine.rs:10-11 shape: 2048 records x 4096 B, n = 2048, five valid requests
    "rows2048_c4096": dict(rows=2048, cols=4096, n=2048, tg=4, seed=1, ld=2048, prefix="row0", table="ntt",
                           frames=("valid", 5, 21)),
    # a small ring with random prefixes, the eight hand-made frames of tests/engine.rs:130-148
    "rows2048_n256_hand": dict(rows=2048, cols=512, n=256, tg=4, seed=5, ld=2048, prefix="random", table="ntt",
                               frames=("hand", 2, 22)),
    # the smallest legal shapes, arbitrary twiddles, ld > rows
    "rows64_n16": dict(rows=64, cols=32, n=16, tg=8, seed=6, ld=96, prefix="random", table="random",
                       frames=("valid", 3, 23)),
    # the mutate fuzz model of tests/engine.rs:29-68 (two batches of 64 frames)
    "rows128_n32_fuzz": dict(rows=128, cols=64, n=32, tg=4, seed=7, ld=128, prefix="random", table="ntt",
                             frames=("fuzz", 4, 24)),
    # tests/engine.rs:196-245 shape: 8192 records
    "rows8192_c4096": dict(rows=8192, cols=4096, n=2048, tg=4, seed=2, ld=8192, prefix="row0", table="ntt",
                           frames=("valid", 2, 25)),
}


def inputs(name: str):
    c = CASES[name]
    return workload.make_inputs(c["rows"], c["cols"], n=c["n"], tg=c["tg"], seed=c["seed"], ld=c["ld"],
                                prefix=c["prefix"], table=c["table"])


def frames(name: str, inp) -> list:
    import refmodel

    kind, count, seed = CASES[name]["frames"]
    base = workload.make_requests(count, inp.rows, inp.kp, seed=seed)
    if kind == "valid":
        return base
    if kind == "hand":
        return refmodel.hand_frames(base[0], base[1], inp.kp)[0]
    out = []
    for b in refmodel.fuzz_batches(base, inp.kp, batches=2):
        out.extend(b)
    return out


def load(name: str):
    z = np.load(HERE / f"{name}.npz", allow_pickle=False)
    return z["status"], z["responses"], int(z["answered"])
