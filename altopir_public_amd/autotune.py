"""Tile-shape tuner for the two contractions of the answer engine.

For a resident engine and a batch size, every kernel variant of the scan and of the
packing contraction (alto_variant_name) is timed on a device-resident batch with the
engine's own HIP events (alto_last_timings); the fastest is recorded in the engine's
tuning table for that batch-size bucket (alto_tune_set).  All variants compute identical
bytes, so tuning changes only speed.  The table can be saved to and loaded from JSON,
keyed by (records, record bytes) and bucket.
"""
from __future__ import annotations

import json
from pathlib import Path

import numpy as np

from . import _lib, workload

SCAN, PACK = 0, 1


def variant_names() -> list[str]:
    lib = _lib.load()
    return [lib.alto_variant_name(i).decode() for i in range(lib.alto_variant_count())]


def bucket_top(b: int) -> int:
    """Largest A-tile count (4 queries a tile) of tuning bucket b: 1, 2, 3, 4, 6, 8, 12, ..."""
    if b < 2:
        return b + 1
    return 2 << (b >> 1) if b & 1 else 3 << ((b >> 1) - 1)


def bucket(B: int) -> int:
    """Tuning bucket of a batch of B queries (the engine's bucket_of on ceil(B / 4))."""
    mt = (B + 3) // 4
    b = 0
    while b + 1 < 30 and bucket_top(b) < mt:
        b += 1
    return b


class DeviceBatch:
    """B seeded request frames resident in device memory (torch tensors), in the layout of
    the device entry point: frame c at slots + c * stride + 7."""

    def __init__(self, engine, inp, B: int, seed: int, stream=None):
        import torch

        self.B = B
        self.L = inp.request_len
        self.rl = inp.response_len
        self.stride = (7 + self.L + 255) // 256 * 256
        self.slots = workload.make_requests_device(B, inp.rows, inp.kp, seed=seed, stride=self.stride)
        self.status = torch.zeros(B, dtype=torch.int32, device="cuda")
        self.out = torch.empty((B, self.rl), dtype=torch.uint8, device="cuda")
        self.engine = engine
        self.stream = stream

    def run(self) -> None:
        import torch

        s = self.stream if self.stream is not None else torch.cuda.current_stream().cuda_stream
        rc = self.engine.answer_device(self.slots.data_ptr(), self.stride, self.B, self.out.data_ptr(), self.rl,
                                       self.status.data_ptr(), s)
        if rc != 0:
            raise RuntimeError(f"alto_answer_device returned {rc} ({_lib.last_error()})")


def time_variants(engine, batch: DeviceBatch, which: int, reps: int = 0) -> dict:
    """Mean device time (ms) of the chosen contraction for every variant, each timed over
    back-to-back batches (no host wait between them, as in a serving or benchmark loop:
    the clocks a contraction sustains under load differ from those of an isolated launch).
    reps = 0 picks the count per variant so that about 50 ms of batches run (3 to 20)."""
    lib = _lib.load()
    stage = 2 if which == SCAN else 4  # index in the stage times
    out = {}
    graphs = engine.graphs
    engine.set_graphs(False)  # eager batches record the stage events the times come from
    fusion = lib.alto_set_fusion(engine.ptr, 0)  # each contraction timed in a launch of its own
    saved = lib.alto_tune_get(engine.ptr, which, batch.B)
    mt = (batch.B + 3) // 4
    for idx, name in enumerate(variant_names()):
        if name.startswith("stream") and mt > 16:
            continue  # register-streaming shapes re-read the database once per 4..32 queries
        lib.alto_tune_set(engine.ptr, which, batch.B, idx)
        batch.run()  # first use: instantiation, workspace
        engine.stage_totals(reset=True)
        batch.run()
        n1, t1 = engine.stage_totals(reset=True)
        step_ms = max(sum(t1) / max(n1, 1), 1e-3)
        k = reps if reps > 0 else int(min(20, max(3, round(50.0 / step_ms))))
        for _ in range(k):
            batch.run()
        n, sums = engine.stage_totals(reset=True)
        out[name] = sums[stage] / n
    lib.alto_tune_set(engine.ptr, which, batch.B, saved)
    lib.alto_set_fusion(engine.ptr, fusion)
    engine.set_graphs(graphs)
    return out


def tune(engine, inp, batch_sizes, seed: int = 7, reps: int = 0, log=None) -> dict:
    """Times every variant for each batch size and records the fastest in the engine.
    Returns {bucket: {"scan": (name, ms), "pack": (name, ms), "B": B}}."""
    lib = _lib.load()
    names = variant_names()
    table = {}
    for B in batch_sizes:
        batch = DeviceBatch(engine, inp, B, seed)
        entry = {"B": B}
        for which, key in ((SCAN, "scan"), (PACK, "pack")):
            ms = time_variants(engine, batch, which, reps)
            best = min(ms, key=ms.get)
            lib.alto_tune_set(engine.ptr, which, B, names.index(best))
            entry[key] = (best, ms[best])
            entry[key + "_all"] = ms
            if log:
                log(f"[tune] B={B} {key}: " + ", ".join(f"{k}={v:.3f}" for k, v in sorted(ms.items(), key=lambda x: x[1])))
        table[bucket(B)] = entry
    return table


def save(table: dict, path: Path, rows: int, cols: int) -> None:
    data = {}
    if path.exists():
        data = json.loads(path.read_text())
    data[f"{rows}x{cols}"] = {str(k): {"scan": v["scan"][0], "pack": v["pack"][0], "B": v["B"]} for k, v in table.items()}
    path.write_text(json.dumps(data, indent=1, sort_keys=True))


def load(engine, path: Path, rows: int, cols: int) -> int:
    """Applies a saved table to an engine; returns the number of entries applied."""
    lib = _lib.load()
    if not path.exists():
        return 0
    names = variant_names()
    ent = json.loads(path.read_text()).get(f"{rows}x{cols}", {})
    for _, v in ent.items():
        for which, key in ((SCAN, "scan"), (PACK, "pack")):
            if v[key] in names:
                lib.alto_tune_set(engine.ptr, which, int(v["B"]), names.index(v[key]))
    return len(ent)
