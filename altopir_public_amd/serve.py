"""Python handle of the device-resident serving loop (alto_server_* in include/altopir_engine.h)
and a Poisson load generator modelled on the reference's `load` command
(src/bin/altopir.rs:594-739): requests are pre-generated, sent on a seeded exponential
schedule, and latency runs from each request's scheduled send time to its completion.
"""
from __future__ import annotations

import ctypes as C
import threading
import time
from dataclasses import dataclass, field

import numpy as np

from . import _lib


class Completion(C.Structure):
    _fields_ = [
        ("tag", C.c_uint64), ("db", C.c_int), ("status", C.c_int), ("slot", C.c_int),
        ("response", C.c_void_p), ("response_len", C.c_size_t), ("t_submit_ns", C.c_int64),
        ("t_done_ns", C.c_int64), ("batch", C.c_int),
    ]


class Stats(C.Structure):
    _fields_ = [("submitted", C.c_longlong), ("queue_full", C.c_longlong), ("batches", C.c_longlong),
                ("answered_in_batches", C.c_longlong), ("max_batch_seen", C.c_longlong),
                ("contexts_admitted", C.c_longlong), ("contexts_refused", C.c_longlong)]


class ReplayResult(C.Structure):
    _fields_ = [("answered", C.c_longlong), ("rejected", C.c_longlong), ("batches", C.c_longlong),
                ("wall_s", C.c_double), ("qps", C.c_double), ("mean_batch", C.c_double), ("p50_ms", C.c_double),
                ("p90_ms", C.c_double), ("p99_ms", C.c_double), ("max_ms", C.c_double)]


def _bind(lib):
    if getattr(lib, "_serve_bound", False):
        return lib
    P = C.c_void_p
    lib.alto_server_new.restype = P
    lib.alto_server_new.argtypes = [P, C.c_int, C.c_int, C.c_int, C.c_int]
    lib.alto_server_submit.restype = C.c_int
    lib.alto_server_submit.argtypes = [P, C.c_int, P, C.c_size_t, C.c_uint64]
    lib.alto_server_poll.restype = C.c_int
    lib.alto_server_poll.argtypes = [P, C.POINTER(Completion), C.c_int, C.c_int]
    lib.alto_server_release.restype = C.c_int
    lib.alto_server_release.argtypes = [P, C.c_int, C.c_int]
    lib.alto_server_contexts.restype = C.c_int
    lib.alto_server_contexts.argtypes = [P, C.c_int]
    lib.alto_server_get_stats.restype = C.c_int
    lib.alto_server_get_stats.argtypes = [P, C.POINTER(Stats)]
    lib.alto_server_error.restype = C.c_char_p
    lib.alto_server_error.argtypes = [P]
    lib.alto_server_free.restype = None
    lib.alto_server_free.argtypes = [P]
    lib.alto_server_set_upload.restype = C.c_int
    lib.alto_server_set_upload.argtypes = [P, C.c_int]
    lib.alto_server_upload_mode.restype = C.c_int
    lib.alto_server_upload_mode.argtypes = [P, C.c_int]
    lib.alto_server_set_graphs.restype = C.c_int
    lib.alto_server_set_graphs.argtypes = [P, C.c_int]
    lib.alto_server_replay.restype = C.c_int
    lib.alto_server_replay.argtypes = [P, C.c_int, P, P, C.c_int, C.c_longlong, C.c_int, C.c_int,
                                       C.POINTER(ReplayResult)]
    lib._serve_bound = True
    return lib


@dataclass
class Done:
    tag: int
    db: int
    status: int
    response: bytes | None
    t_submit_ns: int
    t_done_ns: int
    batch: int
    slot: int = -1


class Server:
    """Continuous-batching request loop over resident engines (one per database)."""

    def __init__(self, engines, max_batch: int = 128, slots: int = 512, contexts: int = 2):
        self.lib = _bind(_lib.load())
        self.engines = list(engines)
        arr = (C.c_void_p * len(self.engines))(*[e.ptr for e in self.engines])
        self.ptr = self.lib.alto_server_new(arr, len(self.engines), max_batch, slots, contexts)
        if not self.ptr:
            raise RuntimeError("alto_server_new failed (see stderr)")
        self._buf = (Completion * 512)()

    def set_graphs(self, mode: int) -> None:
        """1: capture and replay each distinct batch, 0: eager launches, -1: follow the engines."""
        if self.lib.alto_server_set_graphs(self.ptr, mode) != 0:
            raise RuntimeError("alto_server_set_graphs failed")

    def set_upload(self, mode: int) -> None:
        """0: copy frames into HBM as they arrive (an upload thread per database);
        1: kernels read the pinned slots over PCIe;
        2: copy a batch's frames into HBM when it is dispatched (a few large copies);
        3: as 2, and write the responses to HBM and copy them back per batch;
        -1: 0 (the default)."""
        if self.lib.alto_server_set_upload(self.ptr, mode) != 0:
            raise RuntimeError("alto_server_set_upload failed (requests pending or bad mode)")

    def upload_mode(self, db: int = 0) -> int:
        """The upload mode in force for database db (0-3)."""
        return self.lib.alto_server_upload_mode(self.ptr, db)

    def contexts(self, db: int = 0) -> int:
        return self.lib.alto_server_contexts(self.ptr, db)

    def submit(self, db: int, frame, tag: int) -> int:
        a = np.frombuffer(frame, np.uint8) if isinstance(frame, (bytes, bytearray)) else np.ascontiguousarray(frame)
        return self.lib.alto_server_submit(self.ptr, db, a.ctypes.data if len(a) else None, len(a), tag)

    def poll(self, timeout_ms: int = 10, copy: bool = True, release: bool = True) -> list[Done]:
        """Completions so far (waits up to timeout_ms for the first).  A valid request's
        response is copied out (copy=True) and its slot released (release=True; otherwise
        the caller releases Done.slot with release())."""
        k = self.lib.alto_server_poll(self.ptr, self._buf, len(self._buf), timeout_ms)
        out = []
        for i in range(max(k, 0)):
            c = self._buf[i]
            resp = None
            if c.status == 0:
                if copy:
                    resp = C.string_at(c.response, c.response_len)
                if release:
                    self.lib.alto_server_release(self.ptr, c.db, c.slot)
            out.append(Done(c.tag, c.db, c.status, resp, c.t_submit_ns, c.t_done_ns, c.batch, c.slot))
        return out

    def release(self, db: int, slot: int) -> int:
        """Returns a leased response slot (alto_server_release): 0, or -1 for a slot that is
        not leased (never handed out with a response, or released already)."""
        return self.lib.alto_server_release(self.ptr, db, slot)

    def replay(self, frames, db: int = 0, total: int = 1024, threads: int = 8, window: int = 384) -> dict:
        """Native load driver (alto_server_replay): `threads` submitter threads keep up to
        `window` requests of database `db` outstanding (frames round-robin) until `total` are
        done; returns the answered rate and latency percentiles."""
        arrs = [np.frombuffer(f, np.uint8) if isinstance(f, (bytes, bytearray)) else np.ascontiguousarray(f)
                for f in frames]
        fp = (C.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])
        ln = (C.c_size_t * len(arrs))(*[len(a) for a in arrs])
        r = ReplayResult()
        rc = self.lib.alto_server_replay(self.ptr, db, C.addressof(fp), C.addressof(ln), len(arrs), total, threads,
                                         window, C.byref(r))
        if rc != 0:
            raise RuntimeError("alto_server_replay: bad arguments")
        out = {f: getattr(r, f) for f, _ in ReplayResult._fields_}
        for k in ("wall_s", "qps", "mean_batch", "p50_ms", "p90_ms", "p99_ms", "max_ms"):
            out[k] = round(out[k], 4)
        out.update(requests=total, threads=threads, window=window)
        return out

    def error(self) -> str:
        """The last error the serving loop recorded (alto_server_error), '' if none."""
        return (self.lib.alto_server_error(self.ptr) or b"").decode(errors="replace")

    def stats(self) -> dict:
        s = Stats()
        self.lib.alto_server_get_stats(self.ptr, C.byref(s))
        return {f: getattr(s, f) for f, _ in Stats._fields_}

    def close(self):
        if getattr(self, "ptr", None):
            self.lib.alto_server_free(self.ptr)
            self.ptr = None

    def __del__(self):
        self.close()


def now_ns() -> int:
    return time.monotonic_ns()


@dataclass
class LoadResult:
    sent: int
    answered: int
    rejected: dict
    wall_s: float
    latencies_ms: np.ndarray = field(repr=False)
    batches: int = 0
    max_batch: int = 0

    @property
    def qps(self) -> float:
        return self.answered / self.wall_s if self.wall_s else 0.0

    def pct(self, p: float) -> float:
        return float(np.percentile(self.latencies_ms, p)) if len(self.latencies_ms) else float("nan")

    def summary(self) -> dict:
        return {"sent": self.sent, "answered": self.answered, "rejected": self.rejected,
                "answered_qps": round(self.qps, 2), "p50_ms": round(self.pct(50), 3),
                "p90_ms": round(self.pct(90), 3), "p99_ms": round(self.pct(99), 3),
                "max_ms": round(float(self.latencies_ms.max()) if len(self.latencies_ms) else float("nan"), 3),
                "batches": self.batches, "max_batch": self.max_batch, "wall_s": round(self.wall_s, 3)}


def run_load(server: Server, frames: list, dbs: list, rate: float, seed: int, on_done=None) -> LoadResult:
    """Sends frames[i] to database dbs[i] on a Poisson schedule of `rate` requests/s
    (exponential gaps from numpy's PCG64 seeded with `seed`), collects every completion, and
    measures latency from the scheduled send time (src/bin/altopir.rs:612-726)."""
    n = len(frames)
    gaps = np.random.default_rng(seed).exponential(1.0 / rate, size=n)
    sched = np.cumsum(gaps)
    t0 = now_ns() + 2_000_000
    sched_ns = (t0 + sched * 1e9).astype(np.int64)
    lat = np.full(n, np.nan)
    statuses = {}
    got = 0
    stats0 = server.stats()
    done_evt = threading.Event()

    def collector():
        nonlocal got
        while got < n:
            for d in server.poll(20):
                i = d.tag
                statuses[d.status] = statuses.get(d.status, 0) + 1
                if d.status == 0:
                    lat[i] = (d.t_done_ns - sched_ns[i]) / 1e6
                if on_done:
                    on_done(i, d)
                got += 1
        done_evt.set()

    th = threading.Thread(target=collector, daemon=True)
    th.start()
    for i in range(n):
        wait = (sched_ns[i] - now_ns()) / 1e9
        if wait > 0:
            time.sleep(wait)
        server.submit(dbs[i], frames[i], i)
    done_evt.wait()
    wall = (now_ns() - t0) / 1e9
    st = server.stats()
    ok = ~np.isnan(lat)
    return LoadResult(sent=n, answered=int(ok.sum()), rejected={k: v for k, v in statuses.items() if k != 0},
                      wall_s=wall, latencies_ms=lat[ok], batches=st["batches"] - stats0["batches"],
                      max_batch=st["max_batch_seen"])


def saturate(server: Server, frames: list, db: int = 0, total: int = 1024, threads: int = 8,
             window: int = 512) -> dict:
    """Offered load above capacity: `threads` submitters keep up to `window` requests of
    database `db` outstanding (frames reused round-robin) until `total` are answered.
    Returns the answered rate over the whole run and the submit-to-completion latencies."""
    sem = threading.Semaphore(window)
    lat = np.full(total, np.nan)
    statuses = {}
    got = 0
    done_evt = threading.Event()
    t_first = [None]

    def collector():
        nonlocal got
        while got < total:
            for d in server.poll(20, copy=False):
                statuses[d.status] = statuses.get(d.status, 0) + 1
                lat[d.tag] = (d.t_done_ns - d.t_submit_ns) / 1e6
                got += 1
                sem.release()
        done_evt.set()

    def submitter(k):
        for i in range(k, total, threads):
            sem.acquire()
            server.submit(db, frames[i % len(frames)], i)

    stats0 = server.stats()
    th = threading.Thread(target=collector, daemon=True)
    th.start()
    t0 = now_ns()
    subs = [threading.Thread(target=submitter, args=(k,), daemon=True) for k in range(threads)]
    for t in subs:
        t.start()
    for t in subs:
        t.join()
    done_evt.wait()
    wall = (now_ns() - t0) / 1e9
    st = server.stats()
    ok = statuses.get(0, 0)
    batches = st["batches"] - stats0["batches"]
    v = lat[~np.isnan(lat)]
    return {"qps": round(ok / wall, 2), "answered": ok, "requests": total, "wall_s": round(wall, 4),
            "statuses": statuses, "batches": batches, "mean_batch": round(total / max(batches, 1), 1),
            "p50_ms": round(float(np.percentile(v, 50)), 3), "p99_ms": round(float(np.percentile(v, 99)), 3),
            "threads": threads, "window": window}
