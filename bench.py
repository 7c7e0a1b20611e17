#!/usr/bin/env python3
"""Benchmark of the MI355X answer engine: QPS of whole batches with the requests resident
in HBM, the roofline of the dominant kernel and the reference CPU engine beside it.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--records R] [--batch B]

A step is one batch through the device entry point (validate -> scan -> packing -> tail)
with the frames already in device memory.  Default workload: 1 GB database (32,768
records x 32 KB, bytes uniform from seed 0) and a batch of 128 requests (y, z uniform in
[0, Q), query words uniform u32, seed 2), the reference's default parameters (n = 2048,
tg = 4, q1 = 2^10, q2 = 2^18).  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "QPS at 1/8/64 GB (batch 1 & 128) + p99 latency; scan GB/s & MFMA util vs peak"
HBM_PEAK_GBS = 8000.0
INT8_PEAK_TOPS = 5000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def round_up(x, m):
    return (x + m - 1) // m * m


def measured_traffic(kernel: str, records: int, record_bytes: int, batch: int):
    """HBM bytes per launch of `kernel` on this workload, from the counter summaries kept
    under profiles/ (lines `traffic: kernel=... records=... record_bytes=... batch=...
    bytes_per_launch=...`, written from two rocprofv3 --pmc runs, FETCH_SIZE and
    WRITE_SIZE, scaled by a plain copy of known size); None when none was measured."""
    for f in sorted((ROOT / "profiles").glob("*.md")):
        for line in f.read_text().splitlines():
            if not line.startswith("traffic:"):
                continue
            kv = dict(t.split("=", 1) for t in line[len("traffic:"):].split() if "=" in t)
            if (kv.get("kernel") == kernel and int(kv.get("records", -1)) == records
                    and int(kv.get("record_bytes", -1)) == record_bytes and int(kv.get("batch", -1)) == batch):
                return int(kv["bytes_per_launch"])
    return None


def gpu_numa_cores() -> tuple[int, str]:
    """Physical cores of the GPU's NUMA node that this process may run on (one thread per core:
    the first logical CPU of each core in the node's list), and a description.  Falls back to
    every CPU the process may use when the node is not known."""
    allowed = sorted(os.sched_getaffinity(0))
    node = None
    try:
        for dev in Path("/sys/class/drm").glob("card*/device"):
            if (dev / "numa_node").exists() and (dev / "vendor").read_text().strip() == "0x1002":
                node = int((dev / "numa_node").read_text().strip())
                break
    except OSError:
        node = None
    cpus = allowed
    where = f"{len(allowed)} CPUs this process may use"
    if node is not None and node >= 0:
        try:
            spec = Path(f"/sys/devices/system/node/node{node}/cpulist").read_text().strip()
            in_node = set()
            for part in spec.split(","):
                a, _, b = part.partition("-")
                in_node.update(range(int(a), int(b or a) + 1))
            cpus = [c for c in allowed if c in in_node] or allowed
            where = f"NUMA node {node} of the GPU, {len(cpus)} of its CPUs open to this process"
        except OSError:
            pass
    cores = set()
    for c in cpus:
        try:
            sib = Path(f"/sys/devices/system/cpu/cpu{c}/topology/thread_siblings_list").read_text().strip()
            cores.add(sib.split(",")[0].split("-")[0])
        except OSError:
            cores.add(str(c))
    return max(1, len(cores)), where + f", {len(cores)} physical cores"


def cpu_baseline(inp, reqs, threads: int, budget_s: float) -> dict:
    """The reference engine compiled from its own source (oracle/_ref) on the host cores,
    method VNNI (AVX-512) or AMX, timed like `altopir bench` (src/bin/altopir.rs:741-826):
    one warm-up, then trials until the budget is spent (at most 5), median."""
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_ctypes as oc

    B = len(reqs)
    if oc.ref_available() and oc.ref_methods():
        method = 2 if 2 in oc.ref_methods() else 0
        ref = oc.RefEngine(inp, threads=threads)
        out = np.zeros((B, inp.response_len), np.uint8)
        st = np.full(B, -1, np.int32)
        ref.answer(reqs, out, st, method)  # warm-up
        times = []
        t_end = time.time() + budget_s
        while len(times) < 5 and (not times or time.time() + times[-1] < t_end):
            t = time.perf_counter()
            r, _ = ref.answer(reqs, out, st, method)
            times.append(time.perf_counter() - t)
            assert r == B
        ref.close()
        med = statistics.median(times)
        return {"value": round(B / med, 3), "unit": "queries/s", "cores": threads, "kind": "reference",
                "sample": f"reference engine (src/amx_kernels.cpp, method {'vnni' if method == 2 else 'amx'}) "
                          f"on the same {B}-request batch, {len(times)} trials after a warm-up, median "
                          f"{med * 1e3:.1f} ms"}
    # no reference build or no AVX-512 VNNI/AMX: time the oracle port on one request
    out = np.zeros((1, inp.response_len), np.uint8)
    st = np.full(1, -1, np.int32)
    t = time.perf_counter()
    oc.oracle_answer(inp, reqs[:1], out, st)
    dt = time.perf_counter() - t
    return {"value": round(1 / dt, 4), "unit": "queries/s", "cores": os.cpu_count(), "kind": "port",
            "sample": f"oracle restatement, 1 request of the workload, {dt:.2f} s"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--records", type=int, default=32768)
    ap.add_argument("--record-bytes", type=int, default=32768)
    ap.add_argument("--batch", type=int, default=128)
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--no-cpu-numa", dest="cpu_numa", action="store_false",
                    help="skip the second CPU baseline run on every physical core of the GPU's NUMA node")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--also-batch", type=int, nargs="*", default=[1],
                    help="further batch sizes measured on the same engine and reported under 'also'")
    ap.add_argument("--plain-steps", type=int, default=5,
                    help="batches timed through the synchronous host entry point (0: skip)")
    ap.add_argument("--no-graph-replay", dest="graph_pass", action="store_false",
                    help="skip the second pass of the timed steps replayed from captured HIP graphs")
    ap.add_argument("--fusion", type=int, default=-1, choices=[-1, 0, 1],
                    help="scan and packing in one launch: 1 wherever the shape pair allows, 0 never, "
                         "-1 the engine's built-in choice (batches of 1-4 queries on databases up to 4 KP records)")
    ap.add_argument("--scan-variant", default=None, help="force a scan contraction variant (alto_variant_name)")
    ap.add_argument("--pack-variant", default=None, help="force a packing contraction variant")
    ap.add_argument("--ingest-records", action="store_true",
                    help="build the engine from the device-resident bytes taken as records (row-major)")
    ap.add_argument("--seed", type=int, default=2, help="request seed")
    ap.add_argument("--tune", action="store_true",
                    help="time every tile-shape variant per batch size first and keep the fastest")
    ap.add_argument("--serve-batches", type=int, default=32,
                    help="batches' worth of requests pushed through the serving loop (0: skip)")
    ap.add_argument("--serve-threads", type=int, default=16, help="submitter threads of the served form")
    ap.add_argument("--serve-contexts", type=int, default=2, help="batches in flight in the served form")
    ap.add_argument("--serve-window", type=int, default=6, help="outstanding requests, in batches")
    ap.add_argument("--serve-graphs", type=int, default=-1, choices=[-1, 0, 1],
                    help="graph mode of the served form's contexts (-1 follow the engine)")
    ap.add_argument("--serve-upload", type=int, default=-1, choices=[-1, 0, 1, 2, 3],
                    help="served form: 0 copy frames into HBM as they arrive, 1 kernels read the pinned slots, "
                         "2 copy a batch's frames into HBM at dispatch, 3 that and copy its responses back, "
                         "-1 the server's default (0)")
    ap.add_argument("--copy-calib", type=int, default=0,
                    help="also run a plain device copy of this many MiB (calibrates HBM byte counters)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only: this build has no multi-GPU path (asked for --gpus {args.gpus})",
              file=sys.stderr)
        return 2

    import torch

    from altopir_public_amd import synthetic, workload
    from altopir_public_amd.engine import Engine

    assert torch.cuda.is_available(), "bench.py needs a HIP device"
    torch.cuda.set_device(0)
    R, C, B = args.records, args.record_bytes, args.batch
    n, tg = 2048, 4
    t0 = time.time()
    # the database (uniform bytes of stream 0, column-major) is generated straight into
    # device memory; the other setup inputs on the host
    inp = workload.make_inputs(R, C, n=n, tg=tg, seed=0, db=False)
    d_db = synthetic.uniform_bytes_device(0, R * C)
    torch.cuda.synchronize()
    log(f"[bench] inputs {R} x {C} B generated in {time.time() - t0:.1f} s")
    t0 = time.time()
    if args.ingest_records:
        # the same bytes taken as records (row-major), transposed into the engine's layout on
        # the device: a different database of the same distribution
        eng = Engine.from_records(d_db.data_ptr(), True, C, inp, threads=args.cpu_threads)
    else:
        eng = Engine.from_device_db(d_db.data_ptr(), inp, threads=args.cpu_threads)
    del d_db
    torch.cuda.empty_cache()
    log(f"[bench] engine set up in {time.time() - t0:.2f} s from {'records' if args.ingest_records else 'columns'}, "
        f"{eng.device_bytes() / 2**30:.2f} GiB resident")
    L, rl = inp.request_len, inp.response_len
    stride = round_up(7 + L, 256)
    KP = tg * n
    torch.cuda.synchronize()
    # a stream of our own (default priority): the engine launches on it and the timing
    # events below are recorded on it
    work = torch.cuda.Stream()
    torch.cuda.set_stream(work)
    stream = work.cuda_stream
    assert stream, "expected a non-null stream handle"

    def measure(Bm: int, steps: int, warmup: int, seed: int) -> dict:
        # request frames generated in device memory (the same bytes as workload.make_requests)
        slots = workload.make_requests_device(Bm, R, inp.kp, seed=seed, stride=stride)
        status = torch.zeros(Bm, dtype=torch.int32, device="cuda")
        out = torch.empty((Bm, rl), dtype=torch.uint8, device="cuda")
        torch.cuda.synchronize()

        def step():
            rc = eng.answer_device(slots.data_ptr(), stride, Bm, out.data_ptr(), rl, status.data_ptr(), stream)
            if rc != 0:
                raise RuntimeError(f"alto_answer_device returned {rc}")

        for _ in range(warmup):
            step()
        torch.cuda.synchronize()
        assert int((status != 0).sum()) == 0, "a valid request was rejected"
        start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

        def timed(k: int) -> float:
            torch.cuda.synchronize()
            start.record()
            for _ in range(k):
                step()
            end.record()
            torch.cuda.synchronize()
            return start.elapsed_time(end) / k

        # timed region: eager launches, each kernel bracketed by the engine's HIP events on
        # the launch stream (a ring of event sets folded into running sums, so the host never
        # waits inside the region); the per-kernel durations come from the same steps as the
        # throughput.  (Events recorded inside captured graphs cost about 8 us each per replay,
        # more than eager launches lose, so the graph path is timed as a whole below.)
        eng.set_graphs(False)
        step()
        torch.cuda.synchronize()
        eng.stage_totals(reset=True)
        ms = timed(steps)
        nb, sums = eng.stage_totals(reset=True)
        assert nb == steps, (nb, steps)
        st = np.array(sums) / nb
        region = "eager"
        # the same steps replayed from the captured HIP graph (the production path), untimed per kernel
        eng.set_graphs(True)
        ms_graph = None
        if args.graph_pass:
            step()
            ms_graph = timed(steps)
        names = ["validate", "scan_prep", "scan", "pack_prep", "pack", "tail"]
        stage_ms = {k: round(float(v), 4) for k, v in zip(names, st)}
        scan_ops, pack_ops = 8 * R * C * Bm, 32 * KP * C * Bm
        scan_bytes, pack_bytes = C * R + 4 * Bm * R, 4 * KP * C + 4 * Bm * KP
        fused = eng.fused(Bm)
        if fused:
            # both contractions ran in one launch, timed as the "scan" stage
            dom, t_dom = "scan_pack", stage_ms["scan"] / 1e3
            ops, byts = scan_ops + pack_ops, scan_bytes + pack_bytes
        else:
            dom = "scan" if stage_ms["scan"] >= stage_ms["pack"] else "pack"
            t_dom = stage_ms[dom] / 1e3
            ops, byts = (scan_ops, scan_bytes) if dom == "scan" else (pack_ops, pack_bytes)
        # ridge: 5.0e15 op/s / 8.0e12 B/s = 625 op/B
        traffic = measured_traffic(f"{dom}_contraction", R, C, Bm)
        if ops / byts > 625.0:
            achieved = ops / t_dom / 1e12
            roof = {"bound": "mfma", "kernel": f"{dom}_contraction", "achieved": round(achieved, 2),
                    "peak": INT8_PEAK_TOPS, "unit": "TOP/s", "frac": round(achieved / INT8_PEAK_TOPS, 4),
                    "traffic": traffic, "algorithmic_ops": ops, "algorithmic_bytes": byts}
        else:
            achieved = byts / t_dom / 1e9
            roof = {"bound": "hbm", "kernel": f"{dom}_contraction", "achieved": round(achieved, 1),
                    "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                    "traffic": traffic, "algorithmic_bytes": byts, "algorithmic_ops": ops}
        res = {"batch": Bm, "qps": round(Bm / (ms / 1e3), 2), "ms_per_step": round(ms, 4), "stage_ms": stage_ms,
               "fused_scan_pack": fused,
               "timed_region": region,
               "graph_replay": ({"qps": round(Bm / (ms_graph / 1e3), 2), "ms_per_step": round(ms_graph, 4)}
                                if ms_graph else None),
               "roofline": roof}
        if not fused:
            res.update(scan_GBps=round(scan_bytes / (stage_ms["scan"] / 1e3) / 1e9, 1),
                       scan_TOPs=round(scan_ops / (stage_ms["scan"] / 1e3) / 1e12, 1),
                       pack_TOPs=round(pack_ops / (stage_ms["pack"] / 1e3) / 1e12, 1))
        gr = f" ({ms_graph:.3f} graph replay)" if ms_graph else ""
        log(f"[bench] B={Bm}: {ms:.3f} ms/step{gr}, {res['qps']:.1f} QPS, stage ms {stage_ms}")
        res["_slots"] = slots
        return res

    tuned = None
    if args.tune:
        from altopir_public_amd import autotune

        eng.set_graphs(False)
        table = autotune.tune(eng, inp, [B, *args.also_batch], seed=7, log=log)
        eng.set_graphs(True)
        tuned = {str(v["B"]): {"scan": v["scan"][0], "pack": v["pack"][0]} for v in table.values()}

    eng.set_fusion(args.fusion)
    if args.scan_variant or args.pack_variant:
        from altopir_public_amd import autotune

        names = autotune.variant_names()
        for b in [B, *args.also_batch]:
            for which, v in ((autotune.SCAN, args.scan_variant), (autotune.PACK, args.pack_variant)):
                if v:
                    eng._lib.alto_tune_set(eng.ptr, which, b, names.index(v))
    main = measure(B, args.steps, args.warmup, seed=args.seed)
    ms_per_step, qps, stage_ms, roof = main["ms_per_step"], main["qps"], main["stage_ms"], main["roofline"]
    slots = main.pop("_slots")
    also = []
    for b in args.also_batch:
        r = measure(b, args.steps, args.warmup, seed=args.seed)
        r.pop("_slots")
        also.append(r)

    plain = None
    if args.plain_steps:
        # the plain form: the synchronous host entry point of the reference's API, frames in
        # host memory, responses copied back, one batch per call (PCIe transfers included)
        frames = [np.ascontiguousarray(f[7:7 + L]) for f in slots.cpu().numpy()]
        hout = np.empty((B, rl), dtype=np.uint8)
        hst = np.zeros(B, dtype=np.int32)
        eng.answer(frames, hout, hst)
        assert int((hst != 0).sum()) == 0
        t0 = time.perf_counter()
        for _ in range(args.plain_steps):
            eng.answer(frames, hout, hst)
        dt = (time.perf_counter() - t0) / args.plain_steps
        plain = {"qps": round(B / dt, 2), "ms_per_step": round(dt * 1e3, 4), "steps": args.plain_steps,
                 "form": "alto_answer, host frames in pageable memory, responses copied back"}
        log(f"[bench] plain host form B={B}: {dt * 1e3:.3f} ms/batch, {B / dt:.1f} QPS")
        # the same with frames and responses in pinned host memory (the copies go straight to the
        # DMA engines, as the served form's do), so that batching is compared with batching
        pin_in = torch.empty((B, L), dtype=torch.uint8, pin_memory=True)
        pin_out = torch.empty((B, rl), dtype=torch.uint8, pin_memory=True)
        pin_in.copy_(torch.from_numpy(np.stack(frames)))
        pframes = [pin_in[c].numpy() for c in range(B)]
        pout = pin_out.numpy()
        eng.answer(pframes, pout, hst)
        assert int((hst != 0).sum()) == 0 and np.array_equal(pout, hout)
        t0 = time.perf_counter()
        for _ in range(args.plain_steps):
            eng.answer(pframes, pout, hst)
        dtp = (time.perf_counter() - t0) / args.plain_steps
        plain["pinned"] = {"qps": round(B / dtp, 2), "ms_per_step": round(dtp * 1e3, 4),
                           "form": "alto_answer, host frames and responses in pinned memory"}
        log(f"[bench] plain host form, pinned frames B={B}: {dtp * 1e3:.3f} ms/batch, {B / dtp:.1f} QPS")

    served = None
    if args.serve_batches:
        # the served form: the continuous-batching loop fed from host memory by submitter
        # threads, uploads overlapped with device work, at an offered load above capacity
        from altopir_public_amd import serve

        frames = [np.ascontiguousarray(f[7:7 + L]) for f in slots.cpu().numpy()]
        win = args.serve_window * B
        srv = serve.Server([eng], max_batch=B, slots=win, contexts=args.serve_contexts)
        srv.set_graphs(args.serve_graphs)
        srv.set_upload(args.serve_upload)
        srv.replay(frames, total=2 * B, threads=args.serve_threads, window=win)  # warm-up
        served = srv.replay(frames, total=args.serve_batches * B, threads=args.serve_threads, window=win)
        served["form"] = (f"alto_server_* loop driven by alto_server_replay: {args.serve_threads} native submitter "
                          f"threads copy host frames into pinned slots, up to {win} requests outstanding")
        served["contexts"] = srv.contexts(0)
        served["upload"] = {0: "frames copied into HBM as they arrive", 1: "kernels read the pinned slots",
                            2: "batch copied into HBM at dispatch",
                            3: "batch copied into HBM at dispatch, responses copied back"}[srv.upload_mode(0)]
        srv.close()
        log(f"[bench] served form: {served}")

    if args.copy_calib:
        # a plain copy of known size: its FETCH_SIZE / WRITE_SIZE counters scale the
        # counters of the engine's kernels to bytes
        x = torch.empty(args.copy_calib << 20, dtype=torch.uint8, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)  # 4-byte accesses per lane
        for _ in range(3):
            y.view(torch.int32).copy_(x.view(torch.int32))  # 16-byte accesses per lane, as the engine's
        torch.cuda.synchronize()
        log(f"[bench] copy calibration: 3 copies of {args.copy_calib} MiB")

    cpu = None
    if not args.no_cpu_baseline:
        log("[bench] cpu baseline ...")
        reqs = [f[7:7 + L] for f in slots.cpu().numpy()]
        if R * C <= (2 << 30):
            if inp.db_col is None:
                inp.db_col = synthetic.uniform_bytes(0, R * C)  # the same bytes, on the host
            cpu = cpu_baseline(inp, reqs, args.cpu_threads, args.cpu_budget)
            if args.cpu_numa:
                ncores, where = gpu_numa_cores()
                numa = cpu_baseline(inp, reqs, ncores, args.cpu_budget)
                numa["where"] = where
                cpu["numa_node"] = numa
                log(f"[bench] cpu baseline on the GPU's NUMA node: {numa}")
        else:
            # bounded sample: the first n columns (one output part) of the same database, the
            # same requests; scan, packing and tail all scale linearly in the columns
            part = workload.make_inputs(R, n, n=n, tg=tg, seed=0)
            cpu = cpu_baseline(part, reqs, args.cpu_threads, args.cpu_budget)
            scale = C // n
            cpu["value"] = round(cpu["value"] / scale, 4)
            cpu["sample"] = (f"{cpu['sample']}; on {n} of the {C} columns (one of {scale} output parts), "
                             f"QPS divided by {scale}")
        log(f"[bench] cpu baseline {cpu}")

    line = {
        "metric": METRIC,
        "value": round(qps, 2),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int8",
        "data": "synthetic",
        "config": {
            "workload": f"{R * C / 2**30:g} GiB database ({R} records x {C} B, uniform bytes seed 0), "
                        f"batch {B} (uniform requests seed {args.seed}), requests resident in HBM",
            "records": R, "record_bytes": C, "batch": B, "n": n, "tg": tg, "parallelism": "single GPU",
        },
        "stage_ms": stage_ms,
        "roofline": roof,
        "also": also,
        "tuned": tuned,
        "plain_host": plain,
        "served": served,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    eng.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
