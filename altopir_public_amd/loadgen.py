"""Open-loop load against the device-resident serving loop, after the reference's `load`
command (src/bin/altopir.rs:594-739): requests for uniformly chosen databases and records are
sent on a Poisson schedule, a fixed fraction is corrupted to exercise validation, and the
answered rate and the latency percentiles (scheduled send -> completion) are reported.

    python -m altopir_public_amd.loadgen --gib 1 16 32 --rates 50 200 500 1000 2000 --duration 60

Each database holds R = gib * 2^30 / C records of C uniform bytes (C = --record-bytes, default
32,768; seeds 0, 1, 2, ...), generated in device memory, with its own packing matrix, tail keys and
prefixes.  Requests are
drawn from a pool of distinct valid frames per database (seed 5) and 1% are corrupted with
workload.corrupt.  One JSON line per rate is printed on stdout.

With --scheme the databases are set up as a server does (offline.py: the hint, the packing
precomputation and the packing matrix on the device) and the pool holds real queries of the
scheme's client (client.py) for random records; after each rate a sample of the answered
responses is decrypted with their clients' secrets and compared with the records.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_engines(gibs, cols: int, n: int, tg: int, threads: int, scheme_pool: int = 0, seed: int = 5):
    """Engines over databases generated in device memory.  scheme_pool > 0: set up as a server
    does (offline.py) and, per database, scheme_pool (index, client, frame, record) queries of
    the scheme's client (else None)."""
    import torch

    from . import client, offline, synthetic, workload
    from .engine import Engine

    engines, inputs, queries = [], [], []
    for i, g in enumerate(gibs):
        rows = int(round(g * 2**30 / cols))
        rows -= rows % (n if scheme_pool else 64)
        t0 = time.time()
        d_db = synthetic.uniform_bytes_device(i, rows * cols)
        torch.cuda.synchronize()
        if scheme_pool:
            p = client.Params(rows, cols, n, tg)
            hint = offline.device_hint(p, d_db.data_ptr(), on_device=True)
            inp = offline.server_inputs(p, None, offline.device_offline(p, hint), device=True)
            del hint
            idx = [int(x) for x in np.random.default_rng([seed, i]).integers(0, rows, scheme_pool)]
            recs = d_db.view(cols, rows)[:, idx].T.contiguous().cpu().numpy()
            qs = []
            for k, r in enumerate(idx):
                c, f = client.new_request(p, r, seed=seed * 100003 + i * 1009 + k)
                qs.append((r, c, f, recs[k].tobytes()))
            queries.append(qs)
        else:
            inp = workload.make_inputs(rows, cols, n=n, tg=tg, seed=i, db=False)
            queries.append(None)
        eng = Engine.from_device_db(d_db.data_ptr(), inp, threads=threads)
        del d_db
        torch.cuda.empty_cache()
        log(f"[load] database {i}: {g:g} GiB ({rows} records), engine set up in {time.time() - t0:.1f} s")
        engines.append(eng)
        inputs.append(inp)
    return engines, inputs, queries


def frame_pool(inp, count: int, seed: int) -> list:
    """`count` distinct valid frames for one database, generated on the device."""
    from . import workload

    L = inp.request_len
    stride = (7 + L + 255) // 256 * 256
    arena = workload.make_requests_device(count, inp.rows, inp.kp, seed=seed, stride=stride)
    host = arena.cpu().numpy()
    return [np.ascontiguousarray(host[c, 7:7 + L]) for c in range(count)]


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--gib", type=float, nargs="+", default=[1, 16, 32], help="database sizes (GiB)")
    ap.add_argument("--rates", type=float, nargs="+", default=[50, 200, 500, 1000, 2000], help="offered req/s")
    ap.add_argument("--duration", type=float, default=60.0, help="seconds of arrivals per rate")
    ap.add_argument("--corrupt", type=float, default=0.01, help="fraction of corrupted frames")
    ap.add_argument("--pool", type=int, default=64, help="distinct valid frames per database")
    ap.add_argument("--seed", type=int, default=5)
    ap.add_argument("--max-batch", type=int, default=128)
    ap.add_argument("--slots", type=int, default=256, help="request slots per database")
    ap.add_argument("--contexts", type=int, default=2, help="batches in flight per database")
    ap.add_argument("--threads", type=int, default=16)
    ap.add_argument("--record-bytes", type=int, default=32768, help="record size (a multiple of n = 2048)")
    ap.add_argument("--scheme", action="store_true",
                    help="databases set up as a server does, requests from the scheme's client, responses decrypted")
    ap.add_argument("--decrypt-sample", type=int, default=32, help="responses decrypted per rate (--scheme)")
    args = ap.parse_args(argv)

    import torch

    from . import client, serve, workload

    assert torch.cuda.is_available(), "the load generator needs a HIP device"
    torch.cuda.set_device(0)
    cols, n, tg = args.record_bytes, 2048, 4
    engines, inputs, queries = build_engines(args.gib, cols, n, tg, args.threads,
                                             scheme_pool=args.pool if args.scheme else 0, seed=args.seed)
    if args.scheme:
        pools = [[np.asarray(q[2]) for q in qs] for qs in queries]
    else:
        pools = [frame_pool(inp, args.pool, seed=args.seed * 1000 + i) for i, inp in enumerate(inputs)]
    srv = serve.Server(engines, max_batch=args.max_batch, slots=args.slots, contexts=args.contexts)
    ctxs = [srv.contexts(i) for i in range(len(engines))]
    log(f"[load] serving {len(engines)} databases, contexts per database {ctxs} to start with, "
        f"up to {args.contexts} admitted under load")
    for k, rate in enumerate(args.rates):
        count = max(1, int(round(rate * args.duration)))
        rng = np.random.default_rng([args.seed, k])
        dbs = rng.integers(len(engines), size=count)
        pick = rng.integers(args.pool, size=count)
        bad = rng.random(count) < args.corrupt
        frames = []
        for i in range(count):
            f = pools[dbs[i]][pick[i]]
            frames.append(workload.corrupt(f, rng, inputs[dbs[i]].kp) if bad[i] else f)
        expect_bad = {int(i) for i in np.flatnonzero(bad)
                      if workload.frame_status(frames[i], inputs[dbs[i]].rows, inputs[dbs[i]].kp) != 0}
        seen_bad = set()
        sample = []

        def on_done(i, d):
            if d.status != 0:
                seen_bad.add(i)
            elif args.scheme and not bad[i] and len(sample) < args.decrypt_sample:
                sample.append((int(dbs[i]), int(pick[i]), d.response))

        res = serve.run_load(srv, frames, [int(x) for x in dbs], rate, seed=args.seed * 7919 + k, on_done=on_done)
        decrypted = None
        if args.scheme:
            decrypted = all(client.decrypt_response(queries[d][j][1].params, queries[d][j][1], resp) == queries[d][j][3]
                            for d, j, resp in sample)
        out = {"offered_rps": rate, "duration_s": args.duration, "databases_gib": args.gib, **res.summary(),
               "corrupted": int(bad.sum()), "rejected_as_expected": seen_bad == expect_bad,
               "requests": "scheme client queries" if args.scheme else "seeded frames",
               "decrypted_sample": (len(sample) if args.scheme else 0),
               "sample_decrypts_to_the_records": decrypted,
               "contexts": [srv.contexts(i) for i in range(len(engines))],  # after this rate's load
               "max_batch_limit": args.max_batch}
        print(json.dumps(out), flush=True)
        log(f"[load] {rate:g} req/s: {res.summary()}")
    srv.close()
    for e in engines:
        e.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
