"""The server's offline setup for the scheme of client.py: hint, packing precomputation, engine.

From a database and the public values of client.py the server computes (src/service.rs:
129-170, Prepared::new with perform_offline_precomputation_simplepir, src/server.rs:947-1021):
  * the hint: for every column, the sum over the record polynomials of their products with the
    public query polynomials (answer_hint_ring; alto_hint0 on the device);
  * per output part, the InspiRING precomputation from the hint and the packing keys' masks:
    the aggregates bold_t, bold_t_bar (one gadget-decomposed polynomial per generator power),
    bold_t_hat and the packed ciphertext's first row a_hat (src/packing.rs:2260-2359;
    alto_inspiring_offline on the device);
  * the packing matrix U from bold_t / bold_t_bar (sw_pack_build; alto_pack_build on the
    device), the tail keys (bold_t_hat), the row-0 prefixes (a_hat switched to q2 = 2^18,
    src/modulus_switch.rs:136-147) and the inverse-NTT table.
server_inputs() assembles the engine's inputs from these (workload.EngineInputs).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib, client as cl, workload

Q = workload.Q


def automorph_tables(n: int) -> np.ndarray:
    """Every automorphism's slot table, table (t - 1) / 2 for odd t < 2n, n x n u32
    (src/packing.rs:1082-1110, the layout sw_pack_build reads)."""
    return np.stack([cl.auto_slots(t, n) for t in range(1, 2 * n, 2)]).astype(np.uint32).reshape(-1)


def write_fields(buf: np.ndarray, bit0: int, bits: int, vals: np.ndarray) -> None:
    """ORs LSB-first fields of `bits` bits from bit offset bit0 into buf (src/bits.rs:11-41)."""
    vals = np.asarray(vals, dtype=np.uint64)
    b = ((vals[:, None] >> np.arange(bits, dtype=np.uint64)) & np.uint64(1)).astype(np.uint8).reshape(-1)
    allbits = np.unpackbits(buf, bitorder="little")
    allbits[bit0:bit0 + len(b)] |= b
    buf[:] = np.packbits(allbits, bitorder="little")[:len(buf)]


def prefix_from_a_hat(a_hat: np.ndarray, n: int) -> np.ndarray:
    """Row 0 of a response part: a_hat rescaled to q2 = 2^18, 18-bit fields from bit 0
    (switch_prefix_row0, src/modulus_switch.rs:136-147); row 1 left zero."""
    buf = np.zeros(workload.part_len_for(n), np.uint8)
    write_fields(buf, 0, 18, cl.rescale(a_hat, Q, cl.Q2))
    return buf


def device_hint(params: cl.Params, db_col, ld: int | None = None, on_device: bool = False) -> np.ndarray:
    """The hint (n x cols) computed on the device from a column-major database (alto_hint0):
    db_col a host array, or with on_device the address of the database in device memory."""
    ld = params.rows if ld is None else ld
    a = np.ascontiguousarray(cl.hint_query_polys(params), dtype=np.uint64)
    out = np.zeros((params.n, params.cols), np.uint64)
    lib = _lib.load()
    ptr = db_col if on_device else db_col.ctypes.data
    rc = lib.alto_hint0(ptr, 1 if on_device else 0, ld, params.rows, params.cols, params.n, a.ctypes.data, a.shape[0],
                        out.ctypes.data)
    if rc != 0:
        raise RuntimeError(f"alto_hint0 failed ({_lib.last_error()})")
    return out


def device_offline(params: cl.Params, hint: np.ndarray) -> list[tuple]:
    """The packing precomputation of every output part on the device
    (alto_inspiring_offline): a list of (bold_t, bold_t_bar, bold_t_hat, a_hat) per part."""
    n, tg, O = params.n, params.tg, params.outputs
    w, v = (np.ascontiguousarray(x, dtype=np.uint64) for x in cl.masks(params))
    hint = np.ascontiguousarray(hint, dtype=np.uint64)
    m = n // 2 - 1
    bt = np.zeros((O, m * tg * n), np.uint64)
    bb = np.zeros((O, m * tg * n), np.uint64)
    bh = np.zeros((O, tg * n), np.uint64)
    ah = np.zeros((O, n), np.uint64)
    lib = _lib.load()
    rc = lib.alto_inspiring_offline(bt.ctypes.data, bb.ctypes.data, bh.ctypes.data, ah.ctypes.data,
                                    hint.ctypes.data, params.cols, O, w.ctypes.data, v.ctypes.data, n, tg)
    if rc != 0:
        raise RuntimeError(f"alto_inspiring_offline failed ({_lib.last_error()})")
    return [(bt[o], bb[o], bh[o], ah[o]) for o in range(O)]


def device_setup(params: cl.Params, db_col, ld: int | None = None, on_device: bool = False) -> workload.EngineInputs:
    """The whole setup on the device in one call (alto_scheme_setup: hint, packing
    precomputation and packing matrix without the intermediate results leaving the device):
    the engine's inputs.  db_col a host array, or with on_device the address of the database in
    device memory (then the returned inputs carry no database)."""
    n, tg, O = params.n, params.tg, params.outputs
    ld = params.rows if ld is None else ld
    a = np.ascontiguousarray(cl.hint_query_polys(params), dtype=np.uint64)
    w, v = (np.ascontiguousarray(x, dtype=np.uint64) for x in cl.masks(params))
    ub = np.zeros(4 * tg * n * params.cols, np.int8)
    that = np.zeros((O, tg * n), np.uint64)
    a_hat = np.zeros((O, n), np.uint64)
    lib = _lib.load()
    rc = lib.alto_scheme_setup(ub.ctypes.data, that.ctypes.data, a_hat.ctypes.data,
                               db_col if on_device else db_col.ctypes.data, 1 if on_device else 0, ld, params.rows,
                               params.cols, a.ctypes.data, a.shape[0], w.ctypes.data, v.ctypes.data, n, tg)
    if rc != 0:
        raise RuntimeError(f"alto_scheme_setup failed ({_lib.last_error()})")
    parts = [(None, None, that[o], a_hat[o]) for o in range(O)]
    return server_inputs(params, None if on_device else db_col, parts, ub=ub)


def server_inputs(params: cl.Params, db_col: np.ndarray, parts: list[tuple], ub: np.ndarray | None = None,
                  device: bool = True, threads: int = 16) -> workload.EngineInputs:
    """The engine's inputs from the per-part precomputation: U built from bold_t / bold_t_bar
    (on the device, or with sw_pack_build on host threads), that = bold_t_hat, the row-0
    prefixes from a_hat, the inverse-NTT table of the client's NTT.  ub: a packing matrix
    built already (skips the build)."""
    from .engine import build_packing_matrix

    n, tg = params.n, params.tg
    if ub is None:
        ub = np.zeros(4 * tg * n * params.cols, np.int8)
        build_packing_matrix(ub, [p[0] for p in parts], [p[1] for p in parts], automorph_tables(n),
                             cl.gen_pows(n).astype(np.uint32), n // 2 - 1, tg, n, Q, threads, device=device)
    return workload.EngineInputs(rows=params.rows, cols=params.cols, n=n, tg=tg, ld=params.rows, db_col=db_col,
                                 ub=ub, prefixes=[prefix_from_a_hat(p[3], n) for p in parts],
                                 that=[np.ascontiguousarray(p[2], dtype=np.uint64) for p in parts],
                                 inv_table=workload.inverse_table(n), part_len=workload.part_len_for(n),
                                 row1_off=18 * n)


def main(argv=None) -> int:
    """Times the server's setup on the device for a database of --records x --record-bytes
    uniform bytes generated in device memory (seed 0), then answers --queries client queries
    with the HIP engine and checks that each decrypts to its record.  Prints one JSON line."""
    import argparse
    import json
    import time

    import torch

    from . import synthetic
    from .engine import Engine

    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument("--records", type=int, default=32768)
    ap.add_argument("--record-bytes", type=int, default=32768)
    ap.add_argument("--queries", type=int, default=3)
    ap.add_argument("--staged", action="store_true",
                    help="the three steps as separate calls, each result through host memory")
    args = ap.parse_args(argv)
    p = cl.Params(args.records, args.record_bytes)
    torch.cuda.set_device(0)
    t = {}
    d_db = synthetic.uniform_bytes_device(0, p.rows * p.cols)
    torch.cuda.synchronize()
    t0 = time.time()
    if args.staged:
        hint = device_hint(p, d_db.data_ptr(), on_device=True)
        t["hint_s"] = time.time() - t0
        t0 = time.time()
        parts = device_offline(p, hint)
        t["packing_precomputation_s"] = time.time() - t0
        t0 = time.time()
        inp = server_inputs(p, None, parts, device=True)
        t["packing_matrix_s"] = time.time() - t0
        del parts, hint
    else:
        inp = device_setup(p, d_db.data_ptr(), on_device=True)
        t["hint_precomputation_packing_matrix_s"] = time.time() - t0
    t0 = time.time()
    eng = Engine.from_device_db(d_db.data_ptr(), inp, threads=16)
    t["engine_s"] = time.time() - t0
    t["setup_total_s"] = sum(t.values())
    idx = [int(i) for i in np.random.default_rng(3).integers(0, p.rows, args.queries)]
    clients, reqs = [], []
    t0 = time.time()
    for k, i in enumerate(idx):
        c, f = cl.new_request(p, i, seed=50 + k)
        clients.append(c)
        reqs.append(f)
    t["client_queries_s"] = time.time() - t0
    out = np.zeros((len(reqs), p.response_len), np.uint8)
    st = np.full(len(reqs), -1, np.int32)
    answered, _ = eng.answer(reqs, out, st)
    ok = answered == len(idx)
    t0 = time.time()
    for k, i in enumerate(idx):
        rec = d_db.view(p.cols, p.rows)[:, i].cpu().numpy().tobytes()
        ok = ok and cl.decrypt_response(p, clients[k], out[k]) == rec
    t["client_decrypt_s"] = time.time() - t0
    eng.close()
    print(json.dumps({"records": p.rows, "record_bytes": p.cols, "n": p.n, "queries": len(idx),
                      "decrypted_to_the_records": bool(ok), **{k: round(v, 3) for k, v in t.items()}}), flush=True)
    return 0 if ok else 1


if __name__ == "__main__":
    raise SystemExit(main())
