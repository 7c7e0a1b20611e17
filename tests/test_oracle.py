"""Pins the oracle (oracle/liboracle.so) against the reference and against mathematics.

Checks, per oracle function:
  * orc_answer        byte parity with the reference engine compiled from its own source
                      (oracle/_ref, both methods the CPU has) on several shapes, and with
                      the committed golden fixtures (tests/golden) made by that build
  * orc_validate      the status vector of tests/engine.rs:130-148 and the mutate fuzz
                      model of tests/engine.rs:29-68 against the decoder rule
  * orc_scan_acc      brute force with Python integers on tiny shapes; linearity
  * orc_scan          the rescale floor((acc * Q + 2^31) / 2^32) on special accumulators
  * orc_pack          brute force with Python integers; U read back from the tile layout
                      is checked against an independent placement formula
  * orc_intt          INTT(NTT(x)) == x with a naive negacyclic NTT, and equality with a
                      naive inverse transform
  * orc_rescale_q1    exhaustive comparison over every v in [0, Q) with rescale_i64
  * orc_write_bits    the golden patterns of src/bits.rs:115-133
  * orc_pack_build    the reference's sw_pack_build (oracle/_ref)
"""
from __future__ import annotations

import ctypes as C
import random

import numpy as np
import pytest

from altopir_public_amd import synthetic, workload
import oracle_ctypes as oc
import refmodel

Q = workload.Q
needs_ref = pytest.mark.skipif(not oc.ref_available(), reason="oracle/_ref not built (needs the reference source)")


def _u64(a):
    return np.ascontiguousarray(a, dtype=np.uint64)


# ---------------------------------------------------------------- bits / rescale --

def test_write_bits_golden_patterns():
    lib = oc.oracle()
    cases = [  # src/bits.rs:115-133
        (0b11010101, 1, 6, [0b00101010, 0, 0, 0]),
        (0b11111111, 0, 8, [0b11111111, 0, 0, 0]),
        (0b10101010, 4, 4, [0b10100000, 0, 0, 0]),
    ]
    for val, off, nb, want in cases:
        buf = np.zeros(4, np.uint8)
        lib.orc_write_bits(buf.ctypes.data, 4, val, off, nb)
        assert list(buf) == want


def test_write_bits_round_trip_14_bit_fields():
    # src/bits.rs:135-151 with seeded values
    lib = oc.oracle()
    rng = random.Random(3)
    vals = [rng.randrange(1 << 14) for _ in range(42)]
    buf = np.zeros((42 * 14 + 7) // 8, np.uint8)
    for i, v in enumerate(vals):
        lib.orc_write_bits(buf.ctypes.data, len(buf), v, i * 14, 14)
    bits = np.unpackbits(buf, bitorder="little")
    for i, v in enumerate(vals):
        got = int(sum(int(b) << j for j, b in enumerate(bits[i * 14:(i + 1) * 14])))
        assert got == v


def test_rescale_q1_matches_rescale_i64_exhaustively():
    lib = oc.oracle()
    lib.orc_count_rescale_mismatch.restype = C.c_uint64
    lib.orc_count_rescale_mismatch.argtypes = [C.c_uint64] * 3
    assert lib.orc_count_rescale_mismatch(0, Q, 1024) == 0


@pytest.mark.parametrize("q1", [2, 3, 1000, 1 << 18, (1 << 21) - 1])
def test_rescale_q1_other_moduli_sampled(q1):
    lib = oc.oracle()
    rng = random.Random(q1)
    vs = [0, 1, Q // 2 - 1, Q // 2, Q // 2 + 1, Q - 1] + [rng.randrange(Q) for _ in range(20000)]
    for v in vs:
        assert lib.orc_rescale_q1(v, q1) == lib.orc_rescale_i64(v, Q, q1)


def test_rescale_is_rounding_to_nearest():
    # closed form: r = round(v_centred * q1 / Q) mod q1, ties away from zero
    lib = oc.oracle()
    rng = random.Random(5)
    for _ in range(5000):
        v = rng.randrange(Q)
        c = v - Q if v >= Q // 2 else v
        num = c * 1024
        r = (abs(num) * 2 + Q) // (2 * Q)
        r = r if num >= 0 else -r
        assert lib.orc_rescale_q1(v, 1024) == r % 1024


# --------------------------------------------------------------------------- NTT --

@pytest.mark.parametrize("n", [16, 64, 256, 2048])
def test_intt_inverts_the_negacyclic_ntt(n):
    lib = oc.oracle()
    psi = workload.primitive_root_2n(n)
    table = workload.inverse_table(n, psi)
    x = _u64(synthetic.uniform_below_q(n, n))
    X = np.zeros(n, np.uint64)
    lib.orc_ntt_naive(X.ctypes.data, x.ctypes.data, n, psi)
    y = X.copy()
    assert lib.orc_intt(y.ctypes.data, table.ctypes.data, n) == 0
    assert np.array_equal(y, x)
    z = np.zeros(n, np.uint64)
    lib.orc_intt_naive(z.ctypes.data, X.ctypes.data, n, psi)
    assert np.array_equal(z, x)


def test_ntt_slot_order_matches_the_automorphism_tables():
    # src/packing.rs:1082-1110: with slot j = evaluation at psi^(2 bitrev(j) + 1), the
    # automorphism X -> X^t permutes slots by table[j] = bitrev((t (2 bitrev(j)+1) mod 2n - 1) / 2)
    lib = oc.oracle()
    n = 64
    logn = 6
    psi = workload.primitive_root_2n(n)
    x = [int(v) for v in synthetic.uniform_below_q(9, n)]
    X = np.zeros(n, np.uint64)
    lib.orc_ntt_naive(X.ctypes.data, _u64(x).ctypes.data, n, psi)
    for t in (3, 5, 2 * n - 1):
        ax = [0] * n
        for i, v in enumerate(x):  # negacyclic automorphism in coefficient space
            e = (i * t) % (2 * n)
            if e < n:
                ax[e] = (ax[e] + v) % Q
            else:
                ax[e - n] = (ax[e - n] - v) % Q
        AX = np.zeros(n, np.uint64)
        lib.orc_ntt_naive(AX.ctypes.data, _u64(ax).ctypes.data, n, psi)
        br = [workload.bitrev(j, logn) for j in range(n)]
        table = [br[((t * (2 * br[j] + 1)) % (2 * n) - 1) // 2] for j in range(n)]
        assert all(int(AX[j]) == int(X[table[j]]) for j in range(n))


def _auto_tables(n):
    logn = n.bit_length() - 1
    br = [workload.bitrev(j, logn) for j in range(n)]
    tables = np.array([[br[((t * (2 * br[j] + 1)) % (2 * n) - 1) // 2] for j in range(n)]
                       for t in range(1, 2 * n, 2)], dtype=np.uint32).reshape(-1)
    gen_pows = np.array([pow(5, i, 2 * n) for i in range(n)], dtype=np.uint32)
    return tables, gen_pows


def _online_pack_by_automorphisms(y, bt, bb, n, tg, rho, num_iter, gen_pows):
    """The rotation sum of the reference's online packing (src/packing.rs:2526-2566), with each
    rotation computed without the slot tables: back to coefficients (naive inverse NTT), the
    negacyclic automorphism X -> X^t there, forward again (naive NTT)."""
    lib = oc.oracle()
    psi = workload.primitive_root_2n(n)

    def ntt(c):
        out = np.zeros(n, np.uint64)
        lib.orc_ntt_naive(out.ctypes.data, _u64(c).ctypes.data, n, psi)
        return [int(v) for v in out]

    def intt(v):
        out = np.zeros(n, np.uint64)
        lib.orc_intt_naive(out.ctypes.data, _u64(v).ctypes.data, n, psi)
        return [int(x) for x in out]

    def automorph(c, t):
        a = [0] * n
        for i, v in enumerate(c):
            e = (i * t) % (2 * n)
            if e < n:
                a[e] = (a[e] + v) % Q
            else:
                a[e - n] = (a[e - n] - v) % Q
        return a

    coeffs = [intt([int(v) for v in y[k * n:(k + 1) * n]]) for k in range(tg)]
    total = [[0] * n for _ in range(rho)]
    for i in range(num_iter):
        g = int(gen_pows[i])
        for k in range(tg):
            rot = ntt(automorph(coeffs[k], g))
            rotbar = ntt(automorph(coeffs[k], 2 * n - g))
            for o in range(rho):
                for z in range(n):
                    w = (i * tg + k) * n + z
                    total[o][z] = (total[o][z] + rot[z] * int(bt[o][w]) + rotbar[z] * int(bb[o][w])) % Q
    return np.array([v for o in range(rho) for v in total[o]], dtype=np.uint64)


@pytest.mark.parametrize("n,tg,rho", [(16, 8, 2), (32, 4, 2)])  # tg n a multiple of 128, as the engine needs
def test_online_packing_collapses_into_the_packing_matrix(n, tg, rho):
    # the per-client online loop of the reference (rotate the keys by every automorphism of
    # the packing, multiply by bold_t / bold_t_bar, add up) equals one product y . U with the
    # matrix U that sw_pack_build assembles from the same bold_t / bold_t_bar: the identity the
    # batched engine rests on (packing is linear in y and automorphisms are slot permutations)
    lib = oc.oracle()
    num_iter = n // 2 - 1
    tables, gen_pows = _auto_tables(n)
    kp = tg * n
    bt = [synthetic.uniform_below_q(300 + o, num_iter * tg * n).astype(np.uint64) for o in range(rho)]
    bb = [synthetic.uniform_below_q(400 + o, num_iter * tg * n).astype(np.uint64) for o in range(rho)]
    pt = (C.c_void_p * rho)(*[x.ctypes.data for x in bt])
    pb = (C.c_void_p * rho)(*[x.ctypes.data for x in bb])
    for seed in (1, 2):
        y = synthetic.uniform_below_q(500 + seed, kp)
        online = np.zeros(rho * n, np.uint64)
        lib.orc_pack_online(online.ctypes.data, y.ctypes.data, C.addressof(pt), C.addressof(pb), tables.ctypes.data,
                            gen_pows.ctypes.data, num_iter, tg, n, rho)
        # the rotations by coefficient-space automorphisms give the same sums (pins the slot tables'
        # use in the restatement)
        assert np.array_equal(online, _online_pack_by_automorphisms(y, bt, bb, n, tg, rho, num_iter, gen_pows))
        ub = np.zeros(4 * kp * rho * n, np.int8)
        lib.orc_pack_build(ub.ctypes.data, C.addressof(pt), C.addressof(pb), tables.ctypes.data, gen_pows.ctypes.data,
                           num_iter, tg, n, rho)
        beta = np.zeros(rho * n, np.uint64)
        ys = (C.c_void_p * 1)(y.ctypes.data)
        lib.orc_pack(beta.ctypes.data, ub.ctypes.data, C.addressof(ys), 1, kp, rho * n)
        assert np.array_equal(beta, online)
        # a planted mistake (bold_t and bold_t_bar swapped for one rotation) is seen
        bt2 = [x.copy() for x in bt]
        bt2[0][:tg * n] = bb[0][:tg * n]
        pt2 = (C.c_void_p * rho)(*[x.ctypes.data for x in bt2])
        wrong = np.zeros(rho * n, np.uint64)
        lib.orc_pack_online(wrong.ctypes.data, y.ctypes.data, C.addressof(pt2), C.addressof(pb), tables.ctypes.data,
                            gen_pows.ctypes.data, num_iter, tg, n, rho)
        assert not np.array_equal(wrong, beta)


def test_intt_rejects_non_power_of_two():
    lib = oc.oracle()
    a = np.zeros(48, np.uint64)
    assert lib.orc_intt(a.ctypes.data, a.ctypes.data, 48) == -1
    assert lib.orc_intt(a.ctypes.data, a.ctypes.data, 8) == -1


# ------------------------------------------------------------------ scan / pack --

def _qptrs(qs):
    arr = (C.c_void_p * len(qs))(*[q.ctypes.data for q in qs])
    return arr


def test_scan_brute_force_and_rescale():
    lib = oc.oracle()
    rows, cols, B, ld = 64, 48, 3, 80
    db = synthetic.uniform_bytes(1, cols * ld)
    qs = [synthetic.uniform_u32(10 + c, rows) for c in range(B)]
    qs[1][:] = 0xFFFFFFFF
    acc = np.zeros(B * cols, np.uint32)
    qp = _qptrs(qs)
    lib.orc_scan_acc(acc.ctypes.data, db.ctypes.data, ld, rows, cols, C.addressof(qp), B)
    inter = np.zeros(B * cols, np.uint64)
    lib.orc_scan(inter.ctypes.data, db.ctypes.data, ld, rows, cols, C.addressof(qp), B)
    for c in range(B):
        for col in range(cols):
            s = sum(int(db[col * ld + k]) * int(qs[c][k]) for k in range(rows)) % (1 << 32)
            assert int(acc[c * cols + col]) == s
            assert int(inter[c * cols + col]) == (s * Q + (1 << 31)) >> 32
            assert int(inter[c * cols + col]) < Q


def test_scan_is_linear_in_the_query():
    lib = oc.oracle()
    rows, cols = 256, 32
    db = synthetic.uniform_bytes(2, rows * cols)
    a, b = synthetic.uniform_u32(3, rows), synthetic.uniform_u32(4, rows)
    s = (a.astype(np.uint64) + b).astype(np.uint32)
    outs = []
    for q in (a, b, s):
        acc = np.zeros(cols, np.uint32)
        qp = _qptrs([q])
        lib.orc_scan_acc(acc.ctypes.data, db.ctypes.data, rows, rows, cols, C.addressof(qp), 1)
        outs.append(acc.astype(np.uint64))
    assert np.array_equal((outs[0] + outs[1]) % (1 << 32), outs[2])


def test_unpack_u_against_an_independent_placement():
    # The reference stores byte plane j of U[kk][col] at
    #   ((j*(CT/2) + ct/2)*KC + kc)*2048 + (ct%2)*1024 + ((kk%64)/4)*64 + 4*(col%16) + kk%4
    # (src/amx_kernels.cpp:39-41,331-336).  Build U from known values with that placement
    # written out here independently, then read it back.
    lib = oc.oracle()
    kp, cols = 128, 64
    KC, CT = kp // 64, cols // 16
    U = synthetic.uniform_u32(77, kp * cols).reshape(kp, cols)
    ub = np.zeros(4 * kp * cols, np.int8)
    for kk in range(kp):
        for col in range(cols):
            for j in range(4):
                off = (((j * (CT // 2) + (col // 16) // 2) * KC + kk // 64) * 2048 + ((col // 16) % 2) * 1024
                       + ((kk % 64) // 4) * 64 + 4 * (col % 16) + kk % 4)
                ub[off] = np.int8(((int(U[kk, col]) >> (8 * j)) & 0xFF) - 128)
    for kk in range(0, kp, 7):
        for col in range(0, cols, 3):
            assert lib.orc_unpack_u(ub.ctypes.data, kp, cols, kk, col) == int(U[kk, col])


def test_pack_brute_force():
    lib = oc.oracle()
    kp, cols, B = 128, 32, 3
    ub = synthetic.uniform_bytes(5, 4 * kp * cols).view(np.int8)
    ys = [synthetic.uniform_below_q(20 + c, kp) for c in range(B)]
    ys[2][:] = Q - 1
    beta = np.zeros(B * cols, np.uint64)
    yp = _qptrs(ys)
    lib.orc_pack(beta.ctypes.data, ub.ctypes.data, C.addressof(yp), B, kp, cols)
    for c in range(B):
        for col in range(cols):
            s = sum(int(ys[c][kk]) * lib.orc_unpack_u(ub.ctypes.data, kp, cols, kk, col) for kk in range(kp)) % Q
            assert int(beta[c * cols + col]) == s


# -------------------------------------------------------------------- validation --

def test_hand_frames_get_the_reference_statuses():
    lib = oc.oracle()
    rows, kp = 128, 256
    r0, r1 = workload.make_requests(2, rows, kp, seed=8)
    frames, want = refmodel.hand_frames(r0, r1, kp)
    for f, w in zip(frames, want):
        a = np.frombuffer(f, np.uint8)
        assert lib.orc_validate(a.ctypes.data, len(a), rows, kp) == w
        assert refmodel.decoder_accepts(f, rows, kp) == (w == 0)
    assert lib.orc_validate(None, 0, rows, kp) == 1


def test_fuzzed_frames_follow_the_decoder():
    lib = oc.oracle()
    rows, kp = 128, 256
    base = workload.make_requests(4, rows, kp, seed=9)
    accepted = 0
    for frames in refmodel.fuzz_batches(base, kp):
        for f in frames:
            a = np.frombuffer(f, np.uint8)
            s = lib.orc_validate(a.ctypes.data if len(a) else None, len(a), rows, kp)
            assert (s == 0) == refmodel.decoder_accepts(f, rows, kp)
            accepted += s == 0
    assert 0 < accepted < 512


# ------------------------------------------------------- parity with the reference --

REF_SHAPES = [
    # rows, cols, n, tg, B, ld, prefix, table
    (2048, 4096, 2048, 4, 5, 2048, "row0", "ntt"),       # tests/engine.rs:10-11
    (2048, 512, 256, 4, 9, 2048, "random", "ntt"),
    (64, 32, 16, 8, 3, 64, "random", "random"),
    (128, 64, 32, 4, 4, 192, "random", "ntt"),
]


@needs_ref
@pytest.mark.parametrize("shape", REF_SHAPES)
def test_answer_matches_the_compiled_reference(shape):
    rows, cols, n, tg, B, ld, prefix, table = shape
    inp = workload.make_inputs(rows, cols, n=n, tg=tg, seed=rows + n, ld=ld, prefix=prefix, table=table)
    reqs = workload.make_requests(B, rows, inp.kp, seed=31)
    rl = inp.response_len
    out_o = np.zeros((B, rl), np.uint8)
    st_o = np.full(B, -1, np.int32)
    assert oc.oracle_answer(inp, reqs, out_o, st_o) == B
    ref = oc.RefEngine(inp)
    for m in oc.ref_methods():
        out_r = np.zeros((B, rl), np.uint8)
        st_r = np.full(B, -1, np.int32)
        r, _ = ref.answer(reqs, out_r, st_r, m)
        assert r == B and np.array_equal(st_r, st_o)
        assert np.array_equal(out_r, out_o), f"method {m}"
    ref.close()


@needs_ref
def test_rejections_match_the_compiled_reference():
    inp = workload.make_inputs(128, 64, n=32, tg=4, seed=4, prefix="random")
    base = workload.make_requests(4, inp.rows, inp.kp, seed=41)
    ref = oc.RefEngine(inp)
    rl = inp.response_len
    frames, want = refmodel.hand_frames(base[0], base[1], inp.kp)
    batches = [frames] + refmodel.fuzz_batches(base, inp.kp, batches=2)
    for frames in batches:
        B = len(frames)
        out_o = np.full((B, rl), 0xA5, np.uint8)
        st_o = np.full(B, -1, np.int32)
        oc.oracle_answer(inp, frames, out_o, st_o)
        for m in oc.ref_methods():
            out_r = np.full((B, rl), 0xA5, np.uint8)
            st_r = np.full(B, -1, np.int32)
            ref.answer(frames, out_r, st_r, m)
            assert np.array_equal(st_r, st_o) and np.array_equal(out_r, out_o)
    ref.close()


@needs_ref
@pytest.mark.parametrize("rows,cols,ld", [(64, 32, 64), (256, 96, 320), (1024, 64, 1024), (192, 160, 200)])
def test_tile_db_matches_the_compiled_reference(rows, cols, ld):
    # orc_tile_db and the library's host sw_tile_db against the reference's own sw_tile_db
    # (src/amx_kernels.cpp:138-149), including column pitches above the record count
    from altopir_public_amd import _lib

    rlib, orc = oc.ref(), oc.oracle()
    db = synthetic.uniform_bytes(rows + cols + ld, cols * ld)
    want = np.zeros(rows * cols, np.uint8)
    rlib.sw_tile_db(want.ctypes.data, db.ctypes.data, ld, rows, cols, 3)
    got_o = np.full(rows * cols, 0x5A, np.uint8)
    orc.orc_tile_db(got_o.ctypes.data, db.ctypes.data, ld, rows, cols)
    assert np.array_equal(got_o, want)
    got_l = np.full(rows * cols, 0x5A, np.uint8)
    _lib.load().sw_tile_db(got_l.ctypes.data, db.ctypes.data, ld, rows, cols, 2)
    assert np.array_equal(got_l, want)


@needs_ref
def test_pack_build_matches_the_compiled_reference():
    lib, rlib = oc.oracle(), oc.ref()
    n, tg, rho = 32, 4, 2
    num_iter = n // 2 - 1
    logn = 5
    br = [workload.bitrev(j, logn) for j in range(n)]
    tables = np.array([[br[((t * (2 * br[j] + 1)) % (2 * n) - 1) // 2] for j in range(n)]
                       for t in range(1, 2 * n, 2)], dtype=np.uint32).reshape(-1)
    gen_pows = np.array([pow(5, i, 2 * n) for i in range(n)], dtype=np.uint32)
    bt = [synthetic.stream(50 + o, num_iter * tg * n) for o in range(rho)]
    bb = [synthetic.stream(60 + o, num_iter * tg * n) for o in range(rho)]
    pt = (C.c_void_p * rho)(*[b.ctypes.data for b in bt])
    pb = (C.c_void_p * rho)(*[b.ctypes.data for b in bb])
    size = 4 * tg * n * rho * n
    ub_o, ub_r = np.zeros(size, np.int8), np.zeros(size, np.int8)
    lib.orc_pack_build(ub_o.ctypes.data, C.addressof(pt), C.addressof(pb), tables.ctypes.data, gen_pows.ctypes.data,
                       num_iter, tg, n, rho)
    rlib.sw_pack_build(ub_r.ctypes.data, C.addressof(pt), C.addressof(pb), tables.ctypes.data, gen_pows.ctypes.data,
                       num_iter, tg, n, rho, Q, 4)
    assert np.array_equal(ub_o, ub_r)


def _matmul_inputs(rows, cols, bc, seed):
    a = synthetic.uniform_u32(seed, rows * cols)
    b = synthetic.uniform_u32(seed + 1, bc * 4 * cols)
    out = synthetic.uniform_u32(seed + 2, rows * bc + 8)
    return a, b, out


def test_matmul_vec_packed_brute_force():
    lib = oc.oracle()
    lib.orc_matmul_vec_packed.argtypes = [C.c_void_p] * 3 + [C.c_size_t] * 3
    rows, cols = 16, 5
    for bc in (1, 2, 4, 8):
        a, b, out = _matmul_inputs(rows, cols, bc, 40 + bc)
        got = out.copy()
        lib.orc_matmul_vec_packed(got.ctypes.data, a.ctypes.data, b.ctypes.data, rows, cols, bc)
        for i in range(rows):
            for c in range(bc):
                s = int(out[i * bc + c])
                for j in range(cols):
                    w = int(a[i * cols + j])
                    for t in range(4):
                        s += ((w >> (8 * t)) & 0xFF) * int(b[c * 4 * cols + 4 * j + t])
                assert int(got[i * bc + c]) == s % (1 << 32)


@pytest.mark.skipif(not (oc.ORACLE_DIR / "_ref" / "libmatmul_ref.so").exists(), reason="oracle/_ref not built")
def test_matmul_vec_packed_matches_the_compiled_reference():
    lib = oc.oracle()
    lib.orc_matmul_vec_packed.argtypes = [C.c_void_p] * 3 + [C.c_size_t] * 3
    ref = C.CDLL(str(oc.ORACLE_DIR / "_ref" / "libmatmul_ref.so"))
    rows, cols = 64, 96
    for bc, fn in ((1, "matMulVecPacked"), (2, "matMulVecPacked2"), (4, "matMulVecPacked4"), (8, "matMulVecPacked8")):
        a, b, out = _matmul_inputs(rows, cols, bc, 60 + bc)
        o1, o2 = out.copy(), out.copy()
        lib.orc_matmul_vec_packed(o1.ctypes.data, a.ctypes.data, b.ctypes.data, rows, cols, bc)
        f = getattr(ref, fn)
        f.argtypes = [C.c_void_p] * 3 + [C.c_size_t] * 2
        f(o2.ctypes.data, a.ctypes.data, b.ctypes.data, rows, cols)
        assert np.array_equal(o1, o2), fn


def _hint_by_convolution(a, db_col, ld, n, cols):
    # an independent evaluation of the hint: the full linear convolution of each pair of
    # polynomials (numpy), folded with x^n = -1, summed over the records, mod Q
    Q = 4294955009
    npolys = a.size // n
    out = np.zeros((n, cols), np.uint64)
    for col in range(cols):
        acc = np.zeros(n, dtype=object)
        for r in range(npolys):
            c = np.convolve(a[r * n:(r + 1) * n].astype(object), db_col[col * ld + r * n: col * ld + (r + 1) * n].astype(object))
            folded = c[:n].copy()
            folded[:n - 1] -= c[n:]
            acc += folded
        out[:, col] = np.array([int(v) % Q for v in acc], dtype=np.uint64)
    return out


@pytest.mark.parametrize("n, npolys, cols, ld_pad", [(16, 1, 5, 0), (32, 3, 4, 7), (64, 2, 3, 64)])
def test_hint_matches_an_independent_convolution(n, npolys, cols, ld_pad):
    Q = 4294955009
    rng = np.random.default_rng(n + npolys)
    rows = n * npolys
    ld = rows + ld_pad
    db_col = rng.integers(0, 256, size=ld * cols, dtype=np.uint8)
    a = rng.integers(0, Q, size=rows, dtype=np.uint64)
    a[:6] = [0, Q - 1, 0x7F7F7F7F, 0x7F7F7F80, 1, Q // 2]  # the centring boundaries and extremes
    got = oc.oracle_hint0(a, db_col, ld, n)
    assert np.array_equal(got, _hint_by_convolution(a, db_col, ld, n, cols))


def test_hint_of_the_polynomials_one_and_x():
    # a_r = 1: coefficient z is the sum over r of byte r n + z; a_r = x: the bytes shift up by one
    # and the last one wraps around negated (x * x^{n-1} = -1)
    Q = 4294955009
    n, npolys, cols = 32, 4, 3
    rng = np.random.default_rng(5)
    db_col = rng.integers(0, 256, size=n * npolys * cols, dtype=np.uint8)
    d = db_col.reshape(cols, npolys, n).astype(np.int64)
    one = np.zeros(n * npolys, np.uint64)
    one[::n] = 1
    assert np.array_equal(oc.oracle_hint0(one, db_col, n * npolys, n), d.sum(axis=1).T.astype(np.uint64))
    x = np.zeros(n * npolys, np.uint64)
    x[1::n] = 1
    want = np.zeros((n, cols), np.int64)
    want[1:] = d[:, :, :-1].sum(axis=1).T
    want[0] = -d[:, :, -1].sum(axis=1)
    assert np.array_equal(oc.oracle_hint0(x, db_col, n * npolys, n), (want % Q).astype(np.uint64))


def _crt_case(K, a_elems, b_cols, b_bytes, seed, m0=4294967291, m1=4294967279):
    rng = np.random.default_rng(seed)
    A = [[int(x) for x in rng.integers(0, 1 << 62, a_elems)] for _ in range(K)]
    A = [[x % (m0 * m1) for x in row] for row in A]
    a = np.array([[(x % m0) | ((x % m1) << 32) for x in row] for row in A], np.uint64)
    # 4-byte values below 2^26, so that the sums stay below 2^64 (the closed form has no wrap)
    b = rng.integers(0, 1 << min(8 * b_bytes, 26), (b_cols, a_elems), dtype=np.uint64).astype(f"<u{b_bytes}")
    c = rng.integers(0, m0 * m1, (K, b_cols), dtype=np.uint64)
    return A, a, b, c


@pytest.mark.parametrize("b_bytes,K,a_elems", [(1, 1, 64), (2, 3, 100), (4, 2, 37)])
def test_fast_batched_dot_product_is_the_crt_closed_form(b_bytes, K, a_elems):
    # src/kernel.rs:143-192 for CRT-packed a (a = A mod m0 | A mod m1 << 32) and sums below 2^64:
    # c + sum_k A b mod m0 m1, in big integers
    m0, m1 = 4294967291, 4294967279
    b_cols = 5
    A, a, b, c = _crt_case(K, a_elems, b_cols, b_bytes, 11 + b_bytes)
    want = [[(int(c[r, j]) + sum(A[r][k] * int(b[j, k]) for k in range(a_elems))) % (m0 * m1) for j in range(b_cols)]
            for r in range(K)]
    got = c.copy()
    oc.oracle().orc_fast_batched_dot_product(got.ctypes.data, b_cols, a.ctypes.data, a_elems, a_elems, b.ctypes.data,
                                             b_bytes, a_elems, b_cols, K, m0, m1)
    assert [[int(x) for x in row] for row in got] == want


def test_fast_batched_dot_product_wraps_its_partial_sums():
    # the partial sums wrap mod 2^64 before the reduction (wrapping_add, src/kernel.rs:180-181)
    m0, m1, a_elems = 1000003, 998244353, 5000
    a = np.full((1, a_elems), 0xFFFFFFFF | (0xFFFFFFFF << 32), np.uint64)
    b = np.full((1, a_elems), 0xFFFFFFFF, "<u4")
    c = np.zeros((1, 1), np.uint64)
    oc.oracle().orc_fast_batched_dot_product(c.ctypes.data, 1, a.ctypes.data, a_elems, a_elems, b.ctypes.data, 4,
                                             a_elems, 1, 1, m0, m1)
    s = (a_elems * 0xFFFFFFFF * 0xFFFFFFFF) % (1 << 64)
    lo, hi = s % m0, s % m1
    x = next(v for v in range(lo, m0 * m1, m0) if v % m1 == hi)
    assert int(c[0, 0]) == x
