/*
 * alto_oracle.c -- scalar CPU restatement of AltoPIR's batch answer engine.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP engine.  Only tests/, the smoke
 * check and the benchmark's CPU-baseline leg may load it; the product never does.
 *
 * The restatement follows the closed form of the reference engine
 * (src/amx_kernels.cpp:564-615) stage by stage.  It computes every stage directly
 * from its mathematical definition (plain sums, % for every reduction), so it shares
 * no tiling, limb or reduction tricks with either the reference or the HIP engine.
 * Loops over independent outputs are OpenMP-parallel only to keep test times short.
 */
#include "alto_oracle.h"

#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;

static uint32_t ld32(const uint8_t *p) {
    return (uint32_t)p[0] | (uint32_t)p[1] << 8 | (uint32_t)p[2] << 16 | (uint32_t)p[3] << 24;
}

/* Request length: 9 header bytes, then y and z (KP u32 each) and one u32 query value per
 * record.  src/service.rs:35-37, src/amx_kernels.cpp:562. */
size_t orc_request_len(size_t rows, size_t kp) { return 9 + 4 * (2 * kp + rows); }

/* Request validation, src/amx_kernels.cpp:571-581 (decoder semantics src/service.rs:200-225):
 *   1  null frame or length != request_len
 *   2  flag byte != 0, or either key length field != KP
 *   3  any of the 2*KP key words >= Q   (query words are never range-checked)
 *   0  valid */
int orc_validate(const uint8_t *f, size_t len, size_t rows, size_t kp) {
    if (!f || len != orc_request_len(rows, kp)) return 1;
    if (f[0] != 0 || ld32(f + 1) != kp || ld32(f + 5) != kp) return 2;
    for (size_t i = 0; i < 2 * kp; i++)
        if (ld32(f + 9 + 4 * i) >= ORC_Q) return 3;
    return 0;
}

/* One u32 of the packing matrix U[kk][col], read back from the reference's AMX B-tile
 * layout.  The layout is sw_ubt (src/amx_kernels.cpp:39-41) with the placement of
 * src/amx_kernels.cpp:331-336: byte plane j of U[kk][col] sits at
 *   sw_ubt(j, ct, kc) + r*64 + 4*nn + t,  ct = col/16, nn = col%16, kc = kk/64,
 *   r = (kk%64)/4, t = kk%4,
 * and is stored as (byte - 128). */
uint32_t orc_unpack_u(const int8_t *ub, size_t kp, size_t cols, size_t kk, size_t col) {
    const size_t KC = kp / 64, CT = cols / 16;
    const size_t ct = col / 16, nn = col % 16, kc = kk / 64, r = (kk % 64) / 4, t = kk % 4;
    uint32_t v = 0;
    for (size_t j = 0; j < 4; j++) {
        size_t base = ((j * (CT / 2) + ct / 2) * KC + kc) * 2048 + (ct % 2) * 1024;
        uint8_t b = (uint8_t)(ub[base + r * 64 + 4 * nn + t] + 128);
        v |= (uint32_t)b << (8 * j);
    }
    return v;
}

/* Scan accumulator: acc[c][col] = sum_k D[col][k] * q_c[k]  mod 2^32.
 * Closed form of the reference scan (src/amx_kernels.cpp:175-309), identical to the
 * word-wise loop of the reference server, src/server.rs:1588-1599. */
void orc_scan_acc(uint32_t *acc, const uint8_t *db_col, size_t ld, size_t rows, size_t cols,
                  const uint32_t *const *q, size_t B) {
#pragma omp parallel for schedule(static)
    for (size_t col = 0; col < cols; col++) {
        const uint8_t *d = db_col + col * ld;
        for (size_t c = 0; c < B; c++) {
            uint32_t s = 0;
            for (size_t k = 0; k < rows; k++) s += (uint32_t)d[k] * q[c][k];
            acc[c * cols + col] = s;
        }
    }
}

/* Scan with the rescale from Z_{2^32} to Z_Q:
 *   inter[c][col] = floor((acc * Q + 2^31) / 2^32)
 * src/amx_kernels.cpp:306-307 (= modswitch_w32_to_q, src/server.rs:1029-1031). */
void orc_scan(uint64_t *inter, const uint8_t *db_col, size_t ld, size_t rows, size_t cols,
              const uint32_t *const *q, size_t B) {
    uint32_t *acc = (uint32_t *)malloc(sizeof(uint32_t) * B * cols);
    orc_scan_acc(acc, db_col, ld, rows, cols, q, B);
    for (size_t i = 0; i < B * cols; i++) inter[i] = ((uint64_t)acc[i] * ORC_Q + (1ull << 31)) >> 32;
    free(acc);
}

/* Packing: beta[c][col] = sum_kk y_c[kk] * U[kk][col]  mod Q.
 * Closed form of sw_pack_impl (src/amx_kernels.cpp:345-445): the limb GEMM plus the mod-Q
 * recombination of src/amx_kernels.cpp:410-439 is exactly this sum. */
void orc_pack(uint64_t *beta, const int8_t *ub, const uint32_t *const *y, size_t B, size_t kp,
              size_t cols) {
#pragma omp parallel
    {
        uint32_t *ucol = (uint32_t *)malloc(sizeof(uint32_t) * kp);
#pragma omp for schedule(static)
        for (size_t col = 0; col < cols; col++) {
            for (size_t kk = 0; kk < kp; kk++) ucol[kk] = orc_unpack_u(ub, kp, cols, kk, col);
            for (size_t c = 0; c < B; c++) {
                u128 s = 0;
                for (size_t kk = 0; kk < kp; kk++) s += (u128)((uint64_t)y[c][kk] * ucol[kk]);
                beta[c * cols + col] = (uint64_t)(s % ORC_Q);
            }
        }
        free(ucol);
    }
}

uint64_t orc_mulmod(uint64_t a, uint64_t b) { return (uint64_t)(((u128)a * b) % ORC_Q); }

uint64_t orc_powmod(uint64_t b, uint64_t e) {
    uint64_t r = 1;
    b %= ORC_Q;
    while (e) {
        if (e & 1) r = orc_mulmod(r, b);
        b = orc_mulmod(b, b);
        e >>= 1;
    }
    return r;
}

/* Tail, pointwise part: a[z] = beta[z] + sum_{k<tg} z[k*n+z] * that[k*n+z]  mod Q.
 * src/amx_kernels.cpp:500-511; every operand is reduced mod Q first, as mq_red64 does
 * (src/amx_kernels.cpp:447-452). */
void orc_tail_pointwise(uint64_t *a, const uint64_t *beta, const uint32_t *z, const uint64_t *that,
                        size_t tg, size_t n) {
    for (size_t i = 0; i < n; i++) {
        uint64_t s = beta[i] % ORC_Q;
        for (size_t k = 0; k < tg; k++)
            s = (s + orc_mulmod(z[k * n + i] % ORC_Q, that[k * n + i] % ORC_Q)) % ORC_Q;
        a[i] = s;
    }
}

/* Gentleman-Sande inverse NTT over Z_Q, src/amx_kernels.cpp:453-498.
 * Stage sizes t = 1, 2, 4, ..., n/2; at stage t there are h = n/(2t) groups, group i holds
 * the pairs (2t*i + j, 2t*i + j + t), j < t, and uses w = table[h + i].  Butterfly
 * (mq_bfly, src/amx_kernels.cpp:453-460):
 *   x' = (x + y) / 2 mod Q
 *   y' = (x - y) * w mod Q,  where only the low 32 bits of w take part (_mm512_mul_epu32).
 * Inputs are reduced mod Q first (src/amx_kernels.cpp:465-466).  Returns -1 unless n is a
 * power of two >= 16 (src/amx_kernels.cpp:463). */
int orc_intt(uint64_t *a, const uint64_t *table, size_t n) {
    if (n < 16 || (n & (n - 1))) return -1;
    const uint64_t half_inv = (ORC_Q + 1) / 2; /* 2^-1 mod Q */
    for (size_t i = 0; i < n; i++) a[i] %= ORC_Q;
    for (size_t t = 1; t < n; t *= 2) {
        const size_t h = n / (2 * t);
        for (size_t i = 0; i < h; i++) {
            const uint64_t w = table[h + i] & 0xffffffffull;
            for (size_t j = 0; j < t; j++) {
                uint64_t x = a[2 * t * i + j], y = a[2 * t * i + j + t];
                a[2 * t * i + j] = orc_mulmod((x + y) % ORC_Q, half_inv);
                a[2 * t * i + j + t] = orc_mulmod((x + ORC_Q - y) % ORC_Q, w);
            }
        }
    }
    return 0;
}

static size_t bitrev(size_t x, size_t bits) {
    size_t r = 0;
    for (size_t i = 0; i < bits; i++) { r = (r << 1) | (x & 1); x >>= 1; }
    return r;
}

/* Negacyclic NTT with spiral-rs's slot order: slot j holds the evaluation at
 * psi^(2*bitrev(j)+1) (stated at src/packing.rs:1078-1080).  O(n^2); pins orc_intt. */
void orc_ntt_naive(uint64_t *out, const uint64_t *in, size_t n, uint64_t psi) {
    size_t logn = 0;
    while (((size_t)1 << logn) < n) logn++;
    for (size_t j = 0; j < n; j++) {
        uint64_t pt = orc_powmod(psi, 2 * bitrev(j, logn) + 1), acc = 0, pw = 1;
        for (size_t i = 0; i < n; i++) {
            acc = (acc + orc_mulmod(in[i] % ORC_Q, pw)) % ORC_Q;
            pw = orc_mulmod(pw, pt);
        }
        out[j] = acc;
    }
}

/* Inverse of orc_ntt_naive: a_i = n^-1 sum_j A_j psi^-(i*(2*bitrev(j)+1)). */
void orc_intt_naive(uint64_t *out, const uint64_t *in, size_t n, uint64_t psi) {
    size_t logn = 0;
    while (((size_t)1 << logn) < n) logn++;
    const uint64_t psi_inv = orc_powmod(psi, ORC_Q - 2), n_inv = orc_powmod(n, ORC_Q - 2);
    for (size_t i = 0; i < n; i++) out[i] = 0;
    for (size_t j = 0; j < n; j++) {
        uint64_t pt = orc_powmod(psi_inv, 2 * bitrev(j, logn) + 1), pw = 1;
        for (size_t i = 0; i < n; i++) {
            out[i] = (out[i] + orc_mulmod(in[j] % ORC_Q, pw)) % ORC_Q;
            pw = orc_mulmod(pw, pt);
        }
    }
    for (size_t i = 0; i < n; i++) out[i] = orc_mulmod(out[i], n_inv);
}

/* Rescale of one row-1 value from Z_Q to Z_q1 with rounding after centring, as the
 * reference engine computes it (src/amx_kernels.cpp:522-529):
 *   c = v - Q if v >= floor(Q/2) else v
 *   r = trunc((c*q1 +- floor(Q/2)) / Q)   (sign of c; exact integer division here,
 *       where the reference divides in double precision, which is exact for these
 *       magnitudes)
 *   r += q1 if r < 0;  r -= q1 if r >= q1 */
uint64_t orc_rescale_q1(uint64_t v, uint64_t q1) {
    const int64_t Q = (int64_t)ORC_Q, half = (int64_t)(ORC_Q / 2);
    int64_t c = (int64_t)v;
    int neg = v >= (uint64_t)half;
    if (neg) c -= Q;
    int64_t x = c * (int64_t)q1 + (neg ? -half : half);
    int64_t t = x / Q;
    if (t < 0) t += (int64_t)q1;
    if (t >= (int64_t)q1) t -= (int64_t)q1;
    return (uint64_t)t;
}

/* rescale_i64 of src/modulus_switch.rs:120-134, restated as written (used to pin
 * orc_rescale_q1 exhaustively). */
uint64_t orc_rescale_i64(uint64_t a, uint64_t inp_mod, uint64_t out_mod) {
    const int64_t im = (int64_t)inp_mod, om = (int64_t)out_mod;
    int64_t iv = (int64_t)(a % inp_mod);
    if (iv >= im / 2) iv -= im;
    const int64_t sign = iv >= 0 ? 1 : -1;
    const int64_t val = iv * om;
    int64_t result = (val + sign * (im / 2)) / im;
    result = (result + (int64_t)((inp_mod / out_mod) * out_mod) + 2 * om) % om;
    return (uint64_t)((result + om) % om);
}

/* LSB-first bit writer, src/bits.rs:11-41: ORs num_bits bits of val at bit_offs, never
 * past the end of the buffer. */
void orc_write_bits(uint8_t *data, size_t len, uint64_t val, size_t bit_offs, size_t num_bits) {
    size_t byte_index = bit_offs / 8, bit_index = bit_offs % 8;
    while (num_bits > 0 && byte_index < len) {
        size_t w = 8 - bit_index < num_bits ? 8 - bit_index : num_bits;
        uint64_t mask = (1ull << w) - 1;
        data[byte_index] |= (uint8_t)((val & mask) << bit_index);
        num_bits -= w;
        bit_index += w;
        if (bit_index == 8) { byte_index++; bit_index = 0; }
        val >>= w;
    }
}

/* Row-1 switch and bit packing, src/amx_kernels.cpp:515-537:
 * v = inter + coef mod Q, r = orc_rescale_q1(v), and r << (p % 8) is ORed into the four
 * bytes starting at p / 8 (those inside the part), p = bit_off + z * q1_bits.
 * Returns -1 on the configurations the reference rejects (src/amx_kernels.cpp:517). */
int orc_switch_row1(uint8_t *res, size_t res_len, size_t bit_off, const uint64_t *inter,
                    const uint64_t *coef, size_t n, uint64_t q1, int q1_bits) {
    if (q1 >= (1ull << 21) || n % 8 || (bit_off + n * (size_t)q1_bits + 7) / 8 > res_len ||
        q1_bits > 24)
        return -1;
    for (size_t z = 0; z < n; z++) {
        uint64_t v = (inter[z] % ORC_Q + coef[z] % ORC_Q) % ORC_Q;
        uint64_t r = orc_rescale_q1(v, q1);
        size_t p = bit_off + z * (size_t)q1_bits;
        uint32_t w = (uint32_t)r << (p % 8);
        for (size_t k = 0; k < 4 && p / 8 + k < res_len; k++) res[p / 8 + k] |= (uint8_t)(w >> (8 * k));
    }
    return 0;
}

/* Whole batch, src/amx_kernels.cpp:564-615: validate every request, answer the valid
 * ones in order, leave rejected outputs untouched.  Returns the number answered, or -3
 * when the tail configuration is unsupported (n not a power of two >= 16, or the
 * switch checks of src/amx_kernels.cpp:517). */
int orc_answer(const orc_engine *e, const uint8_t *const *req, const size_t *len, size_t B,
               uint8_t *const *out, int *status) {
    const size_t kp = e->tg * e->n, C = e->cols, n = e->n, O = e->cols / e->n;
    size_t Bv = 0;
    for (size_t c = 0; c < B; c++) {
        status[c] = orc_validate(req[c], len[c], e->rows, kp);
        Bv += status[c] == 0;
    }
    if (!Bv) return 0;
    const uint32_t **q = (const uint32_t **)malloc(sizeof(void *) * Bv);
    const uint32_t **y = (const uint32_t **)malloc(sizeof(void *) * Bv);
    const uint32_t **z = (const uint32_t **)malloc(sizeof(void *) * Bv);
    uint8_t **o = (uint8_t **)malloc(sizeof(void *) * Bv);
    uint32_t *words = (uint32_t *)malloc(sizeof(uint32_t) * Bv * (2 * kp + e->rows));
    size_t v = 0;
    for (size_t c = 0; c < B; c++) {
        if (status[c]) continue;
        uint32_t *w = words + v * (2 * kp + e->rows);
        for (size_t i = 0; i < 2 * kp + e->rows; i++) w[i] = ld32(req[c] + 9 + 4 * i);
        y[v] = w; z[v] = w + kp; q[v] = w + 2 * kp; o[v] = out[c];
        v++;
    }
    uint64_t *inter = (uint64_t *)malloc(sizeof(uint64_t) * Bv * C);
    uint64_t *beta = (uint64_t *)malloc(sizeof(uint64_t) * Bv * C);
    orc_scan(inter, e->db_col, e->ld, e->rows, C, q, Bv);
    orc_pack(beta, e->ub, y, Bv, kp, C);
    int bad = (n < 16 || (n & (n - 1))) ||
              e->q1 >= (1ull << 21) || n % 8 ||
              (e->row1_off + n * (size_t)e->q1_bits + 7) / 8 > e->part_len || e->q1_bits > 24;
    if (!bad) {
#pragma omp parallel for schedule(static)
        for (size_t pi = 0; pi < Bv * O; pi++) {
            const size_t c = pi / O, oo = pi % O, off = c * C + oo * n;
            uint64_t *a = (uint64_t *)malloc(sizeof(uint64_t) * n);
            uint8_t *res = o[c] + oo * e->part_len;
            memcpy(res, e->prefix[oo], e->part_len);
            orc_tail_pointwise(a, beta + off, z[c], e->that[oo], e->tg, n);
            orc_intt(a, e->inv_table, n);
            orc_switch_row1(res, e->part_len, e->row1_off, inter + off, a, n, e->q1, e->q1_bits);
            free(a);
        }
    }
    free(inter); free(beta); free(words); free(q); free(y); free(z); free(o);
    return bad ? -3 : (int)Bv;
}

/* Packing-matrix build (setup), src/amx_kernels.cpp:311-341.  For output column
 * oz = (o, z) and k < tg:
 *   U[k*n + zp][oz] += bold_t[o][(i*tg + k)*n + z]
 *   U[k*n + zb][oz] += bold_t_bar[o][(i*tg + k)*n + z]       for i < num_iter,
 * with g = gen_pows[i], zp = tables[((g-1)/2)*n + z], zb = tables[((2n-g-1)/2)*n + z];
 * the u64 sums wrap, then are reduced mod Q and stored in the tile layout of
 * orc_unpack_u. */
void orc_pack_build(int8_t *ub, const uint64_t *const *bold_t, const uint64_t *const *bold_t_bar,
                    const uint32_t *tables, const uint32_t *gen_pows, size_t num_iter, size_t tg,
                    size_t n, size_t rho) {
    const size_t K = tg * n, KC = K / 64, C = rho * n, CT = C / 16;
#pragma omp parallel
    {
        uint64_t *col = (uint64_t *)malloc(sizeof(uint64_t) * K);
#pragma omp for schedule(static)
        for (size_t oz = 0; oz < C; oz++) {
            const size_t o = oz / n, z = oz % n;
            memset(col, 0, sizeof(uint64_t) * K);
            for (size_t i = 0; i < num_iter; i++) {
                const uint32_t g = gen_pows[i];
                const size_t zp = tables[((g - 1) / 2) * n + z], zb = tables[((2 * n - g - 1) / 2) * n + z];
                for (size_t k = 0; k < tg; k++) {
                    col[k * n + zp] += bold_t[o][(i * tg + k) * n + z];
                    col[k * n + zb] += bold_t_bar[o][(i * tg + k) * n + z];
                }
            }
            const size_t ct = oz / 16, nn = oz % 16;
            for (size_t kk = 0; kk < K; kk++) {
                const uint32_t v = (uint32_t)(col[kk] % ORC_Q);
                const size_t kc = kk / 64, r = (kk % 64) / 4, t = kk % 4;
                for (size_t j = 0; j < 4; j++) {
                    size_t base = ((j * (CT / 2) + ct / 2) * KC + kc) * 2048 + (ct % 2) * 1024;
                    ub[base + r * 64 + 4 * nn + t] = (int8_t)((int)((v >> (8 * j)) & 0xff) - 128);
                }
            }
        }
        free(col);
    }
}

/* The reference's per-client online packing, before it collapses into one matrix: the sum
 * over the packing rotations of src/packing.rs:2526-2566 (full_packing_with_preprocessing_online),
 * for output o and NTT slot z of one client,
 *   sum_o[z] = sum_{i < num_iter} sum_{k < tg}  rot_i,k[z] * bold_t[o][(i tg + k) n + z]
 *                                              + rotbar_i,k[z] * bold_t_bar[o][(i tg + k) n + z]  (mod Q)
 * where rot_i,k and rotbar_i,k are key polynomial y_k (NTT slots, y[k n .. k n + n)) under the
 * automorphisms tau_g and tau_{2n-g}, g = gen_pows[i] (generate_rotations_double,
 * src/packing.rs:1794-1816): in the NTT domain an automorphism permutes slots,
 * out[j] = poly[table_t[j]] with table_t = tables[(t - 1) / 2] (apply_automorph_ntt_raw,
 * src/packing.rs:1112-1126).  The rotations are materialised first and then multiplied
 * pointwise, term by term, in the order of the reference loop; every product is reduced mod Q.
 * The z_body x bold_t_hat term of the same sum (src/packing.rs:2556-2566) is the engine's tail
 * key product and is not part of this sum.  beta[o * n + z] receives sum_o[z]. */
void orc_pack_online(uint64_t *beta, const uint32_t *y, const uint64_t *const *bold_t,
                     const uint64_t *const *bold_t_bar, const uint32_t *tables, const uint32_t *gen_pows,
                     size_t num_iter, size_t tg, size_t n, size_t rho) {
    uint64_t *rot = (uint64_t *)malloc(sizeof(uint64_t) * tg * n);
    uint64_t *rotbar = (uint64_t *)malloc(sizeof(uint64_t) * tg * n);
    uint64_t *sum = (uint64_t *)malloc(sizeof(uint64_t) * rho * n);
    memset(sum, 0, sizeof(uint64_t) * rho * n);
    for (size_t i = 0; i < num_iter; i++) {
        const size_t g = gen_pows[i];
        const uint32_t *tab = tables + ((g - 1) / 2) * n, *tabbar = tables + ((2 * n - g - 1) / 2) * n;
        for (size_t k = 0; k < tg; k++)
            for (size_t j = 0; j < n; j++) {
                rot[k * n + j] = y[k * n + tab[j]] % ORC_Q;
                rotbar[k * n + j] = y[k * n + tabbar[j]] % ORC_Q;
            }
        for (size_t o = 0; o < rho; o++)
            for (size_t k = 0; k < tg; k++)
                for (size_t z = 0; z < n; z++) {
                    const size_t w = (i * tg + k) * n + z;
                    sum[o * n + z] = (sum[o * n + z] + orc_mulmod(rot[k * n + z], bold_t[o][w] % ORC_Q)) % ORC_Q;
                    sum[o * n + z] = (sum[o * n + z] + orc_mulmod(rotbar[k * n + z], bold_t_bar[o][w] % ORC_Q)) % ORC_Q;
                }
    }
    memcpy(beta, sum, sizeof(uint64_t) * rho * n);
    free(rot);
    free(rotbar);
    free(sum);
}

/* Database tiling for the reference's AMX path (setup), src/amx_kernels.cpp:138-149 with
 * sw_dbt (:42): 16-column x 64-record tiles, row r of a tile = 64 record bytes of column
 * 16*cb + r, tiles of column-tile pairs interleaved per 64-record chunk. */
void orc_tile_db(uint8_t *dst, const uint8_t *db, size_t ld, size_t rows, size_t cols) {
    const size_t KC = rows / 64;
    for (size_t cb = 0; cb < cols / 16; cb++)
        for (size_t r = 0; r < 16; r++)
            for (size_t kc = 0; kc < KC; kc++)
                memcpy(dst + ((cb / 2) * KC + kc) * 2048 + (cb % 2) * 1024 + r * 64,
                       db + (cb * 16 + r) * ld + kc * 64, 64);
}

/* Exhaustive pin of orc_rescale_q1: counts v in [lo, hi) where it differs from
 * orc_rescale_i64(v, Q, q1), the formula of src/modulus_switch.rs:120-134. */
uint64_t orc_count_rescale_mismatch(uint64_t lo, uint64_t hi, uint64_t q1) {
    uint64_t bad = 0;
#pragma omp parallel for reduction(+ : bad) schedule(static)
    for (uint64_t v = lo; v < hi; v++) bad += orc_rescale_q1(v, q1) != orc_rescale_i64(v, ORC_Q, q1);
    return bad;
}

/* SimplePIR packed product, src/matmul.cpp:105-218 (bc = 1) and its 2/4/8-column forms
 * (:220-388, :391-674, :1548-1559): b column c starts at b + c * 4 * cols, and
 * out[i * bc + c] += sum_j sum_{t<4} byte_t(a[i * cols + j]) * b_c[4 j + t]  (mod 2^32). */
void orc_matmul_vec_packed(uint32_t *out, const uint32_t *a, const uint32_t *b, size_t rows, size_t cols,
                           size_t bc) {
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < rows; i++)
        for (size_t c = 0; c < bc; c++) {
            const uint32_t *bcol = b + c * 4 * cols;
            uint32_t s = 0;
            for (size_t j = 0; j < cols; j++) {
                const uint32_t w = a[i * cols + j];
                for (size_t t = 0; t < 4; t++) s += ((w >> (8 * t)) & 0xffu) * bcol[4 * j + t];
            }
            out[i * bc + c] += s;
        }
}

/* The offline hint, src/server.rs:656-709 (multiply_with_db_ring) for the columns
 * [col0, col0 + ncols): for each column, the sum over the record polynomials r of the product
 * of query polynomial a_r (a[r n .. r n + n), after negacyclic_perm, in coefficient form) with
 * the column's record bytes d_r[j] = db_col[col * ld + r n + j], in Z_Q[x]/(x^n + 1); the
 * reference multiplies in the NTT domain and returns the coefficients, here the negacyclic
 * product is summed term by term (x^{i + j} = -x^{i + j - n} past degree n - 1).
 * hint[z * stride + (col - col0)] receives coefficient z (src/server.rs:702-707 transposes
 * the per-column results into z-major order). */
void orc_hint0(uint64_t *hint, size_t stride, const uint64_t *a, size_t npolys, size_t n, const uint8_t *db_col,
               size_t ld, size_t col0, size_t ncols) {
#pragma omp parallel for collapse(2) schedule(dynamic)
    for (size_t cc = 0; cc < ncols; cc++) {
        for (size_t z = 0; z < n; z++) {
            const uint8_t *d = db_col + (col0 + cc) * ld;
            uint64_t pos = 0, neg = 0; /* each < 2^61: R terms of at most 2^40 */
            for (size_t r = 0; r < npolys; r++) {
                const uint64_t *ar = a + r * n;
                const uint8_t *dr = d + r * n;
                for (size_t j = 0; j <= z; j++) pos += (ar[z - j] % ORC_Q) * dr[j];
                for (size_t j = z + 1; j < n; j++) neg += (ar[n + z - j] % ORC_Q) * dr[j];
            }
            hint[z * stride + cc] = (pos % ORC_Q + ORC_Q - neg % ORC_Q) % ORC_Q;
        }
    }
}

/* ------------------------------------------------------------------------------------- */
/* Offline packing precomputation (InspiRING) of one output part.                          */

/* Forward negacyclic NTT in place, Cooley-Tukey with the powers psi^bitrev(k) (k = m + i in
 * stage m), the output in the slot order of orc_ntt_naive (slot j = evaluation at
 * psi^(2 bitrev(j) + 1)): the transform whose inverse orc_intt computes with the table
 * psi^-bitrev(k) / 2.  The reference takes it from spiral-rs (PolyMatrixRaw::ntt); it is
 * pinned against orc_ntt_naive. */
void orc_ntt(uint64_t *a, size_t n, uint64_t psi) {
    size_t logn = 0;
    while (((size_t)1 << logn) < n) logn++;
    for (size_t i = 0; i < n; i++) a[i] %= ORC_Q;
    size_t t = n;
    for (size_t m = 1; m < n; m *= 2) {
        t /= 2;
        for (size_t i = 0; i < m; i++) {
            const uint64_t s = orc_powmod(psi, bitrev(m + i, logn));
            for (size_t j = 2 * i * t; j < 2 * i * t + t; j++) {
                const uint64_t u = a[j], v = orc_mulmod(a[j + t], s);
                a[j] = (u + v) % ORC_Q;
                a[j + t] = (u + ORC_Q - v) % ORC_Q;
            }
        }
    }
}

/* Slot permutation of the automorphism x -> x^t (t odd) in the NTT domain, the closed form
 * of generate_automorph_tables_brute_force (src/packing.rs:1082-1110):
 * tau_t(a)[j] = a[table_t[j]], table_t[j] = bitrev(((t (2 bitrev(j) + 1)) mod 2n - 1) / 2). */
size_t orc_auto_slot(size_t t, size_t j, size_t n) {
    size_t logn = 0;
    while (((size_t)1 << logn) < n) logn++;
    return bitrev(((t * (2 * bitrev(j, logn) + 1)) % (2 * n) - 1) / 2, logn);
}

/* Coefficients of a (NTT slots) with the table psi^-bitrev(k) / 2 (orc_intt). */
static void intt_psi(uint64_t *a, size_t n, uint64_t psi) {
    size_t logn = 0;
    while (((size_t)1 << logn) < n) logn++;
    uint64_t *tab = malloc(n * sizeof *tab);
    const uint64_t psi_inv = orc_powmod(psi, ORC_Q - 2), half = (ORC_Q + 1) / 2;
    for (size_t k = 0; k < n; k++) tab[k] = orc_mulmod(orc_powmod(psi_inv, bitrev(k, logn)), half);
    orc_intt(a, tab, n);
    free(tab);
}

/* Gadget decomposition of a coefficient polynomial c (values < Q) into tg digit polynomials
 * of base 2^ceil(32 / tg), least significant first (spiral-rs gadget_invert_transposed,
 * called at src/packing.rs:2323/2335/2343 on the raw polynomial), then each digit polynomial
 * to NTT slots: out[k * n + slot]. */
static void gadget_ntt(uint64_t *out, const uint64_t *c, size_t n, size_t tg, uint64_t psi) {
    const unsigned bits = (unsigned)((32 + tg - 1) / tg);
    const uint64_t mask = (bits >= 64) ? ~0ull : ((1ull << bits) - 1);
    for (size_t k = 0; k < tg; k++) {
        for (size_t z = 0; z < n; z++) out[k * n + z] = (k * bits < 64) ? (c[z] >> (k * bits)) & mask : 0;
        orc_ntt(out + k * n, n, psi);
    }
}

/* The offline packing precomputation of output part `part` (columns col0 = part n .. + n):
 *   src/packing.rs:910-942 (prep_pack_lwes): the a-polynomial of column j is
 *     a~_j = NTT(negacyclic_perm(hint[. , j]))  (hint row z, coefficient z);
 *   :2260-2291 (precompute_aggregates_ntt): E[v][slot] = NTT(row_v)[slot] with
 *     row_v[j] = a~_j[v], and for i < n/2, g = 5^i mod 2n:
 *     r_i[slot] = n^-1 E[table_g[slot]][slot],  rbar_i[slot] = n^-1 E[table_{2n-g}[slot]][slot];
 *   :2301-2359 (full_packing_offline_mode), walking i = n/2 - 2 .. 0:
 *     bold_t[i] = NTT(G^-1(INTT(r_{i+1}))),  r_i += sum_k tau_g(w_k) . bold_t[i][k],
 *     bold_t_bar[i] = NTT(G^-1(INTT(rbar_{i+1}))),  rbar_i += sum_k tau_{2n-g}(w_k) . bold_t_bar[i][k],
 *   then bold_t_hat = NTT(G^-1(INTT(rbar_0))), r_0 += sum_k v_k . bold_t_hat[k],
 *   a_hat = INTT(r_0) (coefficients).
 * w, v: tg x n mask slots (the reference draws them from W_SEED / V_SEED, src/scheme.rs:24-31).
 * Outputs: bold_t, bold_t_bar ((n/2 - 1) tg n values, [(i tg + k) n + slot]), bold_t_hat
 * (tg n), a_hat (n).  hint[z * stride + col0 + j].  Returns 0, -1 for n not a power of two
 * of at least 16. */
int orc_inspiring_offline(uint64_t *bold_t, uint64_t *bold_t_bar, uint64_t *bold_t_hat, uint64_t *a_hat,
                          const uint64_t *hint, size_t stride, size_t col0, const uint64_t *w,
                          const uint64_t *v, size_t n, size_t tg, uint64_t psi) {
    if (n < 16 || (n & (n - 1))) return -1;
    const size_t half = n / 2;
    uint64_t *at = malloc(n * n * sizeof *at);   /* a~_j[slot] at [j n + slot] */
    uint64_t *ev = malloc(n * n * sizeof *ev);   /* E[v][slot] */
    uint64_t *r = malloc(half * n * sizeof *r), *rb = malloc(half * n * sizeof *rb);
    uint64_t *c = malloc(n * sizeof *c);
    for (size_t j = 0; j < n; j++) {
        uint64_t *p = at + j * n;
        for (size_t z = 0; z < n; z++) {  /* negacyclic_perm (src/util.rs:5-21) */
            const uint64_t h = hint[(z == 0 ? 0 : n - z) * stride + col0 + j] % ORC_Q;
            p[z] = z == 0 ? h : (ORC_Q - h) % ORC_Q;
        }
        orc_ntt(p, n, psi);
    }
    for (size_t vv = 0; vv < n; vv++) {
        uint64_t *e = ev + vv * n;
        for (size_t j = 0; j < n; j++) e[j] = at[j * n + vv];
        orc_ntt(e, n, psi);
    }
    const uint64_t ninv = orc_powmod(n, ORC_Q - 2);
    uint64_t g = 1;
    for (size_t i = 0; i < half; i++) {
        for (size_t s = 0; s < n; s++) {
            r[i * n + s] = orc_mulmod(ninv, ev[orc_auto_slot(g, s, n) * n + s]);
            rb[i * n + s] = orc_mulmod(ninv, ev[orc_auto_slot(2 * n - g, s, n) * n + s]);
        }
        g = g * 5 % (2 * n);
    }
    size_t *gp = malloc(half * sizeof *gp);  /* gen_pows, src/packing.rs:1695-1703 */
    gp[0] = 1;
    for (size_t i = 1; i < half; i++) gp[i] = gp[i - 1] * 5 % (2 * n);
    for (size_t ii = half - 1; ii-- > 0;) {
        for (int bar = 0; bar < 2; bar++) {
            uint64_t *src = (bar ? rb : r) + (ii + 1) * n, *dst = (bar ? rb : r) + ii * n;
            uint64_t *bt = (bar ? bold_t_bar : bold_t) + ii * tg * n;
            const size_t t = bar ? 2 * n - gp[ii] : gp[ii];
            memcpy(c, src, n * sizeof *c);
            intt_psi(c, n, psi);
            gadget_ntt(bt, c, n, tg, psi);
            for (size_t k = 0; k < tg; k++)
                for (size_t s = 0; s < n; s++)
                    dst[s] = (dst[s] + orc_mulmod(w[k * n + orc_auto_slot(t, s, n)] % ORC_Q, bt[k * n + s])) % ORC_Q;
        }
    }
    memcpy(c, rb, n * sizeof *c);
    intt_psi(c, n, psi);
    gadget_ntt(bold_t_hat, c, n, tg, psi);
    for (size_t k = 0; k < tg; k++)
        for (size_t s = 0; s < n; s++) r[s] = (r[s] + orc_mulmod(v[k * n + s] % ORC_Q, bold_t_hat[k * n + s])) % ORC_Q;
    memcpy(a_hat, r, n * sizeof *a_hat);
    intt_psi(a_hat, n, psi);
    free(at); free(ev); free(r); free(rb); free(c); free(gp);
    return 0;
}

/* The two-modulus batched dot product of src/kernel.rs:143-192 (the reference's scalar form):
 * for batch row r < K and column j, the wrapping u64 sums of (a & 0xffffffff) b and (a >> 32) b
 * over k, reduced mod m0 and m1 (barrett_coeff_u64), composed by CRT into Z_{m0 m1}
 * (crt_compose_2) and added to c mod m0 m1 (barrett_u64).  b elements of b_bytes bytes. */
void orc_fast_batched_dot_product(uint64_t *c, size_t c_stride, const uint64_t *a, size_t a_stride,
                                  size_t a_elems, const void *b_t, int b_bytes, size_t b_rows, size_t b_cols,
                                  size_t K, uint64_t m0, uint64_t m1) {
    uint64_t inv = 0;  /* m0^-1 mod m1 (extended Euclid: m1 need not be prime) */
    {
        __int128 t0 = 0, t1 = 1, r0 = m1, r1 = m0 % m1;
        while (r1) {
            __int128 q = r0 / r1, r2 = r0 - q * r1, t2 = t0 - q * t1;
            r0 = r1; r1 = r2; t0 = t1; t1 = t2;
        }
        inv = (uint64_t)(((t0 % (__int128)m1) + m1) % m1);
    }
    const u128 M = (u128)m0 * m1;
    for (size_t r = 0; r < K; r++)
        for (size_t j = 0; j < b_cols; j++) {
            uint64_t lo = 0, hi = 0;
            for (size_t k = 0; k < a_elems; k++) {
                const size_t bi = j * b_rows + k;
                const uint64_t bv = b_bytes == 1 ? ((const uint8_t *)b_t)[bi]
                                  : b_bytes == 2 ? ((const uint16_t *)b_t)[bi] : ((const uint32_t *)b_t)[bi];
                const uint64_t av = a[r * a_stride + k];
                lo += (av & 0xffffffffull) * bv;
                hi += (av >> 32) * bv;
            }
            const uint64_t l = lo % m0, h = hi % m1;
            const uint64_t t = (uint64_t)((u128)((h + m1 - l % m1) % m1) * inv % m1);
            const u128 res = (u128)l + (u128)m0 * t;
            c[r * c_stride + j] = (uint64_t)(((u128)c[r * c_stride + j] + res) % M);
        }
}
