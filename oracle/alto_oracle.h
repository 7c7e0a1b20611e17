/*
 * alto_oracle.h -- CPU restatement of AltoPIR's batch answer engine.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP engine is
 * compared against: only the test suite, the smoke check and the CPU-baseline leg
 * of the benchmark may load it.  The product path never calls it and has no CPU
 * fallback.
 *
 * Every function restates one piece of the reference engine
 * (src/amx_kernels.cpp, src/service.rs, src/modulus_switch.rs, src/bits.rs) in
 * plain scalar C, written for clarity rather than speed.  The citations are given
 * per function in alto_oracle.c.
 *
 * Pinning (what checks each function against something other than itself):
 *   orc_validate      the status vector of tests/engine.rs:130-148 and the mutate
 *                     fuzz model, plus byte parity with the compiled reference
 *   orc_scan          brute-force u64 arithmetic on tiny shapes, linearity,
 *                     special inputs, compiled reference
 *   orc_pack          brute-force u128 arithmetic, compiled reference
 *   orc_intt          a naive O(n^2) negacyclic inverse transform built from a
 *                     primitive 2n-th root of unity, INTT(NTT(x)) == x, compiled
 *                     reference
 *   orc_rescale_q1    exhaustive comparison with orc_rescale_i64 (the
 *                     modulus_switch.rs formula) over every v in [0, Q)
 *   orc_write_bits    the golden bit patterns of src/bits.rs:115-133
 *   orc_pack_build    compiled reference (sw_pack_build); with orc_pack, against
 *                     orc_pack_online (the per-client rotation loop of the reference's
 *                     online packing, which never builds the matrix)
 *   orc_pack_online   the automorphism tables' closed form (pinned against NTT slot
 *                     evaluation) and a brute-force evaluation of each automorphism on
 *                     the coefficient polynomial
 *   orc_answer        compiled reference, byte for byte, on seeded inputs
 *   orc_matmul_vec_packed  compiled reference (src/matmul.cpp) and brute force
 *   orc_hint0         an independent evaluation (full linear convolution of each pair of
 *                     polynomials, then the fold x^n = -1), and the closed forms for the query
 *                     polynomials 1 and x (sums of record bytes, shifted with one negated term)
 *   orc_ntt           orc_ntt_naive (slot evaluation), and orc_intt inverts it
 *   orc_fast_batched_dot_product  the closed form c + sum_k A b mod m0 m1 for CRT-packed
 *                     A (big integers), and the u64 wrap of the partial sums
 *   orc_inspiring_offline  the aggregates against the reference's direct form (sums of
 *                     monomial products, src/packing.rs:2179-2258, evaluated term by term);
 *                     the whole precomputation by decryption: responses the engine computes
 *                     from it decrypt, with the client's secret, to the database records
 */
#ifndef ALTO_ORACLE_H
#define ALTO_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_Q 4294955009ull /* 2^32 - 12287, src/service.rs:25 */

/* Everything the answer engine borrows at construction (src/amx_kernels.cpp:539-560). */
typedef struct {
    const uint8_t *db_col;        /* column-major database, db_col[col * ld + k], k < rows */
    size_t ld, rows, cols;
    const int8_t *ub;             /* packing matrix in the reference's tile layout */
    size_t n, tg;                 /* ring dimension, gadget digits (KP = tg * n) */
    const uint8_t *const *prefix; /* cols / n buffers of part_len bytes */
    size_t part_len, row1_off;
    uint64_t q1;
    int q1_bits;
    const uint64_t *const *that;  /* cols / n arrays of tg * n values */
    const uint64_t *inv_table;    /* n values */
} orc_engine;

size_t orc_request_len(size_t rows, size_t kp);
int orc_validate(const uint8_t *frame, size_t len, size_t rows, size_t kp);

uint32_t orc_unpack_u(const int8_t *ub, size_t kp, size_t cols, size_t kk, size_t col);

void orc_scan(uint64_t *inter, const uint8_t *db_col, size_t ld, size_t rows, size_t cols,
              const uint32_t *const *q, size_t B);
void orc_scan_acc(uint32_t *acc, const uint8_t *db_col, size_t ld, size_t rows, size_t cols,
                  const uint32_t *const *q, size_t B);
void orc_pack(uint64_t *beta, const int8_t *ub, const uint32_t *const *y, size_t B, size_t kp,
              size_t cols);

uint64_t orc_mulmod(uint64_t a, uint64_t b);
int orc_intt(uint64_t *a, const uint64_t *table, size_t n);
void orc_tail_pointwise(uint64_t *a, const uint64_t *beta, const uint32_t *z, const uint64_t *that,
                        size_t tg, size_t n);
uint64_t orc_rescale_q1(uint64_t v, uint64_t q1);
uint64_t orc_rescale_i64(uint64_t a, uint64_t inp_mod, uint64_t out_mod);
void orc_write_bits(uint8_t *data, size_t len, uint64_t val, size_t bit_offs, size_t num_bits);
uint64_t orc_count_rescale_mismatch(uint64_t lo, uint64_t hi, uint64_t q1);
int orc_switch_row1(uint8_t *res, size_t res_len, size_t bit_off, const uint64_t *inter,
                    const uint64_t *coef, size_t n, uint64_t q1, int q1_bits);

int orc_answer(const orc_engine *e, const uint8_t *const *req, const size_t *len, size_t B,
               uint8_t *const *out, int *status);

void orc_pack_build(int8_t *ub, const uint64_t *const *bold_t, const uint64_t *const *bold_t_bar,
                    const uint32_t *tables, const uint32_t *gen_pows, size_t num_iter, size_t tg,
                    size_t n, size_t rho);
void orc_tile_db(uint8_t *dst, const uint8_t *db, size_t ld, size_t rows, size_t cols);
void orc_pack_online(uint64_t *beta, const uint32_t *y, const uint64_t *const *bold_t,
                     const uint64_t *const *bold_t_bar, const uint32_t *tables, const uint32_t *gen_pows,
                     size_t num_iter, size_t tg, size_t n, size_t rho);

void orc_matmul_vec_packed(uint32_t *out, const uint32_t *a, const uint32_t *b, size_t rows, size_t cols,
                           size_t bc);
void orc_hint0(uint64_t *hint, size_t stride, const uint64_t *a, size_t npolys, size_t n, const uint8_t *db_col,
               size_t ld, size_t col0, size_t ncols);

/* Offline packing precomputation (InspiRING) of one output part, and its NTT helpers. */
void orc_ntt(uint64_t *a, size_t n, uint64_t psi);
size_t orc_auto_slot(size_t t, size_t j, size_t n);
int orc_inspiring_offline(uint64_t *bold_t, uint64_t *bold_t_bar, uint64_t *bold_t_hat, uint64_t *a_hat,
                          const uint64_t *hint, size_t stride, size_t col0, const uint64_t *w,
                          const uint64_t *v, size_t n, size_t tg, uint64_t psi);

void orc_fast_batched_dot_product(uint64_t *c, size_t c_stride, const uint64_t *a, size_t a_stride,
                                  size_t a_elems, const void *b_t, int b_bytes, size_t b_rows, size_t b_cols,
                                  size_t K, uint64_t m0, uint64_t m1);

/* Forward negacyclic NTT under the same slot convention, used only to pin orc_intt. */
void orc_ntt_naive(uint64_t *out, const uint64_t *in, size_t n, uint64_t psi);
void orc_intt_naive(uint64_t *out, const uint64_t *in, size_t n, uint64_t psi);
uint64_t orc_powmod(uint64_t b, uint64_t e);

#ifdef __cplusplus
}
#endif
#endif
